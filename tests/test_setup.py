"""Host table builders (libmaos_setup.so): identities and known values."""
import math

import numpy as np
import pytest

from maos_amd import _lib
from maos_amd._lib import dptr, lptr
from maos_amd.powfs import AperCfg, Powfs, PowfsCfg, SimCfg


def S():
    return _lib.setup()


def test_nextfftsize_values():
    # nextfftsize (math/mathmisc.c:109-146): even, 2-3-5-7 smooth, >= n
    got = [S().mxs_nextfftsize(n) for n in (28, 32, 50, 64, 65, 121, 480, 481)]
    assert got == [28, 32, 50, 64, 70, 126, 480, 486]


def test_fft2_matches_numpy():
    a = np.random.RandomState(1).randn(8, 6) + 1j * np.random.RandomState(2).randn(8, 6)
    b = np.ascontiguousarray(a.copy())
    S().mxs_fft2(dptr(b.view(np.float64)), 6, 8, -1)
    assert np.allclose(b, np.fft.fft2(a))


def test_saloc_and_pruning_ngs():
    p = Powfs(AperCfg(8.0, 1.2), PowfsCfg(order=16, dsa=0.5, dx=1 / 32), SimCfg())
    assert p.nxsa == 16 and p.notf == 32
    assert p.saa.min() >= 0.3 and p.saa.max() <= 1.0 + 1e-12
    # annulus area / subaperture area bounds the number of subapertures
    area = math.pi / 4 * (8.0 ** 2 - 1.2 ** 2) / 0.25
    assert 0.9 * area < p.saa.sum() < 1.1 * area
    # symmetric pupil: subaperture centres average to the pupil centre
    assert abs((p.ptsx + 0.25 - p.dx / 2).mean()) < 1e-12
    assert abs((p.ptsy + 0.25 - p.dx / 2).mean()) < 1e-12


def test_loc_circle_add_area():
    n = 200
    x = (np.arange(n) - n / 2 + 0.5) * 0.05
    X, Y = np.meshgrid(x, x)
    lx, ly = np.ascontiguousarray(X.ravel()), np.ascontiguousarray(Y.ravel())
    amp = np.zeros(n * n)
    S().mxs_loc_circle_add(dptr(amp), dptr(lx), dptr(ly), n * n, 0.05, 0.05, 0.0, 0.0, 4.0, 0.6, 1.0)
    assert math.isclose(amp.sum() * 0.05 ** 2, math.pi * (16 - 0.36), rel_tol=2e-3)
    assert amp.min() >= -1e-12 and amp.max() <= 1 + 1e-12


def test_dtf_treat_is_checkerboard_shift():
    """dotf_treat = fftshift + (-1)^(ix+iy): the treated nominal equals the
    analytic pixel x blur transfer function, shifted and sign-alternated."""
    N = 32
    pd = 0.5 * 4.848e-6 / (0.7e-6 / 1.0)
    nom = np.zeros(N * N)
    S().mxs_dtf_nominal(dptr(nom), N, N, pd, pd, 0.0, 0.3)
    nom = nom.reshape(N, N)
    j = np.arange(N) - N // 2
    du = pd / N
    g = np.sinc(j * du) * pd * np.exp(-2 * (math.pi * 0.3 * du) ** 2 * j ** 2)
    cen = np.outer(g, g)
    ref = np.fft.fftshift(cen) * ((-1.0) ** np.add.outer(np.arange(N), np.arange(N)))
    assert np.allclose(nom, ref, atol=1e-13)


def test_si_bilinear_weights():
    N, npx = 32, 8
    dth = 0.7e-6 / 1.0
    pix = 0.5 * 4.848137e-6
    pp = np.zeros(N * N + 1, dtype=np.int64)
    pi = np.zeros(4 * npx * npx, dtype=np.int64)
    px = np.zeros(4 * npx * npx)
    nnz = S().mxs_si(N, N, dth, npx, npx, pix, pix, 0.0, 0.0, 0.0, lptr(pp), lptr(pi), dptr(px))
    assert 0 < nnz <= 4 * npx * npx
    H = np.zeros((npx * npx, N * N))
    for col in range(N * N):
        for k in range(pp[col], pp[col + 1]):
            H[pi[k], col] = px[k]
    # interior pixels: weights sum to one and reproduce a linear image exactly
    assert np.allclose(H.sum(1), 1.0)
    yy, xx = np.mgrid[0:N, 0:N]
    lin = (xx - N // 2) * dth + 2 * (yy - N // 2) * dth
    cx = (np.arange(npx) - (npx / 2 - 0.5)) * pix
    expect = np.add.outer(2 * cx, cx).ravel()
    assert np.allclose(H @ lin.ravel(), expect, atol=1e-12 * pix)


def test_smooth_preserves_sum():
    h = np.arange(80e3, 100e3 + 1, 10.0)
    na = np.exp(-0.5 * ((h - 90e3) / 3e3) ** 2)
    prof = np.ascontiguousarray(np.concatenate([h, na]))
    out = np.zeros(2 * h.size)
    m = S().mxs_smooth(dptr(prof), h.size, 250.0, dptr(out), h.size)
    assert m == math.ceil((h[-1] - h[0]) / 250.0)
    assert math.isclose(out[m:2 * m].sum(), na.sum(), rel_tol=1e-12)
    assert np.allclose(out[:m], h[0] + 250.0 * np.arange(m))


def test_etf_limits():
    """ETF with rsa=0 is 1 (only the nominal remains); a narrow profile
    gives |ETF| ~ 1 near the origin; DC of the ETF is sum(profile)."""
    N = 64
    dth = 0.589e-6 / 1.0
    nominal = np.ones(N * N)
    h = np.arange(85e3, 95e3 + 1, 100.0)
    na = np.exp(-0.5 * ((h - 90e3) / 3e3) ** 2)
    na /= na.sum()
    prof = np.ascontiguousarray(np.concatenate([h, na]))
    srot = np.array([0.0, 0.7])
    srsa = np.array([0.0, 10.0])
    etf = np.zeros(2 * 2 * N * N)
    rc = S().mxs_mketf(dptr(etf), None, dptr(nominal), N, dth, dptr(prof), h.size, dptr(srot),
                       dptr(srsa), 2, 90e3, 0.0, 0.0, 0)
    assert rc == 0
    e = (etf[0::2] + 1j * etf[1::2]).reshape(2, N, N)
    assert np.allclose(e[0], 1.0)
    assert math.isclose(e[1][0, 0].real, 1.0, rel_tol=1e-9) and abs(e[1][0, 0].imag) < 1e-12
    assert np.abs(e[1]).max() <= 1 + 1e-9


def test_llt_amp_normalised():
    amp = np.zeros(4096)
    o = _lib.C.c_double(0)
    n = S().mxs_llt_amp(0.4, 0.6, 1 / 64, 0.5, dptr(amp), 4096, _lib.C.byref(o))
    assert n == 32 and math.isclose(o.value, (1 / 64 - 0.5) / 2)
    a = amp[:n * n]
    assert (a ** 2).sum() < 1.0 and (a ** 2).sum() > 0.95   # clipped Gaussian
    assert math.isclose(a.reshape(n, n)[n // 2, n // 2] ** 2 / a.max() ** 2, 1.0, rel_tol=0.05)


def test_genscreen_structure_function():
    """The screens are the von Karman PSD (spatial_psd, lib/turbulence.c:
    436-480) times unit complex Gaussian noise, so their structure function
    must equal its expectation 2 sum PSD df^2 (1 - cos(2 pi f r)) over the
    discrete frequency grid (checked at short lags, where one realisation
    averages well)."""
    n, dx, r0, L0 = 1024, 1 / 32, 0.2, 30.0
    scr = np.zeros(n * n)
    wt = np.array([1.0])
    S().mxs_genscreen(dptr(scr), n, n, dx, r0, L0, dptr(wt), 1, 3)
    s = scr.reshape(n, n)
    f = np.fft.fftfreq(n, dx)
    fx, fy = np.meshgrid(f, f)
    strength = 0.0229 * r0 ** (-5 / 3) * (0.5e-6 / (2 * math.pi)) ** 2
    psd = strength * (fx ** 2 + fy ** 2 + L0 ** -2) ** (-11 / 6)
    psd[0, 0] = 0
    df2 = (1 / (n * dx)) ** 2
    for lag in (1, 4):
        d = np.mean((s[:, lag:] - s[:, :-lag]) ** 2)
        expect = 2 * np.sum(psd * df2 * (1 - np.cos(2 * np.pi * fx * lag * dx)))
        assert 0.9 < d / expect < 1.1
    # independent seeds give independent screens; the same seed repeats
    scr2 = np.zeros(n * n)
    S().mxs_genscreen(dptr(scr2), n, n, dx, r0, L0, dptr(wt), 1, 3)
    assert np.array_equal(scr, scr2)
    S().mxs_genscreen(dptr(scr2), n, n, dx, r0, L0, dptr(wt), 1, 4)
    d1, d2 = np.diff(scr.reshape(n, n), axis=1), np.diff(scr2.reshape(n, n), axis=1)
    assert abs(np.corrcoef(d1.ravel(), d2.ravel())[0, 1]) < 0.03


def test_mtch_is_weighted_pinv():
    """mtche = i0m * pinv(i0g, W): compare one subaperture with a direct
    weighted least-squares solve in numpy (no constraint columns)."""
    rng = np.random.RandomState(3)
    npx = 6
    yy, xx = np.mgrid[0:npx, 0:npx]
    i0 = 200 * np.exp(-((xx - 2.6) ** 2 + (yy - 2.4) ** 2) / 2.0) + 1.0
    i0 = np.ascontiguousarray(i0.reshape(1, -1))
    pix = 1e-6
    # derivatives have the scale of i0/pixtheta, as the FFT derivative does
    gx = np.ascontiguousarray(rng.randn(1, npx * npx) * 50 / pix)
    gy = np.ascontiguousarray(rng.randn(1, npx * npx) * 50 / pix)
    m = np.zeros(2 * npx * npx)
    san = np.zeros(3)
    isum = np.zeros(1)
    ss = np.zeros(1)
    rc = S().mxs_mtch(dptr(m), dptr(san), dptr(isum), dptr(ss), dptr(i0), dptr(gx), dptr(gy), 1,
                      npx, npx, 0.0, 0.0, 3.0, pix, pix, 0, 1.0)
    assert rc == 0
    A = np.stack([gx[0], gy[0], i0[0] / pix], axis=1)
    W = np.diag(1.0 / (9.0 + i0[0]))
    P = np.linalg.solve(A.T @ W @ A, A.T @ W)
    assert np.allclose(m.reshape(npx * npx, 2).T, P[:2], rtol=1e-9, atol=1e-12 * np.abs(P).max())
    nea = P[:2] @ np.linalg.inv(W) @ P[:2].T
    assert np.allclose([san[0], san[1], san[2]], [nea[0, 0], nea[1, 1], nea[0, 1]], rtol=1e-9)
    assert math.isclose(isum[0], i0.sum())
