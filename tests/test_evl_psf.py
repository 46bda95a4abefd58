"""Science PSF (locfft_psf, lib/locfft.c:90-188; psfcomp_r,
cuda/sim/perfevl.cu:255-290) and the aperture embedding (loc_create_embed,
math/loc.c:147-179): the oracle and the host setup code, pinned by

- numpy's FFT (a library routine independent of the oracle's) on the same
  embedded field, over a ragged annular pupil with a duplicated index;
- closed forms: a flat wavefront's PSF peaks at the window centre
  (psfsize/2, psfsize/2) with value (sum amp)^2; a tilt of m whole cycles
  across the FFT grid moves that peak by exactly m pixels, in +x for a phase
  growing along +x (the forward transform's sign); Parseval's identity for
  a full-size window;
- the Strehl-only form equals the window's centre pixel;
- the embedding's size rule (nextfftsize of twice the extent), its centring
  (the odd pad on the left) and its errors;
- the wavefront-error decomposition (loc_calc_ptt, math/loc.c:653-709):
  against numpy's weighted least-squares residuals (the piston-removed and
  the piston/tip/tilt-removed variances as residual norms, the fit by
  lstsq), and exact on planes and pistons.
"""
import numpy as np
import pytest


def annulus(D, din, dx):
    """Points of a square grid of step dx inside an annulus (x fastest)."""
    n = int(round(D / dx)) + 1
    c = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(c, c)
    r = np.hypot(X, Y)
    m = (r <= D / 2 + 1e-9) & (r >= din / 2)
    return X[m].copy(), Y[m].copy()


def _embed(locx, locy, dx, nembed=0, oversize=2.0, fftpad=True):
    from maos_amd import engine
    return engine.loc_embed(locx, locy, dx, nembed, oversize, fftpad)


def _numpy_psf(opd, amp, embed, N, P, wvl):
    f = np.zeros(N * N, dtype=complex)
    for i in range(len(opd)):  # point order: a later point at one index wins
        f[embed[i]] = amp[i] * np.exp(2j * np.pi / wvl * opd[i])
    F = np.fft.fft2(f.reshape(N, N))  # rows are y, columns x
    h = P // 2
    idx = (np.arange(P) - h) % N
    return np.abs(F[np.ix_(idx, idx)]) ** 2


def test_embed_size_centring_and_errors():
    from maos_amd import _lib
    locx, locy = annulus(2.0, 0.3, 1.0 / 16)  # 33 points across
    emb, N = _embed(locx, locy, 1.0 / 16)
    assert N == _lib.setup().mxs_nextfftsize(66) == 70
    ix, iy = emb % N, emb // N
    # 33 points in 70: 37 of pad, 19 on the left and 18 on the right
    assert ix.min() == (N - 33 + 1) // 2 and iy.min() == (N - 33 + 1) // 2
    assert ix.max() - ix.min() == 32 and len(np.unique(emb)) == len(emb)
    # an explicit size is kept, and the grid step is respected
    emb2, N2 = _embed(locx, locy, 1.0 / 16, nembed=64)
    assert N2 == 64
    np.testing.assert_array_equal(np.round((locx - locx.min()) * 16), emb2 % 64 - (emb2 % 64).min())
    with pytest.raises(ValueError):
        _embed(locx, locy, 1.0 / 16, nembed=30)  # below 0.6 of the minimum (66)


def test_oracle_psf_matches_numpy_fft():
    import oracle
    locx, locy = annulus(2.0, 0.4, 1.0 / 16)
    emb, N = _embed(locx, locy, 1.0 / 16, nembed=64)
    rng = np.random.default_rng(5)
    amp = rng.uniform(0.5, 1.0, len(locx))
    opd = rng.normal(size=len(locx)) * 2e-7
    emb = emb.copy()
    emb[7] = emb[3]  # two points at one index: the later one (7) is used
    for P in (64, 32, 6):
        got = oracle.locfft_psf(opd, amp, emb, N, P, 1.0e-6)
        want = _numpy_psf(opd, amp, emb, N, P, 1.0e-6)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * want.max())


def test_oracle_psf_closed_forms():
    import oracle
    locx, locy = annulus(2.0, 0.0, 1.0 / 16)
    emb, N = _embed(locx, locy, 1.0 / 16, nembed=64)
    amp = np.full(len(locx), 1.0 / len(locx))
    wvl = 0.8e-6
    P = 32
    flat = oracle.locfft_psf(np.zeros(len(locx)), amp, emb, N, P, wvl)
    assert np.unravel_index(np.argmax(flat), flat.shape) == (P // 2, P // 2)
    assert flat[P // 2, P // 2] == pytest.approx(amp.sum() ** 2, rel=1e-12)
    # phase 2 pi m ix / N: the peak moves to x = P/2 + m
    for m in (3, -5):
        ix = emb % N
        opd = wvl * m * ix / N
        t = oracle.locfft_psf(opd, amp, emb, N, P, wvl)
        assert np.unravel_index(np.argmax(t), t.shape) == (P // 2, P // 2 + m)
        assert t.max() == pytest.approx(amp.sum() ** 2, rel=1e-10)
    # along y the same
    iy = emb // N
    t = oracle.locfft_psf(wvl * 2 * iy / N, amp, emb, N, P, wvl)
    assert np.unravel_index(np.argmax(t), t.shape) == (P // 2 + 2, P // 2)
    # Parseval: sum |F|^2 = N^2 sum |f|^2
    rng = np.random.default_rng(1)
    opd = rng.normal(size=len(locx)) * 3e-7
    full = oracle.locfft_psf(opd, amp, emb, N, N, wvl)
    assert full.sum() == pytest.approx(N * N * np.sum(amp ** 2), rel=1e-11)


def test_oracle_strehl_is_centre_pixel():
    import oracle
    locx, locy = annulus(2.0, 0.4, 1.0 / 16)
    emb, N = _embed(locx, locy, 1.0 / 16, nembed=64)
    rng = np.random.default_rng(2)
    amp = rng.uniform(0.2, 1.0, len(locx))
    opd = rng.normal(size=len(locx)) * 1e-7
    s = oracle.locfft_psf(opd, amp, emb, N, 1, 1.2e-6)[0, 0]
    c = oracle.locfft_psf(opd, amp, emb, N, 8, 1.2e-6)[4, 4]
    assert s == pytest.approx(c, rel=1e-11)
    assert s == pytest.approx(abs(np.sum(amp * np.exp(2j * np.pi / 1.2e-6 * opd))) ** 2, rel=1e-12)


def test_oracle_wfe_against_least_squares():
    import oracle
    locx, locy = annulus(4.0, 0.8, 1.0 / 16)
    rng = np.random.default_rng(6)
    amp = rng.uniform(0.3, 1.0, len(locx))
    opd = rng.normal(size=len(locx)) * 1e-7 + 3e-7 * locx - 2e-7 * locy + 5e-7
    rms, coeff = oracle.loc_calc_ptt(locx, locy, amp, opd)
    w = np.sqrt(amp)
    A = np.stack([np.ones_like(locx), locx, locy], 1)
    fit = np.linalg.lstsq(A * w[:, None], opd * w, rcond=None)[0]
    np.testing.assert_allclose(coeff, fit, rtol=1e-10)
    mean = np.sum(amp * opd) / amp.sum()
    pr = np.sum(amp * (opd - mean) ** 2)
    pttr = np.sum(amp * (opd - A @ fit) ** 2)
    np.testing.assert_allclose(rms, [pr, pr - pttr, pttr], rtol=1e-9)
    # a plane has no residual; a piston has neither residual nor tip/tilt
    rms, coeff = oracle.loc_calc_ptt(locx, locy, amp, 2e-7 + 4e-7 * locx + 1e-7 * locy)
    np.testing.assert_allclose(coeff, [2e-7, 4e-7, 1e-7], rtol=1e-10)
    assert abs(rms[2]) < 1e-12 * rms[1]
    piston = np.full(len(locx), 3e-7)
    rms, _ = oracle.loc_calc_ptt(locx, locy, amp, piston)
    tot = np.sum(amp * piston ** 2)
    assert np.all(np.abs(rms) < 1e-12 * tot)  # rounding of the differences only


def test_evl_init_rejects_unsupported_sizes():
    """mx_evl_init checks its sizes before any GPU call: a side that is not a
    power of two from 32 to 4096, an odd or oversized window, and direction
    or wavelength counts out of range are errors with a message."""
    import ctypes as C
    from maos_amd import _lib
    from maos_amd._lib import dptr, lptr
    lib = _lib.hip()
    amp = np.ones(10)
    wvl = np.array([1e-6])
    emb = np.arange(10, dtype=np.int64)

    def init(nembed, psfsize, ndir=1, nwvl=1):
        h = C.c_void_p()
        ne = np.array([nembed] * max(nwvl, 1), dtype=np.int64)
        ps = np.array([psfsize] * max(nwvl, 1), dtype=np.int64)
        w = np.full(max(nwvl, 1), 1e-6)
        return lib.mx_evl_init(C.byref(h), 10, None, None, dptr(amp), ndir, None, None, None, nwvl, dptr(w),
                               lptr(ne), lptr(ps), lptr(np.tile(emb, max(nwvl, 1))))

    assert init(96, 32) == -1 and b"powers of two" in lib.mx_last_error()
    assert init(8192, 32) == -1
    assert init(64, 7) == -1 and b"psfsize" in lib.mx_last_error()
    assert init(64, 128) == -1
    assert init(64, 32, ndir=0) == -1 and init(64, 32, ndir=65) == -1
    assert init(64, 32, nwvl=5) == -1
