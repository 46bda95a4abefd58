"""Pins of the sodium elongation transfer function (mxs_mketf, mxs_smooth)
against what the projection geometry fixes.

The ETF of a subaperture at distance rsa from the launch axis, at polar
angle srot, is the Fourier transform along the radial direction
(cos srot, sin srot) of the sodium column seen on the sky: a sample at
range r sits at the radial angle theta = rsa/r - rsa/hs from the focus
(lib/mkdtf.c:255-278 and, inverted, ih = hs rsa / (rsa + theta hs) at
:323-324).  The interpolating form (the default, llt.na_interp=1) weighs
the profile per unit angle and renormalises (:281-378); the direct form
sums the phasor of each sample.  Expected values come from a quadrature of
that definition with the exact Gaussian profile, and from the fact that a
single sample has a pure phase-ramp transform.
"""
import math

import numpy as np
import pytest

from maos_amd import _lib
from maos_amd._lib import dptr

N = 64
DTH = 0.589e-6          # lambda / (nwvf dx) of the lgs30m powfs (rad)
HS = 90e3


def S():
    return _lib.setup()


def mketf(h, na, srot, srsa, hs=HS, htel=0.0, za=0.0, no_interp=0, nominal=None):
    srot = np.ascontiguousarray(srot, dtype=np.float64)
    srsa = np.ascontiguousarray(srsa, dtype=np.float64)
    prof = np.ascontiguousarray(np.concatenate([h, na]).astype(np.float64))
    nsa = srot.size
    etf = np.zeros(2 * nsa * N * N)
    nom = None if nominal is None else np.ascontiguousarray(nominal, dtype=np.float64)
    rc = S().mxs_mketf(dptr(etf), None, dptr(nom) if nom is not None else None, N, DTH, dptr(prof), h.size,
                       dptr(srot), dptr(srsa), nsa, hs, htel, za, no_interp)
    assert rc == 0
    return (etf[0::2] + 1j * etf[1::2]).reshape(nsa, N, N)   # [isa, ky, kx]


def kgrid():
    k = np.fft.fftfreq(N, d=DTH)     # signed cycles/rad in FFT (corner) order
    ky, kx = np.meshgrid(k, k, indexing="ij")
    return kx, ky


def gauss(h, h0, sig):
    return np.exp(-0.5 * ((h - h0) / sig) ** 2)


def quadrature_etf(srot, rsa, h0, sig, hs=HS):
    """i0scale * int P(h(theta)) exp(-2 pi i k_r theta) dtheta /
    int P(h(theta)) dtheta, h(theta) = hs rsa / (rsa + theta hs), on a fine
    angle grid (the interpolating form's definition)."""
    th = np.linspace(-N / 2 * DTH, N / 2 * DTH, 16001)
    P = gauss(hs * rsa / (rsa + th * hs), h0, sig)
    kx, ky = kgrid()
    kr = (math.cos(srot) * kx + math.sin(srot) * ky).ravel()
    out = np.empty(kr.size, complex)
    for i in range(0, kr.size, 256):
        out[i:i + 256] = np.exp(-2j * np.pi * np.outer(kr[i:i + 256], th)) @ P
    return (out / P.sum()).reshape(N, N)


# sodium layer sampled every 10 m, 2 km rms, centred 2 km below the focus:
# the image is offset by rsa (1/88 - 1/90 km) and elongated along srot
H = np.arange(78e3, 98e3 + 1, 10.0)
H0, SIG = 88e3, 2e3
NA = gauss(H, H0, SIG) / gauss(H, H0, SIG).sum()


@pytest.mark.parametrize("srot", [0.0, math.pi / 2, -math.pi])
def test_interp_etf_on_axis_equals_the_quadrature(srot):
    """srot along an axis: the 2-D table samples the 1-D transform at its
    own grid points (no rotation interpolation), so it equals the
    quadrature closely; the phase is -2 pi k_r theta0 and the magnitude
    falls along srot only."""
    rsa = 10.0
    e = mketf(H, NA, [srot], [rsa])[0]
    q = quadrature_etf(srot, rsa, H0, SIG)
    assert np.abs(e - q).max() < 2e-4
    kx, ky = kgrid()
    kr = math.cos(srot) * kx + math.sin(srot) * ky
    theta0 = rsa / H0 - rsa / HS
    low = np.abs(kr) * 2 * math.pi * SIG * rsa / HS ** 2 < 0.5   # where the spot width does not matter yet
    ph = np.angle(e[low] * np.exp(2j * np.pi * kr[low] * theta0))
    assert np.abs(ph).max() < 0.05
    across = np.abs(kr) < 1e-3 / DTH
    assert np.abs(np.abs(e[across]) - 1).max() < 1e-9   # no elongation across srot


@pytest.mark.parametrize("srot,rsa", [(math.pi / 6, 10.0), (-2.3, 6.0), (1.1, 14.0)])
def test_interp_etf_rotated_equals_the_quadrature(srot, rsa):
    """Rotated: the table interpolates linearly between 1-D samples half a
    frequency step apart (h = 1 / (2 notf dtheta)), so it is within the
    linear-interpolation bound h^2/8 max|f''| of the quadrature, with
    |f''| <= (2 pi)^2 (theta0^2 + sigma_theta^2) for the offset Gaussian
    spot; a swapped axis or a reversed angle is O(1) off."""
    e = mketf(H, NA, [srot], [rsa])[0]
    q = quadrature_etf(srot, rsa, H0, SIG)
    h = 1.0 / (2 * N * DTH)
    theta0, sig_theta = rsa / H0 - rsa / HS, rsa * SIG / H0 ** 2
    bound = h ** 2 / 8 * (2 * math.pi) ** 2 * (theta0 ** 2 + sig_theta ** 2)
    assert np.abs(e - q).max() < 1.5 * bound
    for wrong in (quadrature_etf(-srot, rsa, H0, SIG), quadrature_etf(math.pi / 2 - srot, rsa, H0, SIG)):
        assert np.abs(e - wrong).max() > 0.1


def test_interp_etf_zero_distance_and_nominal_fusion():
    nom = np.random.default_rng(1).normal(size=N * N)
    e = mketf(H, NA, [0.3, 0.3], [0.0, 10.0], nominal=nom)
    assert np.abs(e[0] - nom.reshape(N, N)).max() == 0     # rsa = 0: ETF = 1, times the nominal
    e1 = mketf(H, NA, [0.3], [10.0])[0]
    assert np.abs(e[1] - e1 * nom.reshape(N, N)).max() <= 1e-15 * np.abs(e1).max() * np.abs(nom).max()


def test_interp_etf_rejects_uneven_heights():
    h = np.concatenate([np.arange(80e3, 90e3, 100.0), np.arange(90e3, 100e3, 200.0)])
    na = gauss(h, 90e3, 2e3)
    prof = np.ascontiguousarray(np.concatenate([h, na]))
    etf = np.zeros(2 * N * N)
    rc = S().mxs_mketf(dptr(etf), None, None, N, DTH, dptr(prof), h.size, dptr(np.zeros(1)),
                       dptr(np.full(1, 5.0)), 1, HS, 0.0, 0.0, 0)
    assert rc == -2
    assert mketf(h, na / na.sum(), [0.0], [5.0], no_interp=1).shape == (1, N, N)   # the direct form takes any spacing


@pytest.mark.parametrize("za,htel", [(0.0, 0.0), (0.4, 2600.0)])
def test_direct_etf_of_one_sample_is_a_phase_ramp(za, htel):
    """A one-sample profile (MAOS takes the direct form for it): ETF =
    w exp(-2 pi i k_r theta) with theta from the sample's range
    (h - htel)/cos(za); at the focus range it is w everywhere."""
    srot = np.array([0.0, 0.9, -2.0])
    srsa = np.array([4.0, 9.0, 13.0])
    kx, ky = kgrid()
    at_focus = np.array([HS * math.cos(za) + htel])
    e = mketf(at_focus, np.array([0.7]), srot, srsa, htel=htel, za=za)
    assert np.abs(e - 0.7).max() < 1e-12
    h = np.array([86.5e3])
    e = mketf(h, np.array([1.0]), srot, srsa, htel=htel, za=za)
    for i in range(3):
        theta = srsa[i] / ((h[0] - htel) / math.cos(za)) - srsa[i] / HS
        kr = math.cos(srot[i]) * kx + math.sin(srot[i]) * ky
        assert np.abs(e[i] - np.exp(-2j * np.pi * kr * theta)).max() < 1e-12


def test_direct_etf_is_linear_in_the_profile():
    srot, srsa = np.array([0.4, 2.0]), np.array([5.0, 12.0])
    h = np.array([84e3, 91e3, 99e3])
    w = np.array([0.2, 0.5, 0.3])
    e = mketf(h, w, srot, srsa, no_interp=1)
    parts = sum(mketf(h[i:i + 1], w[i:i + 1], srot, srsa) for i in range(3))
    assert np.abs(e - parts).max() < 1e-12
    assert np.abs(e[:, 0, 0] - 1.0).max() < 1e-12            # ETF(0) = sum of the strengths


def test_direct_form_with_angle_weights_equals_the_interpolating_form():
    """The two forms are independent codings of the same geometry.  The
    interpolating form weighs the profile per unit angle, the direct form
    per sample; giving each height sample the weight P(h) |dtheta/dh| =
    P(h) rsa / h^2 makes the direct sum the same integral.  Subapertures on
    the axes, so that the interpolating form adds no rotation error."""
    for srot, rsa in ((0.0, 7.0), (math.pi / 2, 11.0), (-math.pi, 14.5)):
        ei = mketf(H, NA, [srot], [rsa], no_interp=0)[0]
        w = NA * rsa / H ** 2
        ed = mketf(H, w / w.sum(), [srot], [rsa], no_interp=1)[0]
        assert np.abs(ei - ed).max() < 1e-3


def test_smooth_keeps_the_first_moment_and_splits_a_sample_linearly():
    """Linear binning (smooth, lib/mkdtf.c:421-452): each input sample goes
    to its two neighbours with weights 1-f and f, which keeps the sum and
    the first moment of the profile whenever both neighbours exist (the
    profile vanishes near the ends, where a neighbour is missing)."""
    h = np.arange(80e3, 100e3 + 1, 10.0)
    na = gauss(h, 90e3, 1.5e3)
    prof = np.ascontiguousarray(np.concatenate([h, na]))
    out = np.zeros(2 * h.size)
    m = S().mxs_smooth(dptr(prof), h.size, 250.0, dptr(out), h.size)
    ho, no = out[:m], out[m:2 * m]
    assert math.isclose((ho * no).sum() / no.sum(), (h * na).sum() / na.sum(), rel_tol=1e-9)
    # one bright sample at 80.8 km: between output samples 3 and 4 at fraction 0.2
    na1 = np.zeros_like(h)
    na1[80] = 1.0
    prof = np.ascontiguousarray(np.concatenate([h, na1]))
    m = S().mxs_smooth(dptr(prof), h.size, 250.0, dptr(out), h.size)
    got = out[m:2 * m]
    assert np.isclose(got[3], 0.8, atol=1e-12) and np.isclose(got[4], 0.2, atol=1e-12)
    assert np.abs(np.delete(got, [3, 4])).max() == 0
