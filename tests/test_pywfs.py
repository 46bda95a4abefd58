"""Pins of the pyramid WFS tables (maos_amd/host/pywfs.c) and of the oracle's
pyramid imaging (orc_pywfs_ints, orc_pywfs_grad; lib/pywfs.c:370-522).

Expected values come from what the optics fixes:
- a facet of the pyramid is a phase ramp of a quarter wave per sample, so
  a single ramp over the whole focal plane moves the pupil image by a
  quarter of the grid, and without any mask the pupil image is the pupil
  intensity |amp|^2 (normalised to the total) at the embedded points;
- the detector transfer function has the zeros of its pixel sinc where
  j dsa / (notf dx) is an integer, and the Gaussian blur's transform;
- a flat wavefront on a symmetric pupil lights the four pupil images
  equally (mirror images of each other), so the flat-wavefront gradients
  vanish;
- a modulated pyramid's signal is linear in a small tilt, and its inverse
  optical gain is close to (pi/2) times the modulation radius (the
  geometric-optics limit).
"""
import math

import numpy as np
import pytest

import oracle
from maos_amd import _lib
from maos_amd._lib import dptr
from maos_amd.pywfs import Pywfs, PywfsCfg


def small(**kw):
    c = dict(D=2.0, din=0.3, order=8, dx=1.0 / 16, saat=0.3, siglev=1000.0)
    c.update(kw)
    return Pywfs(PywfsCfg(**c))


def oracle_make(pw):
    return lambda opd: oracle.pywfs_image(pw, opd)


def test_pyramid_mask_facets():
    N = 64
    m = np.zeros(2 * N * N)
    assert _lib.setup().mxs_pywfs_pyramid(dptr(m), N, 4, N, 0.0, 0.0) == 0
    M = (m[0::2] + 1j * m[1::2]).reshape(N, N)   # [y, x]
    y, x = np.mgrid[0:N, 0:N] - N // 2
    inside = x ** 2 + y ** 2 < (N / 2) ** 2
    assert np.all(M[~inside] == 0) and np.allclose(np.abs(M[inside]), 1, atol=1e-15)
    # each facet: a quarter wave per sample away from the apex
    want = np.exp(0.5j * np.pi * (np.abs(x) + np.abs(y)))
    assert np.allclose(M[inside], want[inside], atol=1e-12)
    # a field stop of 20 samples and flat edges of half-width 2
    assert _lib.setup().mxs_pywfs_pyramid(dptr(m), N, 4, 20, 2.0, 0.0) == 0
    M = (m[0::2] + 1j * m[1::2]).reshape(N, N)
    assert np.all(M[x ** 2 + y ** 2 >= 100] == 0)
    flat = (np.abs(x) < 2) | (np.abs(y) < 2)
    assert np.allclose(M[flat & (x ** 2 + y ** 2 < 100)], 1)
    # a roof: |x| only
    assert _lib.setup().mxs_pywfs_pyramid(dptr(m), N, 2, N, 0.0, 0.0) == 0
    M = (m[0::2] + 1j * m[1::2]).reshape(N, N)
    assert np.allclose(M[inside], np.exp(0.5j * np.pi * np.abs(x))[inside], atol=1e-12)


def test_nominal_zeros_and_blur():
    N, dx, dsa = 100, 1.0 / 16, 0.25       # sinc zeros every N dx / dsa = 25 samples
    S = _lib.setup()
    a = np.zeros(N * N)
    S.mxs_pywfs_nominal(dptr(a), N, dsa, dx, 0.0)
    A = a.reshape(N, N)
    assert math.isclose(A[0, 0], (dsa / dx) ** 2, rel_tol=1e-14)   # DC: pixel area in samples
    for j in (25, 50, 75):
        assert abs(A[0, j]) < 1e-12 * A[0, 0] and abs(A[j, 0]) < 1e-12 * A[0, 0]
    assert np.allclose(A, A.T) and np.allclose(A[1:, 1:], A[1:, 1:][::-1, ::-1])
    b = np.zeros(N * N)
    S.mxs_pywfs_nominal(dptr(b), N, dsa, dx, 0.3)
    B = b.reshape(N, N)
    j = np.fft.fftfreq(N) * N
    k2 = (j[:, None] ** 2 + j[None, :] ** 2) / (N * dx) ** 2
    gauss = np.exp(-2 * (math.pi * 0.3 * dsa) ** 2 * k2)      # transform of a Gaussian of rms 0.3 dsa
    nz = np.abs(A) > 1e-9 * A[0, 0]
    assert np.allclose(B[nz] / A[nz], gauss[nz], rtol=1e-12)


def test_saloc_and_image_offsets():
    S = _lib.setup()
    x, y = np.zeros(16), np.zeros(16)
    assert S.mxs_pywfs_saloc(dptr(x), dptr(y), 16, 4, 0.5) == 16
    assert np.allclose(np.unique(x), [-0.75, -0.25, 0.25, 0.75]) and np.allclose(x.mean(), 0)
    off = np.zeros(8)
    S.mxs_pywfs_sioff(dptr(off), 4)
    assert np.allclose(off.reshape(4, 2), [[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]])
    S.mxs_pywfs_sioff(dptr(off), 3)
    o3 = off[:6].reshape(3, 2)
    assert np.allclose(np.hypot(o3[:, 0], o3[:, 1]), 0.5)     # three images on a circle of radius 1/2


def _no_pyramid(pw, phase=None):
    """Replace the mask by 1 (or a given phase) everywhere, no modulation,
    an identity detector transfer function, every side sampled at the
    centre, and subapertures at the embedded grid points of chosen pupil
    points."""
    N = pw.N
    M = np.ones((N, N), complex) if phase is None else np.exp(1j * phase)
    pw.mask = np.ascontiguousarray(np.stack([M.real.ravel(), M.imag.ravel()], 1).ravel()[None, :])
    pw.nominal = np.ones(N * N)
    pw.modulate = 0.0
    pw.sioff = np.zeros_like(pw.sioff)
    rng = np.random.default_rng(4)
    pick = rng.choice(pw.nloc, 30, replace=False)
    ix, iy = pw.embed[pick] % N, pw.embed[pick] // N
    pw.salocx = (ix - N // 2 + 0.5) * pw.cfg.dx
    pw.salocy = (iy - N // 2 + 0.5) * pw.cfg.dx
    pw.saa = np.ones(30)
    return pick


def test_without_a_pyramid_the_image_is_the_pupil_intensity():
    pw = small(modulpos=4)
    pick = _no_pyramid(pw)
    opd = np.random.default_rng(1).normal(size=pw.nloc) * 2e-7
    ints, _ = oracle.pywfs_image(pw, opd, siglev=500.0)
    want = pw.nsa * 500.0 * pw.amp[pick] ** 2 / (pw.amp ** 2).sum()
    for side in range(4):
        assert np.allclose(ints[side], want, rtol=1e-10, atol=1e-12 * want.max())


@pytest.mark.parametrize("mx,my", [(3, 0), (0, -5), (2, 7)])
def test_a_ramp_mask_moves_the_pupil_image(mx, my):
    """mask exp(2 pi i (x mx + y my) / N): the image moves by -(mx, my)."""
    pw = small(modulpos=4)
    N = pw.N
    y, x = np.mgrid[0:N, 0:N]
    pick = _no_pyramid(pw, 2 * np.pi * (x * mx + y * my) / N)
    pw.salocx = pw.salocx - mx * pw.cfg.dx
    pw.salocy = pw.salocy - my * pw.cfg.dx
    opd = np.random.default_rng(2).normal(size=pw.nloc) * 2e-7
    ints, _ = oracle.pywfs_image(pw, opd, siglev=500.0)
    want = pw.nsa * 500.0 * pw.amp[pick] ** 2 / (pw.amp ** 2).sum()
    assert np.allclose(ints[0], want, rtol=1e-10, atol=1e-12 * want.max())


def test_flat_wavefront_lights_the_four_images_alike():
    """The pupil's centre point lands on grid sample N/2 and the pyramid's
    apex is the PSF's centre; with the detector grid centred on that sample
    too (MAOS's grid sits half a sample lower, lib/pywfs.c:279) the four
    images are mirror images.  Unmodulated the symmetry is exact; with
    modulation the PSF's Nyquist row and column (grid index 0 after
    centring), which has no mirror partner on an even grid, enters the
    shifted window on one side only, a few 1e-4 of the image."""
    for mod, tol in ((0.0, 1e-10), (5.0, 1e-3)):
        pw = small()
        pw.modulate = mod * pw.wvl.max() / pw.cfg.D
        pw.saa = np.ones(pw.nsa)
        pw.salocx = pw.salocx + 0.5 * pw.cfg.dx
        pw.salocy = pw.salocy + 0.5 * pw.cfg.dx
        ints, grad = oracle.pywfs_image(pw, np.zeros(pw.nloc))
        # image 0 (-x, -y) mirrored in x is image 1, mirrored in y image 2
        n = pw.cfg.order
        I = ints.reshape(4, n, n)                 # [side, row (y), column (x)]
        for k, m in ((1, I[0][:, ::-1]), (2, I[0][::-1, :]), (3, I[0][::-1, ::-1])):
            assert np.abs(I[k] - m).max() < tol * I.max()
        assert np.abs(grad).max() < 10 * tol  # gain 1: 1 is a full-scale quad-cell difference
        assert np.allclose(I.sum(axis=(1, 2)) / I.sum(), 0.25, atol=tol)


def test_calibrated_signal_is_linear_and_gain_follows_modulation():
    """With the modulation circle sampled densely (points less than 0.6
    lambda/D apart: 3 lambda/D with 32 points, 6 lambda/D with 64) the
    inverse optical gain is the geometric (pi/2) r; sparser circles respond
    through the few points that straddle an edge."""
    gains = []
    for mod, npos in ((3.0, 32), (6.0, 64)):
        pw = small(modulate=mod, modulpos=npos).calibrate(oracle_make)
        gains.append(pw.gain)
        alpha = 0.005 * math.pi / 180 / 3600
        _, g1 = oracle.pywfs_image(pw, alpha * pw.locx)
        _, g2 = oracle.pywfs_image(pw, 0.5 * alpha * pw.locx)
        nsa = pw.nsa
        # a tip moves the x gradients (by about the tip) and leaves y alone
        assert 0.7 < g1[:nsa].mean() / alpha < 1.3
        assert abs(g1[nsa:].mean()) < 0.05 * alpha
        assert np.allclose(g2, 0.5 * g1, rtol=2e-3, atol=2e-3 * np.abs(g1).max())
        r = mod * pw.wvl.max() / pw.cfg.D
        assert 0.95 < pw.gain / (0.5 * math.pi * r) < 1.05
    assert 1.9 < gains[1] / gains[0] < 2.1
