"""Pins of the oracle's imaging chain on the LGS branch: the uplink OTF and
its conjugate, the ETF multiply, the two-profile blend and the normalisation.

orc_wfsints follows maos/wfsints.c:79-302 (uplink OTF :149-199, ccwmc
:277-279 with math/matcomp.c:30-36, apply_dtf_etf :38-64 and :200-223,
dspmulcreal :286).  Each expected value here is fixed by mathematics, not
by a restatement of that chain:

- a flat wavefront on a full square subaperture has the Dirichlet-kernel
  PSF |D(kx)|^2 |D(ky)|^2, D(k) = sum_{n<nxsa} exp(-2 pi i k n / nwvf);
- a tilt of m PSF samples is an exact circular shift by m samples;
- an uplink tilt of the same OPD as the downlink tilt cancels it exactly
  (the OTFs carry opposite phase ramps because the uplink OTF enters
  conjugated), so the uplink moves the image the other way;
- the image is linear in the ETF, so the blend of two ETFs is the blend of
  the two images, and an ETF that equals the nominal reproduces the nominal
  branch;
- an ETF phase ramp exp(-2 pi i (ux mx + uy my) / notf) moves the image by
  +(mx, my) samples.

The detector sampler is the identity (one pixel per image sample), so the
intensities are the full notf x notf image.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import OrcPowfs, OrcWfs, dp, lp


def _d(a):
    return a.ctypes.data_as(dp) if a is not None else None


class Mini:
    """A powfs of nsa full square subapertures (amp = 1) with an identity
    detector sampler: npix = notf^2 pixels, pixel p = image sample p."""

    def __init__(self, nsa=2, nxsa=8, embfac=2, notf=None, wvl=0.589e-6, sigsim=1000.0,
                 areascale=1.0, nlx=None):
        self.nsa, self.nxsa, self.embfac = nsa, nxsa, embfac
        self.nwvf = nxsa * embfac
        self.notf = notf or self.nwvf
        N = self.notf
        self.wvl = np.array([wvl])
        self.wvlwts = np.array([1.0])
        self.sigsim, self.areascale = sigsim, areascale
        self.dx = 1.0 / 64
        self.amp = np.ones(nsa * nxsa * nxsa)
        self.saa = np.ones(nsa)
        self.ptsx = np.zeros(nsa)
        self.ptsy = np.zeros(nsa)
        iy, ix = np.mgrid[0:N, 0:N]
        # treat(ones) = (-1)^(ix+iy) fftshift(ones) (lib/mkdtf.c:26-39)
        self.nominal = np.ascontiguousarray(((-1.0) ** (ix + iy)).reshape(-1))
        self.si_pp = np.arange(N * N + 1, dtype=np.int64)
        self.si_pi = np.arange(N * N, dtype=np.int64)
        self.si_px = np.ones(N * N)
        self.nlx = nlx if nlx is not None else nxsa
        self.llt_amp = np.ones(self.nlx * self.nlx)

    def ints(self, opd, lotf=None, etf=None, etf2=None, etf2wt=0.0, llt=False):
        s = OrcPowfs()
        s.nsa, s.nxsa, s.embfac, s.notf = self.nsa, self.nxsa, self.embfac, self.notf
        s.pixpsax = s.pixpsay = self.notf
        s.nwvl = 1
        s.dx, s.areascale, s.saat = self.dx, self.areascale, 0.3
        s.ptsx, s.ptsy, s.amp, s.saa, s.wvl = (_d(a) for a in (self.ptsx, self.ptsy, self.amp, self.saa,
                                                               self.wvl))
        s.nominal = _d(self.nominal)
        s.si_pp, s.si_pi = self.si_pp.ctypes.data_as(lp), self.si_pi.ctypes.data_as(lp)
        s.si_px = _d(self.si_px)
        s.si_nzmax = self.si_px.size
        keep = [np.ascontiguousarray(x) if x is not None else None for x in (etf, etf2, lotf)]
        etf, etf2, lotf = [k.view(np.float64).reshape(-1) if k is not None else None for k in keep]
        s.etf, s.etf2, s.etf2wt = _d(etf), _d(etf2), etf2wt
        s.llt_nx = self.nlx if llt else 0
        w = OrcWfs(0.0, 0.0, 90e3, self.sigsim, 0.0, 0.0, _d(self.wvlwts))
        out = np.zeros(self.nsa * self.notf * self.notf)
        o = np.ascontiguousarray(opd, dtype=np.float64).reshape(-1)
        oracle.lib().orc_wfsints(_d(out), _d(o), _d(lotf), C.byref(s), C.byref(w), 0, self.nsa)
        return out.reshape(self.nsa, self.notf, self.notf)

    def lotf(self, lltopd):
        N = self.notf
        out = np.zeros(2 * N * N)
        o = np.ascontiguousarray(lltopd, dtype=np.float64)
        oracle.lib().orc_lotf(_d(out), _d(o), _d(self.llt_amp), self.nlx, self.embfac, self.nwvf, N,
                              float(self.wvl[0]))
        return (out[0::2] + 1j * out[1::2]).reshape(N, N)

    def tilt_opd(self, n, mx=0.0, my=0.0):
        """OPD of an n x n grid (x fastest) that tilts the PSF by (mx, my)
        samples of lambda / (nwvf dx)."""
        a = np.asarray(self.wvl[0]) / (self.nwvf * self.dx)
        iy, ix = np.mgrid[0:n, 0:n].astype(float)
        return (a * self.dx * (mx * ix + my * iy)).reshape(-1)


def dirichlet_psf(nxsa, nwvf, notf):
    """|D(kx)|^2 |D(ky)|^2 on nwvf samples, padded to notf by the corner
    quadrants (ccpcorner, math/matcomp.c:176-216), as [ky, kx]."""
    k = np.arange(nwvf)
    D = np.exp(-2j * np.pi * np.outer(k, np.arange(nxsa)) / nwvf).sum(axis=1)
    p1 = np.abs(D) ** 2
    if notf != nwvf:
        h = nwvf // 2
        q = np.zeros(notf)
        q[:h] = p1[:h]
        q[notf - h:] = p1[h:]
        p1 = q
    return np.outer(p1, p1)


def centred(img):
    n = img.shape[0]
    return np.roll(np.roll(img, n // 2, 0), n // 2, 1)


@pytest.mark.parametrize("nxsa,notf", [(8, 16), (8, 32), (6, 12)])
def test_flat_wavefront_is_the_dirichlet_image(nxsa, notf):
    """ints = sigsim areascale |D|^2 |D|^2 / (nxsa nwvf)^2, centred by the
    treated nominal; notf > nwvf exercises ccpcorner's zero padding."""
    m = Mini(nsa=1, nxsa=nxsa, notf=notf, areascale=1.3)
    got = m.ints(np.zeros(nxsa * nxsa))[0]
    want = m.sigsim * m.areascale * centred(dirichlet_psf(nxsa, m.nwvf, notf)) / (nxsa * m.nwvf) ** 2
    assert np.abs(got - want).max() <= 1e-12 * want.max()
    assert np.isclose(got.sum(), m.sigsim * m.areascale, rtol=1e-12)  # Parseval: every sample kept


def test_flat_wavefront_through_the_bilinear_sampler(ngs_small):
    """With the powfs's own si (bilinear sampler at the pixel centres), the
    intensities of a full subaperture are si applied to the Dirichlet image."""
    cfg, p, layers = ngs_small
    from maos_amd.powfs import Layer
    from maos_amd import configs
    o = oracle.Oracle(p)
    o.atm([Layer(np.zeros_like(layers[0].p), layers[0].dx, 0.0, 0.0, 0.0)])
    w = configs.wfs_specs(cfg, p)[0]
    ints, _, _, _ = o.step(w, 0, cfg["sim"].dt)
    N, nxsa = p.notf, p.nxsa
    nwvf = nxsa * cfg["powfs"].embfac
    full = [i for i in range(p.nsa) if np.all(p.amp[i * nxsa * nxsa:(i + 1) * nxsa * nxsa] == 1.0)]
    assert full
    # the image before sampling: nominal = treat(DTF) multiplies the OTF
    P = dirichlet_psf(nxsa, nwvf, N)
    img = np.fft.ifft2(np.fft.fft2(P) * p.nominal[0].reshape(N, N)).real * N * N
    img = img.reshape(-1) * p.sigsim * p.areascale / (nxsa * nwvf) ** 2 / (N * N)
    want = np.zeros(p.npix)
    pp, pi, px = p.si_pp[0], p.si_pi[0], p.si_px[0]
    for col in range(N * N):
        for k in range(pp[col], pp[col + 1]):
            want[pi[k]] += px[k] * img[col]
    for isa in full:
        assert np.abs(ints[isa] - want).max() <= 1e-10 * want.max()


@pytest.mark.parametrize("mx,my", [(1, 0), (0, 2), (-3, 1)])
def test_downlink_tilt_shifts_the_image(mx, my):
    m = Mini(nsa=1)
    base = m.ints(np.zeros(64))[0]
    got = m.ints(m.tilt_opd(8, mx, my))[0]
    assert np.abs(got - np.roll(np.roll(base, my, 0), mx, 1)).max() <= 1e-11 * base.max()


@pytest.mark.parametrize("mx,my", [(1, 0), (0, 1), (2, -1)])
def test_uplink_tilt_moves_the_image_the_other_way(mx, my):
    """The uplink OTF enters conjugated (maos/wfsints.c:277-279): an uplink
    tilt of m samples moves the LGS image by -m, and uplink and downlink
    tilts of the same OPD cancel exactly."""
    m = Mini(nsa=1)
    flat = np.zeros(64)
    L0 = m.lotf(flat)
    Lt = m.lotf(m.tilt_opd(m.nlx, mx, my))
    base = m.ints(flat, lotf=L0)[0]
    up = m.ints(flat, lotf=Lt)[0]
    assert np.abs(up - np.roll(np.roll(base, -my, 0), -mx, 1)).max() <= 1e-11 * base.max()
    both = m.ints(m.tilt_opd(8, mx, my), lotf=Lt)[0]
    assert np.abs(both - base).max() <= 1e-11 * base.max()
    down = m.ints(m.tilt_opd(8, -mx, -my), lotf=L0)[0]
    assert np.abs(up - down).max() <= 1e-11 * base.max()


def test_uplink_otf_is_the_correlation_with_the_uplink_psf():
    """Flat downlink: the image is the circular cross-correlation of the
    downlink PSF with the uplink PSF (a sum over shifts, no FFT), the uplink
    PSF being |DFT(llt amp exp(2 pi i lltopd / wvl))|^2 / nlwvf^2.  The uplink
    OPD is random, so the uplink PSF is not centro-symmetric and the
    correlation differs from the convolution a missing conjugate would give."""
    m = Mini(nsa=1)
    m.llt_amp = np.linspace(0.2, 1.0, m.nlx * m.nlx) ** 2
    N = m.notf
    flat = np.zeros(64)
    lltopd = np.random.default_rng(11).normal(size=m.nlx * m.nlx) * 1e-7
    got = m.ints(flat, lotf=m.lotf(lltopd))[0]
    lw = np.zeros((N, N), complex)
    s = (N - m.nlx) // 2
    lw[s:s + m.nlx, s:s + m.nlx] = (m.llt_amp * np.exp(2j * np.pi * lltopd / m.wvl[0])).reshape(m.nlx, m.nlx)
    k = np.arange(N)
    F = np.exp(-2j * np.pi * np.outer(k, k) / N)   # explicit DFT matrix
    Q = np.abs(F @ lw @ F.T) ** 2 / (N * N)
    P = dirichlet_psf(m.nxsa, m.nwvf, N)
    corr = np.zeros((N, N))
    for y in range(N):
        for x in range(N):
            corr += Q[y, x] * np.roll(np.roll(P, -y, 0), -x, 1)  # sum_y P(x+y) Q(y)
    want = m.sigsim * centred(corr) / (m.nxsa * m.nwvf) ** 2
    assert np.abs(got - want).max() <= 1e-11 * want.max()
    conv = np.zeros((N, N))
    for y in range(N):
        for x in range(N):
            conv += Q[y, x] * np.roll(np.roll(P, y, 0), x, 1)    # without the conjugate
    assert np.abs(m.sigsim * centred(conv) / (m.nxsa * m.nwvf) ** 2 - want).max() > 1e-3 * want.max()


def test_uplink_grid_smaller_than_the_subaperture():
    """nlx < nxsa: the uplink OTF is cropped/padded by ccpcorner
    (maos/wfsints.c:170-178); uplink and downlink tilts still cancel."""
    m = Mini(nsa=1, nxsa=8, nlx=4)
    flat = np.zeros(64)
    base = m.ints(flat, lotf=m.lotf(np.zeros(16)))[0]
    Lt = m.lotf(m.tilt_opd(4, 1, 0))
    both = m.ints(m.tilt_opd(8, 1, 0), lotf=Lt)[0]
    assert np.abs(both - base).max() <= 1e-11 * base.max()
    up = m.ints(flat, lotf=Lt)[0]
    assert np.abs(up - np.roll(base, -1, 1)).max() <= 1e-11 * base.max()


def _etf_tables(m, rng):
    N = m.notf
    a = rng.normal(size=(m.nsa, N, N)) + 1j * rng.normal(size=(m.nsa, N, N))
    return a * m.nominal.reshape(N, N)


def test_etf_equal_to_nominal_reproduces_the_nominal_branch():
    """With an LLT the nominal is not applied separately (it is fused into
    the ETF, lib/mkdtf.c:389-393): an ETF table equal to the nominal gives
    the image of the no-ETF nominal branch."""
    m = Mini(nsa=2)
    rng = np.random.default_rng(3)
    opd = rng.normal(size=2 * 64) * 5e-8
    L = m.lotf(rng.normal(size=64) * 5e-8)
    nominal_branch = m.ints(opd, lotf=L)
    etf = np.tile(m.nominal.astype(complex), (2, 1))
    etf_branch = m.ints(opd, lotf=L, etf=etf, llt=True)
    assert np.abs(etf_branch - nominal_branch).max() <= 1e-12 * nominal_branch.max()


@pytest.mark.parametrize("w", [0.0, 0.3, 1.0])
def test_etf_blend_is_linear(w):
    """apply_dtf_etf with two ETFs (maos/wfsints.c:211-223): the image of
    (1-w) etf1 + w etf2 is (1-w) image(etf1) + w image(etf2)."""
    m = Mini(nsa=2)
    rng = np.random.default_rng(5)
    opd = rng.normal(size=2 * 64) * 5e-8
    L = m.lotf(rng.normal(size=64) * 5e-8)
    e1, e2 = _etf_tables(m, rng), _etf_tables(m, rng)
    i1 = m.ints(opd, lotf=L, etf=e1, llt=True)
    i2 = m.ints(opd, lotf=L, etf=e2, llt=True)
    ib = m.ints(opd, lotf=L, etf=e1, etf2=e2, etf2wt=w, llt=True)
    assert np.abs(ib - ((1 - w) * i1 + w * i2)).max() <= 1e-12 * np.abs(i1).max()


def test_etf_phase_ramp_moves_the_image_per_subaperture():
    """ETF(u) = nominal(u) exp(-2 pi i (ux mx + uy my) / notf) moves the
    image of subaperture isa by +(mx, my) samples: x is the fast axis of
    the table and each subaperture reads its own table."""
    m = Mini(nsa=3)
    N = m.notf
    uy, ux = np.mgrid[0:N, 0:N]
    shifts = [(1, 0), (0, 2), (-2, 3)]
    etf = np.stack([m.nominal.reshape(N, N) * np.exp(-2j * np.pi * (ux * sx + uy * sy) / N)
                    for sx, sy in shifts])
    rng = np.random.default_rng(9)
    opd = rng.normal(size=3 * 64) * 5e-8
    L = m.lotf(np.zeros(64))
    base = m.ints(opd, lotf=L)
    got = m.ints(opd, lotf=L, etf=etf, llt=True)
    for isa, (sx, sy) in enumerate(shifts):
        want = np.roll(np.roll(base[isa], sy, 0), sx, 1)
        assert np.abs(got[isa] - want).max() <= 1e-11 * base[isa].max()
