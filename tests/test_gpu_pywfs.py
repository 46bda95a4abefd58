"""Pyramid WFS on the GPU (mx_pywfs_*) against the FP64 oracle
(orc_pywfs_ints / orc_pywfs_grad, lib/pywfs.c:370-522) on identical tables.

Bars: intensities relative L2 <= 2e-5, gradients <= 0.1 mas absolute.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import RAD2MAS, rel_l2

pytestmark = pytest.mark.gpu

INTS_TOL = 2e-5
GRAD_TOL_MAS = 0.1


def _pw(**kw):
    from maos_amd.pywfs import Pywfs, PywfsCfg
    c = dict(D=2.0, din=0.3, order=8, dx=1.0 / 16, saat=0.3, siglev=1000.0)
    c.update(kw)
    return Pywfs(PywfsCfg(**c))


def _opd(pw, seed, rms=1e-7):
    """A smooth random OPD: low-order polynomials plus a little noise."""
    rng = np.random.default_rng(seed)
    x, y = pw.locx / (pw.cfg.D / 2), pw.locy / (pw.cfg.D / 2)
    o = sum(rng.normal() * x ** i * y ** j for i in range(4) for j in range(4 - i))
    o = o + 0.05 * rng.normal(size=o.size)
    return o * rms / o.std()


def _check(tag, got, want):
    (gi, gg), (ri, rg) = got, want
    e_int = rel_l2(gi, ri)
    e_g = np.abs(gg - rg).max() * RAD2MAS
    msg = "%s ints=%.2e grad=%.3e mas (grad rms %.3e mas)" % (tag, e_int, e_g, np.sqrt((rg ** 2).mean()) * RAD2MAS)
    print(msg, flush=True)
    assert e_int <= INTS_TOL, msg
    assert e_g <= GRAD_TOL_MAS, msg


@pytest.mark.parametrize("nside", [4, 3])
def test_pywfs_images_match_oracle(nside):
    from maos_amd import engine
    pw = _pw(nside=nside)
    pw.calibrate(lambda p: (lambda opd: oracle.pywfs_image(p, opd)))
    print("case: pyramid 2 m, N=%d, nsa=%d, %d sides" % (pw.N, pw.nsa, nside), flush=True)
    eng = engine.PywfsEngine(pw)
    for seed in (1, 2):
        opd = _opd(pw, seed)
        _check("seed %d" % seed, eng.image(opd), oracle.pywfs_image(pw, opd))
    _check("flat", eng.image(np.zeros(pw.nloc)), oracle.pywfs_image(pw, np.zeros(pw.nloc)))
    eng.close()


def test_pywfs_calibration_on_the_gpu():
    """pywfs_new's calibration run through the GPU imaging equals the
    oracle's: the pruned subapertures, saa, gradoff and gain."""
    from maos_amd import engine
    ref = _pw().calibrate(lambda p: (lambda opd: oracle.pywfs_image(p, opd)))

    def gpu(p):
        def f(opd):
            e = engine.PywfsEngine(p)
            out = e.image(opd)
            e.close()
            return out
        return f
    got = _pw().calibrate(gpu)
    assert got.nsa == ref.nsa and np.array_equal(got.salocx, ref.salocx)
    assert np.abs(got.saa - ref.saa).max() < 1e-5
    assert abs(got.gain / ref.gain - 1) < 1e-4
    assert np.abs(got.gradoff - ref.gradoff).max() * RAD2MAS < GRAD_TOL_MAS


def test_pywfs_radix7_grid():
    """An 8 m pupil at 1/32 m embeds in a 784 = 2^4 7^2 grid (MAOS's
    nextfftsize), which takes the radix-7 stages; 8 modulation points."""
    from maos_amd import engine
    pw = _pw(D=8.0, din=1.2, order=16, dx=1.0 / 32, modulpos=8, saat=0.0)
    assert pw.N == 784
    # gain and offsets calibrated (through the oracle's images), so that the
    # gradients are angles and the 0.1 mas bar applies
    pw.calibrate(lambda p: (lambda opd: oracle.pywfs_image(p, opd)))
    print("case: pyramid 8 m, N=784, nsa=%d" % pw.nsa, flush=True)
    eng = engine.PywfsEngine(pw)
    opd = _opd(pw, 5, rms=2e-7)
    _check("784", eng.image(opd), oracle.pywfs_image(pw, opd))
    eng.close()


def test_pywfs_step_traces_the_screens(ngs_small):
    """mx_pywfs_queue: the screens traced onto the pupil points at a step
    (FP64 coordinates), then imaged, against the oracle's trace and image."""
    import ctypes as C
    from maos_amd import engine
    cfg, _, layers = ngs_small
    pw = _pw(thetax=2e-6, thetay=-1e-6)
    pw.calibrate(lambda p: (lambda opd: oracle.pywfs_image(p, opd)))
    engine.init(0)
    engine.atm2gpu(layers)
    eng = engine.PywfsEngine(pw)
    dt, isim = cfg["sim"].dt, 77
    eng.queue(isim, dt)
    got = eng.sync()
    opd = np.zeros(pw.nloc)
    dp = C.POINTER(C.c_double)
    for L in layers:
        scr = np.ascontiguousarray(L.p)
        oracle.lib().orc_prop(scr.ctypes.data_as(dp), L.nx, L.ny, L.dx, L.ox, L.oy, L.ht,
                              pw.locx.ctypes.data_as(dp), pw.locy.ctypes.data_as(dp), pw.nloc, 1, L.dx,
                              opd.ctypes.data_as(dp), 1.0, 0.0, 2e-6, -1e-6, 0.0, 0.0,
                              -L.vx * dt * isim, -L.vy * dt * isim, 1)
    e_opd = np.abs(eng.opd() - opd).max() / np.abs(opd).max()
    print("trace opd max rel %.2e" % e_opd, flush=True)
    assert e_opd < 1e-5
    _check("step", got, oracle.pywfs_image(pw, opd))
    eng.close()
