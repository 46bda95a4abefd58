"""Science-field evaluation on the GPU (mx_evl_*, through the C ABI) against
the FP64 oracle: the accumulated PSF (locfft_psf / psfcomp_r), the
Strehl-only form, and the ray trace of screens and DM surfaces onto the
aperture points.

Tolerances: PSF relative L2 <= 1e-5 of the oracle's (FP32 FFT over at most
4096 points, FP32 OPD; the phase is reduced in FP64), peak within 1e-5; OPD
within 1e-5 of its RMS (FP32 screens and sums).
"""
import numpy as np
import pytest

from conftest import rel_l2
from test_evl_psf import annulus

pytestmark = pytest.mark.gpu

PSF_TOL = 1e-5


def _case(D, din, dx, N, seed=0):
    from maos_amd import engine
    locx, locy = annulus(D, din, dx)
    emb, n = engine.loc_embed(locx, locy, dx, nembed=N)
    assert n == N
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.6, 1.0, len(locx))
    amp /= amp.sum()
    return locx, locy, amp, emb


def _oracle_mean(opds, amp, emb, N, P, wvl):
    import oracle
    acc = sum(oracle.locfft_psf(o, amp, emb, N, P, wvl) for o in opds)
    return acc / (len(opds) * amp.sum() ** 2)


@pytest.mark.parametrize("D,din,dx,N,P", [
    (1.0, 0.2, 1 / 16, 32, 32),
    (1.0, 0.2, 1 / 16, 32, 2),
    (2.0, 0.4, 1 / 16, 64, 64),
    (2.0, 0.4, 1 / 16, 64, 6),
    (3.5, 0.5, 1 / 16, 128, 128),
    (4.0, 0.5, 1 / 32, 256, 96),
    (8.0, 1.0, 1 / 32, 512, 512),
    (8.0, 1.0, 1 / 64, 1024, 256),
    (16.0, 2.0, 1 / 64, 2048, 2048),
    (30.0, 3.6, 1 / 64, 4096, 256),
])
def test_psf_parity(D, din, dx, N, P):
    from maos_amd import engine
    print("case: psf D=%g N=%d P=%d" % (D, N, P), flush=True)
    locx, locy, amp, emb = _case(D, din, dx, N)
    nloc = len(amp)
    if N <= 256:
        emb = emb.copy()
        emb[11] = emb[5]  # a duplicated index: the later point is used
    rng = np.random.default_rng(N + P)
    wvl = 1.0e-6
    ndir, nstep = 2, 2
    opds = [rng.normal(size=(ndir, nloc)) * 1.5e-7 for _ in range(nstep)]
    ev = engine.Evl(amp, [wvl], [N], [P], emb, dirs=[(0, 0, 0)] * ndir)
    for o in opds:
        ev.psf_add(o)
    for d in range(ndir):
        got, nacc = ev.psf(d, 0)
        assert nacc == nstep
        want = _oracle_mean([o[d] for o in opds], amp, emb, N, P, wvl)
        e = rel_l2(got, want)
        print("  dir %d rel-L2 %.2e" % (d, e), flush=True)
        assert e <= PSF_TOL
        assert got.max() == pytest.approx(want.max(), rel=PSF_TOL)
    ev.close()


def test_psf_two_wavelengths_and_strehl():
    """Two wavelengths with FFT sizes 64 and 128 and a Strehl-only third, one
    handle; the flat wavefront's peak is the Strehl ratio 1."""
    from maos_amd import engine
    locx, locy = annulus(2.0, 0.4, 1 / 16)
    amp = np.full(len(locx), 1.0 / len(locx))
    e64, _ = engine.loc_embed(locx, locy, 1 / 16, nembed=64)
    e128, _ = engine.loc_embed(locx, locy, 1 / 16, nembed=128)
    wvls = [0.9e-6, 1.6e-6, 2.2e-6]
    ev = engine.Evl(amp, wvls, [64, 128, 64], [32, 64, 1], np.stack([e64, e128, e64]),
                    dirs=[(0, 0, 0), (1e-5, 0, 0), (0, 2e-5, 9e4)])
    ev.psf_add(np.zeros((3, len(amp))))
    flat, _ = ev.psf(1, 0)
    assert np.argmax(flat) == 16 + 16 * 32
    assert flat.max() == pytest.approx(1.0, rel=1e-6)
    ev.reset()
    _, n = ev.psf(0, 0)
    assert n == 0 and not ev.psf(0, 2)[0].any()
    rng = np.random.default_rng(9)
    opds = [rng.normal(size=(3, len(amp))) * 2e-7 for _ in range(3)]
    for o in opds:
        ev.psf_add(o)
    for d in range(3):
        for iw, (emb, N, P) in enumerate([(e64, 64, 32), (e128, 128, 64), (e64, 64, 1)]):
            got, nacc = ev.psf(d, iw)
            assert nacc == 3
            want = _oracle_mean([o[d] for o in opds], amp, emb, N, P, wvls[iw])
            assert rel_l2(got, want) <= PSF_TOL, (d, iw)
    ev.close()


def _hip_runtime():
    """The HIP runtime the library links (libamdhip64.so.7 of /opt/rocm),
    for device buffers owned by the test."""
    import ctypes as C
    from maos_amd import _lib
    _lib.hip()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    return rt


def test_psf_from_device_opd():
    """An FP32 OPD already in device memory gives the PSF the host entry
    point gives."""
    import ctypes as C
    from maos_amd import engine
    locx, locy, amp, emb = _case(4.0, 0.5, 1 / 32, 256, seed=4)
    rng = np.random.default_rng(4)
    opd = np.ascontiguousarray((rng.normal(size=(1, len(amp))) * 1e-7).astype(np.float32))
    a = engine.Evl(amp, [1.2e-6], [256], [128], emb)
    b = engine.Evl(amp, [1.2e-6], [256], [128], emb)
    a.psf_add(opd.astype(np.float64))
    rt = _hip_runtime()
    d = C.c_void_p()
    assert rt.hipMalloc(C.byref(d), opd.nbytes) == 0
    try:
        assert rt.hipMemcpy(d, opd.ctypes.data_as(C.c_void_p), opd.nbytes, 1) == 0  # host to device
        b.psf_add_device(d.value)
        b.wait()
        np.testing.assert_array_equal(a.psf()[0], b.psf()[0])
    finally:
        b.wait()
        rt.hipFree(d)
    a.close()
    b.close()


def _oracle_opd(layers, dms, locx, locy, d, isim, dt, atmscale=1.0):
    import ctypes as C
    import oracle
    dp = C.POINTER(C.c_double)
    p = lambda x: np.ascontiguousarray(x, dtype=np.float64).ctypes.data_as(dp)
    tx, ty, hs = d
    out = np.zeros(len(locx))
    lx, ly = np.ascontiguousarray(locx), np.ascontiguousarray(locy)
    for L in layers:
        if hs and L.ht > hs:
            continue
        scr = np.ascontiguousarray(L.p)
        oracle.lib().orc_prop(p(scr), L.nx, L.ny, L.dx, L.ox, L.oy, L.ht, p(lx), p(ly), len(lx), 1, L.dx,
                              out.ctypes.data_as(dp), atmscale, hs, tx, ty, 0.0, 0.0,
                              -L.vx * dt * isim, -L.vy * dt * isim, 1)
    for (m, mdx, ox, oy, ht) in dms:
        a = np.ascontiguousarray(m)
        oracle.lib().orc_prop(p(a), a.shape[1], a.shape[0], mdx, ox, oy, ht, p(lx), p(ly), len(lx), 1, mdx,
                              out.ctypes.data_as(dp), -1.0, hs, tx, ty, 0.0, 0.0, 0.0, 0.0, 0)
    return out


def test_trace_and_psf_end_to_end(lgs_small):
    """Screens (7 layers, frozen flow at steps 0 and 5) and two DM surfaces
    traced onto a 10 m aperture for an on-axis, an off-axis and a
    finite-height direction; then the PSF of the traced OPD."""
    from maos_amd import engine
    cfg, p, layers = lgs_small
    print("case: evl trace 10 m, %d layers" % len(layers), flush=True)
    dx = 1 / 64
    locx, locy = annulus(10.0, 1.2, dx)
    emb, N = engine.loc_embed(locx, locy, dx, nembed=2048)
    amp = np.full(len(locx), 1.0 / len(locx))
    rng = np.random.default_rng(12)
    dms = [(rng.normal(size=(340, 340)) * 5e-8, 1 / 16, -170 / 16, -170 / 16, 0.0),
           (rng.normal(size=(420, 420)) * 5e-8, 1 / 16, -210 / 16, -210 / 16, 11000.0)]
    engine.init(0)
    engine.atm2gpu(layers, atmscale=1.1)
    engine.dm2gpu(dms)
    try:
        _trace_checks(engine, layers, dms, locx, locy, amp, emb, N)
    finally:
        engine.dm2gpu([])  # the DM surfaces are global: later tests run without them


def _trace_checks(engine, layers, dms, locx, locy, amp, emb, N):
    import oracle
    dirs = [(0.0, 0.0, 0.0), (3e-5, -2e-5, 0.0), (1e-5, 1e-5, 90e3)]
    ev = engine.Evl(amp, [1.65e-6], [N], [512], emb, locx=locx, locy=locy, dirs=dirs)
    dt = 1 / 800
    opds = []
    for isim in (0, 5):
        ev.trace(isim, dt, add_psf=True)
        got = ev.opd()
        want = np.stack([_oracle_opd(layers, dms, locx, locy, d, isim, dt, 1.1) for d in dirs])
        rms = np.sqrt(np.mean(want ** 2))
        err = np.abs(got - want).max()
        print("  isim %d OPD max err %.2e of rms %.2e" % (isim, err, rms), flush=True)
        assert err <= 1e-5 * rms
        opds.append(want)
        # wavefront error of the traced OPD against the oracle's, on the oracle OPD
        g_rms, g_coeff = ev.wfe()
        for d in range(len(dirs)):
            o_rms, o_coeff = oracle.loc_calc_ptt(locx, locy, amp, want[d])
            np.testing.assert_allclose(g_rms[d], o_rms, rtol=1e-5)
            np.testing.assert_allclose(g_coeff[d], o_coeff, rtol=1e-5, atol=1e-6 * np.abs(o_coeff).max())
    for d in range(len(dirs)):
        got, nacc = ev.psf(d, 0)
        assert nacc == 2
        want = _oracle_mean([o[d] for o in opds], amp, emb, N, 512, 1.65e-6)
        e = rel_l2(got, want)
        print("  dir %d PSF rel-L2 %.2e" % (d, e), flush=True)
        assert e <= PSF_TOL
    ev.close()


def test_psf_edge_cases():
    """One aperture point: a flat PSF, every pixel (amp / sum amp)^2 = 1.
    The largest direction count (64) at a small FFT size, against the
    oracle."""
    from maos_amd import engine
    ev = engine.Evl([0.37], [1e-6], [32], [8], [5 + 7 * 32])
    ev.psf_add(np.array([[2.3e-7]]))
    got, n = ev.psf(0, 0)
    assert n == 1
    np.testing.assert_allclose(got, 1.0, rtol=1e-6)
    ev.close()
    locx, locy, amp, emb = _case(1.0, 0.2, 1 / 16, 32, seed=8)
    rng = np.random.default_rng(8)
    ndir = 64
    opd = rng.normal(size=(ndir, len(amp))) * 2e-7
    ev = engine.Evl(amp, [0.8e-6], [32], [16], emb, dirs=[(0, 0, 0)] * ndir)
    ev.psf_add(opd)
    for d in (0, 31, 63):
        got, _ = ev.psf(d, 0)
        want = _oracle_mean([opd[d]], amp, emb, 32, 16, 0.8e-6)
        assert rel_l2(got, want) <= PSF_TOL, d
    ev.close()


def test_trace_other_realisation(lgs_small):
    """mx_evl_set_real: the trace reads realisation 1's screens (here the
    negated screens of realisation 0), then realisation 0's again."""
    from maos_amd import engine
    from maos_amd.powfs import Layer
    cfg, p, layers = lgs_small
    neg = [Layer(-L.p, L.dx, L.ht, L.vx, L.vy) for L in layers]
    engine.init(0)
    engine.atm2gpu(layers)
    engine.atm2gpu(neg, ireal=1)
    locx, locy = annulus(4.0, 0.5, 1 / 64)
    emb, N = engine.loc_embed(locx, locy, 1 / 64, nembed=1024)
    amp = np.full(len(locx), 1.0 / len(locx))
    dirs = [(2e-5, 1e-5, 0.0)]
    ev = engine.Evl(amp, [1e-6], [N], [64], emb, locx=locx, locy=locy, dirs=dirs)
    ev.set_real(1)
    ev.trace(3, 1 / 800, add_psf=False)
    want = _oracle_opd(neg, [], locx, locy, dirs[0], 3, 1 / 800)
    got = ev.opd()[0]
    assert np.abs(got - want).max() <= 1e-5 * np.sqrt(np.mean(want ** 2))
    ev.set_real(0)
    ev.trace(3, 1 / 800, add_psf=False)
    np.testing.assert_allclose(ev.opd()[0], -got, rtol=0, atol=1e-12)
    ev.close()


def test_reupload_waits_for_queued_trace(lgs_small):
    """Screens are overwritten in place: an upload right after an
    asynchronous trace (and a queued WFS step) waits for them, so the traced
    OPD is that of the old screens and the next trace sees the new ones."""
    import dataclasses
    from maos_amd import configs, engine
    from maos_amd.powfs import genatm
    cfg, p, layers = lgs_small
    print("case: screen upload after a queued trace", flush=True)
    layers2 = genatm(dataclasses.replace(cfg["atm"], seeds=[s + 50 for s in cfg["atm"].seeds]))
    dx = 1 / 64
    locx, locy = annulus(10.0, 1.2, dx)
    emb, N = engine.loc_embed(locx, locy, dx, nembed=2048)
    amp = np.full(len(locx), 1.0 / len(locx))
    engine.init(0)
    engine.atm2gpu(layers)
    dirs = [(0.0, 0.0, 0.0), (2e-5, 1e-5, 0.0)]
    ev = engine.Evl(amp, [1.65e-6], [N], [256], emb, locx=locx, locy=locy, dirs=dirs)
    wfs = configs.wfs_specs(cfg, p)
    eng = engine.WfsGrad(p, wfs)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    dt = 1 / 800
    for step, (old, new) in enumerate([(layers, layers2), (layers2, layers)]):
        isim = 3 + step
        ev.trace(isim, dt, add_psf=False)
        eng.queue(isim, dt)
        engine.atm2gpu(new)  # must wait for the trace and the step above
        got = ev.opd()
        want = np.stack([_oracle_opd(old, [], locx, locy, d, isim, dt, 1.0) for d in dirs])
        err = np.abs(got - want).max() / np.sqrt(np.mean(want ** 2))
        other = np.abs(got - np.stack([_oracle_opd(new, [], locx, locy, d, isim, dt, 1.0) for d in dirs])).max()
        print("  step %d: OPD error %.2e of rms (against the new screens %.2e)" % (step, err, other), flush=True)
        assert err <= 1e-5 and other > 100 * err * np.sqrt(np.mean(want ** 2))
        wopd = eng.debug_opd(0)
        import oracle
        o = oracle.Oracle(p)
        o.atm(old)
        _, _, ropd, _ = o.step(wfs[0], isim, dt, want_opd=True)
        assert np.abs(wopd - ropd).max() <= 1e-5 * np.abs(ropd).max()
    eng.close()
    ev.close()
    engine.atm2gpu(layers)
