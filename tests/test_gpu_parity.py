"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: pixel intensities relative L2 <= 2e-5 per
frame, gradients <= 0.1 mas absolute.  Inputs are identical seeded screens
and identical setup tables for both paths.
"""
import numpy as np
import pytest

from conftest import RAD2MAS, rel_l2

pytestmark = pytest.mark.gpu

INTS_TOL = 2e-5
GRAD_TOL_MAS = 0.1


def _engine(cfg, p, layers):
    from maos_amd import configs, engine
    engine.init(0)
    engine.atm2gpu(layers)
    wfs = configs.wfs_specs(cfg, p)
    return engine.WfsGrad(p, wfs), wfs


def _oracle(p, layers, mtche=None, i0sum=None):
    import oracle
    o = oracle.Oracle(p, mtche, i0sum)
    o.atm(layers)
    return o


def test_ngs8m_step_parity(ngs_small):
    from maos_amd import _lib
    cfg, p, layers = ngs_small
    print("case: ngs8m nsa=%d notf=%d 512^2 screen" % (p.nsa, p.notf), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    assert eng.variant() == _lib.VARIANT_FFT32  # notf=32: the MFMA form does not apply
    with pytest.raises(RuntimeError):
        eng.set_variant(_lib.VARIANT_MFMA_BF16)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    for isim in (0, 57, 99):
        eng.queue(isim, dt)
        g, ints = eng.sync(0)
        opd = eng.debug_opd(0)
        ri, rg, ropd, _ = o.step(wfs[0], isim, dt, want_opd=True)
        e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
        e_int = rel_l2(ints, ri)
        e_g = np.abs(g - rg).max() * RAD2MAS
        print("isim=%d opd maxrel=%.2e ints relL2=%.2e grad maxabs=%.3e mas" % (isim, e_opd, e_int, e_g))
        assert e_opd < 1e-5
        assert e_int <= INTS_TOL
        assert e_g <= GRAD_TOL_MAS
    eng.close()


def _lgs_ref(p, o, wfs, dt, isim=3):
    """Matched filter built from the oracle's images of one frame."""
    ri, _, _, _ = o.step(wfs[0], isim, dt)
    mtche, i0sum, _ = p.mtch(ri)
    return mtche, i0sum


# _lib.VARIANT_*; fft32_reg is the FP32 form with its register-FFT kernel
# (MX_FFT_REG=1), which it also runs where the half-plane tables do not apply
VARIANTS = {"fft32": 0, "mfma_bf16": 1, "mfma_fp16": 2, "fft32_reg": 0, "mfma_bf16_fused": 1,
            "mfma_fp16_fused": 2}


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16", "mfma_fp16", "fft32_reg", "mfma_bf16_fused",
                                     "mfma_fp16_fused"])
def test_lgs_step_parity(lgs_small, variant, monkeypatch):
    """Every imaging form against the oracle; the _fused variants
    (MX_FUSED=1) trace the subaperture and LLT grids inside the imaging and
    uplink kernels."""
    cfg, p, layers = lgs_small
    if variant == "fft32_reg":
        monkeypatch.setenv("MX_FFT_REG", "1")
    if variant.endswith("_fused"):
        monkeypatch.setenv("MX_FUSED", "1")
    print("case: lgs order20 nsa=%d notf=%d nwfs=%d 1024^2 screens, %s"
          % (p.nsa, p.notf, len(cfg["wfs"]), variant), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    want = VARIANTS[variant]
    eng.set_variant(want)
    assert eng.variant() == want
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    o.set_ref(mtche, i0sum)
    eng.update_ref(mtche, i0sum)
    for isim in (0, 123):
        eng.queue(isim, dt)
        for iwfs in range(len(wfs)):
            g, ints = eng.sync(iwfs)
            ri, rg, ropd, rlotf = o.step(wfs[iwfs], isim, dt, want_opd=True, want_lotf=True)
            opd = eng.debug_opd(iwfs)
            lotf = eng.debug_lotf(iwfs)
            e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
            e_lotf = rel_l2(lotf, rlotf)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d opd=%.2e lotf=%.2e ints=%.2e grad=%.3e mas"
                  % (isim, iwfs, e_opd, e_lotf, e_int, e_g))
            assert e_opd < 1e-5
            # the MFMA forms compute the uplink OTF with the same split
            # products as the image, so they get the image's bar
            assert e_lotf < (1e-5 if variant.startswith("fft32") else INTS_TOL)
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16", "mfma_fp16", "fft32_reg"])
def test_ngs_two_wavelengths(ngs_two_wvl, variant, monkeypatch):
    """N=64 without LLT or ETF (nominal DTF only), two wavelengths with their
    own si and weights, two WFS off axis: both imaging forms against the
    oracle."""
    from maos_amd import _lib
    cfg, p, layers = ngs_two_wvl
    print("case: ngs two wavelengths nsa=%d notf=%d, %s" % (p.nsa, p.notf, variant), flush=True)
    if variant == "fft32_reg":
        monkeypatch.setenv("MX_FFT_REG", "1")
    eng, wfs = _engine(cfg, p, layers)
    assert eng.variant() == _lib.VARIANT_MFMA_BF16  # AUTO picks the MFMA form here
    want = VARIANTS[variant]
    eng.set_variant(want)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    for isim in (0, 41):
        eng.queue(isim, dt)
        for iwfs in range(len(wfs)):
            g, ints = eng.sync(iwfs)
            ri, rg, _, _ = o.step(wfs[iwfs], isim, dt)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d ints=%.2e grad=%.3e mas" % (isim, iwfs, e_int, e_g))
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


def test_i0_accumulation_fp64(lgs_small):
    """Long-exposure i0 is accumulated in FP64 on the GPU: it equals the sum
    of the per-frame images to double rounding, and the oracle's sum to the
    intensity tolerance (maos/wfsgrad.c:449-467)."""
    cfg, p, layers = lgs_small
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    eng.update_ref(mtche, i0sum)
    eng.i0_control(enable=True, reset=True)
    nstep = 5
    acc = np.zeros((p.nsa, p.npix))
    ref = np.zeros((p.nsa, p.npix))
    for isim in range(nstep):
        eng.queue(isim, dt)
        _, ints = eng.sync(2)
        acc += ints
        ref += o.step(wfs[2], isim, dt)[0]
    i0, nframe = eng.i0_get(2)
    assert nframe == nstep
    assert np.abs(i0 - acc).max() <= 1e-12 * np.abs(acc).max()
    assert rel_l2(i0, ref) <= INTS_TOL
    eng.i0_control(enable=False, reset=True)
    i0b, n0 = eng.i0_get(2)
    assert n0 == 0 and np.all(i0b == 0)
    eng.close()


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16"])
def test_graph_matches_direct(lgs_small, variant):
    """Steps launched as one HIP graph give bit-identical intensities,
    gradients and FP64 i0 sums to the same steps launched kernel by kernel.
    Graph replays alternate with the direct handle's eager launches and host
    synchronisations, and the graph's nodes get new arguments every step."""
    from maos_amd import _lib
    cfg, p, layers = lgs_small
    print("case: graph vs direct, lgs order20, %s" % variant, flush=True)
    want = {"fft32": _lib.VARIANT_FFT32, "mfma_bf16": _lib.VARIANT_MFMA_BF16}[variant]
    direct, wfs = _engine(cfg, p, layers)
    graph, _ = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    for eng in (direct, graph):
        eng.set_variant(want)
        eng.update_ref(mtche, i0sum)
        eng.i0_control(enable=True, reset=True)
    graph.set_graph(True)
    for isim in (0, 1, 2, 7, 8):
        direct.queue(isim, dt)
        gd, idd = direct.sync(isim % len(wfs))
        graph.queue(isim, dt)
        gg, ig = graph.sync(isim % len(wfs))
        print("isim=%d max |dgrad|=%.3e max |dints|=%.3e" % (isim, np.abs(gd - gg).max(), np.abs(idd - ig).max()))
        assert np.array_equal(gd, gg)
        assert np.array_equal(idd, ig)
    for iwfs in range(len(wfs)):
        i0d, nd = direct.i0_get(iwfs)
        i0g, ng = graph.i0_get(iwfs)
        assert nd == ng == 5
        assert np.array_equal(i0d, i0g)
    # a change of the kernel set rebuilds the graph
    other = _lib.VARIANT_MFMA_BF16 if want == _lib.VARIANT_FFT32 else _lib.VARIANT_FFT32
    for eng in (direct, graph):
        eng.set_variant(other)
        eng.queue(9, dt)
    gd, idd = direct.sync(0)
    gg, ig = graph.sync(0)
    assert np.array_equal(gd, gg) and np.array_equal(idd, ig)
    direct.close()
    graph.close()


def test_batch_realisations(lgs_small):
    """Two realisations in one engine (one launch per kernel per step):
    realisation 0 is bit-identical to a one-realisation engine on the same
    screens, realisation 1 (other seeds and r0) matches the oracle on its own
    screens."""
    import dataclasses
    from maos_amd import configs, engine
    from maos_amd.powfs import genatm
    cfg, p, layers0 = lgs_small
    atm1 = dataclasses.replace(cfg["atm"], r0=0.14, seeds=[s + 100 for s in cfg["atm"].seeds])
    layers1 = genatm(atm1)
    print("case: batch of 2 realisations, lgs order20", flush=True)
    engine.init(0)
    engine.atm2gpu(layers0, ireal=0)
    engine.atm2gpu(layers1, ireal=1)
    wfs = configs.wfs_specs(cfg, p)
    nw = len(wfs)
    batch = engine.WfsGrad(p, wfs, nreal=2)
    single = engine.WfsGrad(p, wfs)
    o0, o1 = _oracle(p, layers0), _oracle(p, layers1)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o0, wfs, dt)
    for eng in (batch, single):
        eng.update_ref(mtche, i0sum)
    o1.set_ref(mtche, i0sum)
    isim = 17
    batch.queue(isim, dt)
    single.queue(isim, dt)
    for iwfs in range(nw):
        gb, ib = batch.sync(iwfs)
        gs, is_ = single.sync(iwfs)
        assert np.array_equal(gb, gs) and np.array_equal(ib, is_)
        g1, i1 = batch.sync(nw + iwfs)
        ri, rg, _, _ = o1.step(wfs[iwfs], isim, dt)
        e_int = rel_l2(i1, ri)
        e_g = np.abs(g1 - rg).max() * RAD2MAS
        print("real 1 wfs=%d ints=%.2e grad=%.3e mas" % (iwfs, e_int, e_g))
        assert e_int <= INTS_TOL
        assert e_g <= GRAD_TOL_MAS
        assert rel_l2(i1, ib) > 1e-2  # the realisations differ
    batch.close()
    single.close()


@pytest.mark.parametrize("case", ["lf_mixed", "ngs_lo_small", "ngs_tt_small"])
def test_large_field_parity(case, request):
    """The large-field form (line FFTs of any 2/3/5-smooth length <= 1024,
    ccpcorner padding, screens finer than the OPD grid) against the oracle,
    every WFS, two steps; and its FP64 i0 accumulation."""
    from maos_amd import _lib
    cfg, p, layers = request.getfixturevalue(case)
    print("case: %s nsa=%d nxsa=%d nwvf=%d notf=%d nwvl=%d nwfs=%d"
          % (case, p.nsa, p.nxsa, p.nxsa * p.cfg.embfac, p.notf, p.nwvl, len(cfg["wfs"])), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    assert eng.variant() == _lib.VARIANT_FFT32
    assert not eng.variant_ok(_lib.VARIANT_MFMA_BF16)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    eng.i0_control(True, reset=True)
    acc = [np.zeros((p.nsa, p.npix)) for _ in wfs]
    for isim in (0, 311):
        eng.queue(isim, dt)
        for iw in range(len(wfs)):
            g, ints = eng.sync(iw)
            opd = eng.debug_opd(iw)
            ri, rg, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
            acc[iw] += ints
            e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d opd maxrel=%.2e ints relL2=%.2e grad maxabs=%.3e mas"
                  % (isim, iw, e_opd, e_int, e_g))
            assert e_opd < 1e-5
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    for iw in range(len(wfs)):
        i0, n = eng.i0_get(iw)
        assert n == 2
        assert np.array_equal(i0, acc[iw])  # FP32 frames summed in FP64 exactly
    eng.close()


def test_group_matches_single(lgs_small, ngs_lo_small, ngs_tt_small):
    """A heterogeneous step (LGS on the MFMA form, low-order NGS and TT NGS
    on the large-field form) queued as one group graph gives bit-identical
    gradients, intensities and i0 sums to each engine queued on its own,
    over steps whose kernel arguments change, and the group is rebuilt when
    an engine's kernel set changes.  The three fixtures share one
    atmosphere."""
    from maos_amd import _lib, configs, engine
    cfg_l, p_l, layers = lgs_small
    print("case: group of lgs order20 (mfma) + ngs_lo + ngs_tt", flush=True)
    engine.init(0)
    engine.atm2gpu(layers)
    cases = [(cfg_l, p_l), ngs_lo_small[:2], ngs_tt_small[:2]]
    o = _oracle(p_l, layers)
    dt = cfg_l["sim"].dt
    wfs_l = configs.wfs_specs(cfg_l, p_l)
    mtche, i0sum = _lgs_ref(p_l, o, wfs_l, dt)
    pairs = []
    for cfg, p in cases:
        wfs = configs.wfs_specs(cfg, p)
        a, b = engine.WfsGrad(p, wfs), engine.WfsGrad(p, wfs)
        for eng in (a, b):
            if p.llt_nx:
                eng.update_ref(mtche, i0sum)
            eng.i0_control(enable=True, reset=True)
        pairs.append((a, b, len(wfs)))
    group = engine.WfsGroup([b for _, b, _ in pairs])
    for isim in (0, 1, 5, 6):
        if isim == 5:  # another kernel set for the LGS engine: the group graph is rebuilt
            for eng in pairs[0][:2]:
                eng.set_variant(_lib.VARIANT_FFT32)
        for a, _, _ in pairs:
            a.queue(isim, dt)
        group.queue(isim, dt)
        for a, b, nw in pairs:
            for iw in range(nw):
                ga, ia = a.sync(iw)
                gb, ib = b.sync(iw)
                assert np.array_equal(ga, gb) and np.array_equal(ia, ib)
    for a, b, nw in pairs:
        for iw in range(nw):
            i0a, na = a.i0_get(iw)
            i0b, nb = b.i0_get(iw)
            assert na == nb == 4
            assert np.array_equal(i0a, i0b)
    group.close()
    for a, b, _ in pairs:
        a.close()
        b.close()


@pytest.mark.parametrize("case", ["lgs_small", "ngs_lo_small"])
def test_autotune(case, request):
    """The autotuner times every candidate, keeps one, leaves the i0 sums
    and frame count alone, and the chosen configuration matches the
    oracle."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: autotune %s" % case, flush=True)
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    if p.llt_nx:
        mtche, i0sum = _lgs_ref(p, o, wfs, dt)
        o.set_ref(mtche, i0sum)
        eng.update_ref(mtche, i0sum)
    eng.i0_control(enable=True, reset=True)
    eng.queue(0, dt)
    i0a, na = eng.i0_get(0)
    tune = eng.autotune(3, dt, nrep=3)
    print(tune)
    assert sum(t["best"] for t in tune) == 1
    assert len(tune) >= 2 and all(t["step_us"] > 0 for t in tune)
    i0b, nb = eng.i0_get(0)
    assert na == nb == 1 and np.array_equal(i0a, i0b)
    eng.queue(40, dt)
    g, ints = eng.sync(0)
    ri, rg, _, _ = o.step(wfs[0], 40, dt)
    assert rel_l2(ints, ri) <= INTS_TOL
    assert np.abs(g - rg).max() * RAD2MAS <= GRAD_TOL_MAS
    eng.close()


def test_lgs_pad128_parity(lgs_pad128):
    """The 128^2 padding variant of the LGS (embfac 4) on the large-field
    form: uplink OTF from the LLT grid, x conj(lotf) x ETF per subaperture,
    matched filter; against the oracle for every WFS at two steps."""
    from maos_amd import _lib
    cfg, p, layers = lgs_pad128
    print("case: lgs order20 embfac4 nsa=%d notf=%d nwfs=%d" % (p.nsa, p.notf, len(cfg["wfs"])), flush=True)
    assert p.notf == 128 and p.llt_nx * p.cfg.embfac == 128
    eng, wfs = _engine(cfg, p, layers)
    assert eng.variant() == _lib.VARIANT_FFT32
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    o.set_ref(mtche, i0sum)
    eng.update_ref(mtche, i0sum)
    for isim in (0, 123):
        eng.queue(isim, dt)
        for iwfs in range(len(wfs)):
            g, ints = eng.sync(iwfs)
            ri, rg, _, rlotf = o.step(wfs[iwfs], isim, dt, want_lotf=True)
            lotf = eng.debug_lotf(iwfs)
            e_lotf = rel_l2(lotf, rlotf)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d lotf=%.2e ints=%.2e grad=%.3e mas" % (isim, iwfs, e_lotf, e_int, e_g))
            assert e_lotf < 1e-5
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("case,variant", [("ngs_two_wvl", "mfma_bf16"), ("ngs_two_wvl", "fft32"),
                                          ("lgs_small", "mfma_bf16"), ("lf_mixed", "fft32")])
def test_noise_matches_oracle(case, variant, request):
    """Photon and read noise: a noisy engine against the oracle's addnoise
    (the same Philox streams) applied to a noise-free engine's intensities of
    the same step; gradients and i0 sums use the noisy intensities."""
    from maos_amd import configs, engine
    cfg, p, layers = request.getfixturevalue(case)
    print("case: noise %s %s" % (case, variant), flush=True)
    engine.init(0)
    engine.atm2gpu(layers)
    wfs = configs.wfs_specs(cfg, p)
    quiet, noisy = engine.WfsGrad(p, wfs), engine.WfsGrad(p, wfs)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    if p.llt_nx:
        mtche, i0sum = _lgs_ref(p, o, wfs, dt)
        o.set_ref(mtche, i0sum)
    seed, rne, bkgrnd, bkgrndc = (1 << 40) + 12345, 3.0, 5.0, 4.0
    for eng in (quiet, noisy):
        eng.set_variant(VARIANTS[variant])
        if p.llt_nx:
            eng.update_ref(mtche, i0sum)
        eng.i0_control(enable=True, reset=True)
    noisy.set_noise(True, rne=rne, bkgrnd=bkgrnd, bkgrndc=bkgrndc, seed=seed)
    isim = 9
    quiet.queue(isim, dt)
    noisy.queue(isim, dt)
    import oracle
    for iw in range(len(wfs)):
        _, ints_q = quiet.sync(iw)
        g_n, ints_n = noisy.sync(iw)
        ref = np.stack([oracle.addnoise(ints_q[isa], bkgrnd, bkgrndc, rne, seed, iw * p.nsa + isa, isim)
                        for isa in range(p.nsa)])
        err = np.abs(ints_n - ref) / np.maximum(1.0, np.abs(ref))
        print("wfs=%d max rel diff %.2e, mean noise %.3f" % (iw, err.max(), (ints_n - ints_q).mean()))
        assert err.max() < 1e-5
        assert np.abs(ints_n - ints_q).max() > 1.0  # noise present
        g_ref = o.grad(ints_n)
        vign = p.saa < p.cfg.saat
        e_g = np.abs(g_n - g_ref)[np.concatenate([~vign, ~vign])].max() * RAD2MAS
        print("   grad from the noisy ints: %.3e mas" % e_g)
        assert e_g < 0.01
        i0, n = noisy.i0_get(iw)
        assert n == 1 and np.allclose(i0, ints_n, rtol=0, atol=0)
    quiet.close()
    noisy.close()


def test_recorder_writes_maos_telemetry(ngs_small, tmp_path):
    """Recorder writes each step's gradients and intensities as MAOS zfarr
    files that read back equal to what sync returned."""
    from maos_amd import telemetry
    cfg, p, layers = ngs_small
    print("case: telemetry recorder ngs8m", flush=True)
    eng, wfs = _engine(cfg, p, layers)
    rec = telemetry.Recorder(str(tmp_path), eng, 4, seed=1, ints=True)
    got = []
    for isim in range(3):
        eng.queue(isim, cfg["sim"].dt)
        rec.record(isim)
        got.append(eng.sync(0))
    rec.close()
    grads = telemetry.read_bin(str(tmp_path / "wfs0_gradcl_1.bin"))
    ints = telemetry.read_bin(str(tmp_path / "wfs0_intsnf_1.bin"))
    assert len(grads) == 4 and grads[3] is None
    for isim in range(3):
        assert np.array_equal(grads[isim].ravel(), got[isim][0])
        assert np.array_equal(ints[isim].reshape(p.nsa, p.npix), got[isim][1])
    eng.close()


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16"])
def test_opd_inputs_surface_and_tip_tilt(ngs_two_wvl, variant):
    """A static surface OPD per WFS and a per-step piston/tip/tilt enter the
    WFS OPD (debug_opd equals the oracle's traced OPD plus both), and the
    intensities and gradients equal the oracle's from that OPD."""
    cfg, p, layers = ngs_two_wvl
    print("case: opdadd + ptt, ngs two wavelengths, %s" % variant, flush=True)
    eng, wfs = _engine(cfg, p, layers)
    eng.set_variant(VARIANTS[variant])
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    n = p.nxsa
    rs = np.random.RandomState(17)
    surf = [5e-8 * rs.standard_normal(p.nsa * n * n) for _ in wfs]
    ptt = [np.array([1e-8, 2e-7, -1.5e-7]), np.array([0.0, -3e-7, 1e-7])]
    for iw in range(len(wfs)):
        eng.set_opdadd(surf[iw], iwfs=iw)
        eng.set_ptt(ptt[iw], iwfs=iw)
    g = np.arange(n) * p.dx
    X = (p.ptsx[:, None, None] + g[None, None, :] + 0 * g[None, :, None]).reshape(-1)
    Y = (p.ptsy[:, None, None] + g[None, :, None] + 0 * g[None, None, :]).reshape(-1)
    isim = 21
    eng.queue(isim, dt)
    for iw in range(len(wfs)):
        gg, ints = eng.sync(iw)
        opd = eng.debug_opd(iw)
        _, _, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
        ropd = ropd + surf[iw] + ptt[iw][0] + ptt[iw][1] * X + ptt[iw][2] * Y
        ri, rg = o.image(wfs[iw], ropd)
        e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
        e_int = rel_l2(ints, ri)
        e_g = np.abs(gg - rg).max() * RAD2MAS
        print("wfs=%d opd=%.2e ints=%.2e grad=%.3e mas" % (iw, e_opd, e_int, e_g))
        assert e_opd < 1e-5 and e_int <= INTS_TOL and e_g <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("case", ["lgs_small", "ngs_small"])
def test_geometric_ztilt(case, request):
    """Geometric Zernike-tilt gradients (set_geom(1)) against the oracle's
    pts_ztilt of the oracle's OPD; a tip/tilt input moves them by exactly
    that tilt."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: ztilt %s" % case, flush=True)
    eng, wfs = _engine(cfg, p, layers)
    eng.set_geom(1)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    isim = 33
    eng.queue(isim, dt)
    g0 = []
    for iw in range(len(wfs)):
        gg, _ = eng.sync(iw, want_ints=False)
        _, _, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
        rg = o.ztilt(ropd)
        vign = np.concatenate([p.saa < p.cfg.saat] * 2)
        e_g = np.abs(gg - rg)[~vign].max() * RAD2MAS
        print("wfs=%d ztilt err %.3e mas (|g| max %.1f mas)" % (iw, e_g, np.abs(rg).max() * RAD2MAS))
        assert e_g < 1e-3
        g0.append(gg)
    tilt = np.array([0.0, 2e-6, -1e-6])
    eng.set_ptt(tilt)
    eng.queue(isim, dt)
    for iw in range(len(wfs)):
        gg, _ = eng.sync(iw, want_ints=False)
        d = gg - g0[iw]
        ok = ~(p.saa < p.cfg.saat)
        assert np.allclose(d[:p.nsa][ok], 2e-6, atol=2e-6 * 1e-4)
        assert np.allclose(d[p.nsa:][ok], -1e-6, atol=1e-6 * 1e-4)
    eng.close()


@pytest.mark.parametrize("case", ["lgs_small", "ngs_small"])
def test_geometric_gtilt(case, request):
    """Geometric average-gradient gradients (set_geom(2), GS0 from mkg)
    against the oracle's cell-by-cell evaluation on the GPU's own traced OPD
    (FP32) and on the oracle's OPD; a tip/tilt input moves every
    illuminated subaperture by exactly that tilt."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: gtilt %s nsa=%d" % (case, p.nsa), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    rowp, coli, val = p.gs0()
    eng.set_gs0(rowp, coli, val)
    eng.set_geom(2)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    isim = 21
    eng.queue(isim, dt)
    g0 = []
    for iw in range(len(wfs)):
        gg, _ = eng.sync(iw, want_ints=False)
        own = o.gtilt(eng.debug_opd(iw))
        _, _, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
        rg = o.gtilt(ropd)
        scale = np.abs(rg).max()
        e_own = np.abs(gg - own).max() / scale
        e_ref = np.abs(gg - rg).max() * RAD2MAS
        print("wfs=%d gtilt vs own OPD %.2e rel, vs oracle OPD %.3e mas (|g| max %.1f mas)"
              % (iw, e_own, e_ref, scale * RAD2MAS))
        assert e_own < 1e-6
        assert e_ref < 1e-3
        g0.append(gg)
    tilt = np.array([0.0, 2e-6, -1e-6])
    eng.set_ptt(tilt)
    eng.queue(isim, dt)
    full = np.abs(val).sum() > 0
    for iw in range(len(wfs)):
        gg, _ = eng.sync(iw, want_ints=False)
        d = gg - g0[iw]
        lit = np.diff(rowp) > 0
        assert full
        assert np.allclose(d[lit[:p.nsa].nonzero()[0]], 2e-6, atol=2e-6 * 1e-4)
        assert np.allclose(d[p.nsa + lit[p.nsa:].nonzero()[0]], -1e-6, atol=1e-6 * 1e-4)
    eng.close()


def _dm_maps(rs):
    """two DM surfaces: 30 m at 1/64 m on the ground and a 3.5 m map at 1/32 m
    conjugated to 8 km (smaller than the footprints there, so that some
    points fall outside and are left alone)"""
    out = []
    for n, dx, ht in ((1920, 1.0 / 64, 0.0), (112, 1.0 / 32, 8000.0)):
        y, x = np.mgrid[0:n, 0:n] * dx
        p = 1e-7 * (np.sin(2 * np.pi * x / 3.1 + rs.uniform(0, 6)) * np.cos(2 * np.pi * y / 2.3)
                    + 0.3 * rs.standard_normal((n, n)))
        out.append((p, dx, -(n // 2) * dx, -(n // 2) * dx, ht))
    return out


@pytest.mark.parametrize("case", ["ngs_two_wvl", "lgs_small"])
def test_dm_surfaces(case, request):
    """DM surfaces traced onto the WFS (no wrap, alpha = -1): the OPD equals
    the oracle's traced OPD minus the oracle's non-wrapping prop of every DM,
    and the images and gradients follow from it."""
    import ctypes as C
    import oracle
    from maos_amd import engine
    cfg, p, layers = request.getfixturevalue(case)
    print("case: DM surfaces %s" % case, flush=True)
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    if p.llt_nx:
        mtche, i0sum = _lgs_ref(p, o, wfs, dt)
        o.set_ref(mtche, i0sum)
        eng.update_ref(mtche, i0sum)
    eng.set_variant(VARIANTS["fft32"])
    dms = _dm_maps(np.random.RandomState(4))
    engine.dm2gpu(dms)
    try:
        isim = 12
        eng.queue(isim, dt)
        L = oracle.lib()
        n = p.nxsa
        for iw in range(len(wfs)):
            gg, ints = eng.sync(iw)
            opd = eng.debug_opd(iw)
            _, _, ropd, rlotf = o.step(wfs[iw], isim, dt, want_opd=True, want_lotf=True)
            missing = 0
            for (m, dx, ox, oy, ht) in dms:
                m = np.ascontiguousarray(m)
                missing += L.orc_prop(m.ctypes.data_as(oracle.dp), m.shape[1], m.shape[0], dx, ox, oy, ht,
                                      np.ascontiguousarray(p.ptsx).ctypes.data_as(oracle.dp),
                                      np.ascontiguousarray(p.ptsy).ctypes.data_as(oracle.dp), p.nsa, n, p.dx,
                                      ropd.ctypes.data_as(oracle.dp), -1.0, wfs[iw].hs, wfs[iw].thetax,
                                      wfs[iw].thetay, wfs[iw].misregx, wfs[iw].misregy, 0.0, 0.0, 0)
            ri, rg = o.image(wfs[iw], ropd, rlotf)
            e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
            e_int = rel_l2(ints, ri)
            e_g = np.abs(gg - rg).max() * RAD2MAS
            print("wfs=%d points outside the DM maps %d, opd=%.2e ints=%.2e grad=%.3e mas"
                  % (iw, missing, e_opd, e_int, e_g))
            assert missing > 0  # the edge rule is exercised
            assert e_opd < 1e-5 and e_int <= INTS_TOL and e_g <= GRAD_TOL_MAS
    finally:
        engine.dm2gpu([])
    eng.close()


@pytest.mark.parametrize("case", ["lgs_small", "ngs_two_wvl", "lgs_pad128"])
def test_raytrace_two_row_form(case, request, monkeypatch):
    """The ray trace carries each lane's upper x-interpolated tile row to the
    next output row; a layer whose row step passes over a tile row takes the
    plain two-row interpolation instead.  MX_RT_TWOROW=1 forces that form on
    every layer: both forms do the same arithmetic on the same values, so
    the traced OPD is bit-identical, and it matches the oracle's."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: %s nsa=%d nxsa=%d, both interpolation forms" % (case, p.nsa, p.nxsa), flush=True)
    from maos_amd.powfs import PTYPE_MF
    eng, wfs = _engine(cfg, p, layers)
    if cfg["powfs"].phytype == PTYPE_MF:  # the step needs a filter; the OPD does not use it
        eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    for isim in (0, 77):
        monkeypatch.delenv("MX_RT_TWOROW", raising=False)
        eng.queue(isim, dt)
        eng.wait()
        carried = [eng.debug_opd(iw) for iw in range(len(wfs))]
        monkeypatch.setenv("MX_RT_TWOROW", "1")
        eng.queue(isim, dt)
        eng.wait()
        for iw in range(len(wfs)):
            two = eng.debug_opd(iw)
            _, _, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
            e_opd = np.abs(two - ropd).max() / np.abs(ropd).max()
            print("isim=%d wfs=%d forms equal=%s opd vs oracle %.2e"
                  % (isim, iw, np.array_equal(two, carried[iw]), e_opd))
            assert np.array_equal(two, carried[iw])
            assert e_opd < 1e-5
    monkeypatch.delenv("MX_RT_TWOROW", raising=False)
    eng.close()


def test_screens_smaller_than_tile(ngs_two_wvl):
    """Screens of 32 x 32 pixels (narrower than the staged 34-pixel tile of a
    32-point grid, so a tile would wrap more than once) are traced by the
    direct gather; the OPD matches the oracle's periodic trace."""
    import dataclasses
    from maos_amd.powfs import genatm
    cfg, p, _ = ngs_two_wvl
    atm = dataclasses.replace(cfg["atm"], nx=32)
    layers = genatm(atm)
    print("case: ngs_two_wvl geometry, %d layers of %s" % (len(layers), layers[0].p.shape), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    for isim in (0, 31):
        eng.queue(isim, dt)
        eng.wait()
        for iw in range(len(wfs)):
            opd = eng.debug_opd(iw)
            _, _, ropd, _ = o.step(wfs[iw], isim, dt, want_opd=True)
            e_opd = np.abs(opd - ropd).max() / np.abs(ropd).max()
            print("isim=%d wfs=%d opd vs oracle %.2e" % (isim, iw, e_opd))
            assert e_opd < 1e-5
    eng.close()


def test_dither_statistics_and_gpu_matched_filter(lgs_small):
    """TT-dither statistics and the on-GPU matched-filter update:
    - each step adds ints, cs ints and ss ints (FP64) per WFS; the sums
      equal the host's FP64 sums of the synced images;
    - a fold moves s0 imb, sg imx, sg imy into the folded sums and clears
      the step sums;
    - the update blends keep*model + acc*folded and rebuilds mtche, i0sum
      and NEA^2 on the GPU, equal to the host mtch of the same blended
      model (mxs_mtch, FP64) to FP32 rounding of the filter;
    - the next step's gradients with that filter equal those with the host
      filter uploaded by mx_wfsgrad_update_ref."""
    cfg, p, layers = lgs_small
    print("case: lgs order20 nsa=%d nwfs=%d, dither + GPU MF update" % (p.nsa, len(cfg["wfs"])), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    nw = len(wfs)
    i0m, gxm, gym = p.i0_model(cfg["atm"].r0, cfg["atm"].L0)
    mt0, is0, _ = p.mtch(i0m, gxm, gym)
    eng.update_ref(mt0, is0)
    eng.dither_control(True, True)
    for iw in range(nw):
        eng.set_i0model(iw, i0m, gxm, gym)
    dt = cfg["sim"].dt
    rng = np.random.default_rng(5)
    n = p.nsa * p.npix
    imb = np.zeros((nw, n))
    imx = np.zeros((nw, n))
    imy = np.zeros((nw, n))
    for isim in range(3):
        cs = rng.uniform(-1, 1, nw)
        ss = rng.uniform(-1, 1, nw)
        for iw in range(nw):
            eng.set_dither(cs[iw], ss[iw], iwfs=iw)
        eng.queue(isim, dt)
        for iw in range(nw):
            _, ints = eng.sync(iw)
            x = ints.reshape(-1)
            imb[iw] += x
            imx[iw] += cs[iw] * x
            imy[iw] += ss[iw] * x
    for iw in range(nw):
        b, gx_, gy_, nf = eng.dither_get(iw)
        assert nf == 3
        np.testing.assert_allclose(b.reshape(-1), imb[iw], rtol=1e-14)
        np.testing.assert_allclose(gx_.reshape(-1), imx[iw], rtol=1e-13, atol=1e-12 * np.abs(imx[iw]).max())
        np.testing.assert_allclose(gy_.reshape(-1), imy[iw], rtol=1e-13, atol=1e-12 * np.abs(imy[iw]).max())
    s0, sg = 1.0 / 3, 2.0 / (3 * 0.05)
    eng.dither_fold(s0, sg)
    b, gx_, gy_, _ = eng.dither_get(0)
    assert not b.any() and not gx_.any() and not gy_.any()
    keep, acc = 0.6, 0.4
    eng.mtch_update(keep, acc)
    for iw in range(nw):
        m, i0sum, sanea, i0, gx, gy = eng.mtch_get(iw)
        want_i0 = keep * i0m.reshape(-1) + acc * s0 * imb[iw]
        want_gx = keep * gxm.reshape(-1) + acc * sg * imx[iw]
        want_gy = keep * gym.reshape(-1) + acc * sg * imy[iw]
        np.testing.assert_allclose(i0.reshape(-1), want_i0, rtol=1e-12, atol=1e-12 * np.abs(want_i0).max())
        np.testing.assert_allclose(gx.reshape(-1), want_gx, rtol=1e-12, atol=1e-12 * np.abs(want_gx).max())
        np.testing.assert_allclose(gy.reshape(-1), want_gy, rtol=1e-12, atol=1e-12 * np.abs(want_gy).max())
        hm, hs, hsanea = p.mtch(i0, gx, gy)
        hm = hm.reshape(p.nsa, 2 * p.npix)
        err = np.abs(m - hm).max() / np.abs(hm).max()
        print("wfs=%d mtche GPU vs host max rel %.2e, zeroed %d/%d" % (iw, err, int((~m.any(1)).sum()), p.nsa))
        assert err < 1e-6
        np.testing.assert_array_equal(~m.any(1), ~hm.any(1))
        np.testing.assert_allclose(i0sum, hs, rtol=1e-12)
        np.testing.assert_allclose(sanea, hsanea, rtol=1e-7, atol=1e-9 * np.abs(hsanea).max())
    # the next step with the GPU filter equals the host filter uploaded
    eng.queue(3, dt)
    g_gpu = [eng.sync(iw, want_ints=False)[0] for iw in range(nw)]
    for iw in range(nw):
        _, _, _, i0, gx, gy = eng.mtch_get(iw)
        hm, hs, _ = p.mtch(i0, gx, gy)
        eng.update_ref(hm, hs, iwfs=iw)
    eng.queue(3, dt)
    for iw in range(nw):
        g_host = eng.sync(iw, want_ints=False)[0]
        e = np.abs(g_gpu[iw] - g_host).max() * RAD2MAS
        print("wfs=%d gradients GPU-filter vs host-filter %.2e mas" % (iw, e))
        assert e < 1e-3
    eng.close()


@pytest.mark.parametrize("case,variant", [("lgs_small", "mfma_bf16"), ("lgs_small", "fft32"),
                                          ("ngs_two_wvl", "fft32"), ("lf_mixed", "fft32")])
def test_sigmatch_global(case, variant, request):
    """powfs.sigmatch = 2 (flux normalised over the whole WFS): the matched
    filter's gradients scaled by i0sumsum / sum(ints), the CoG's moments
    divided by sum(ints)/saasum * saa (imaging epilogues of every form leave
    the raw moments, one pass per WFS finishes them); parity with the
    oracle's shwfs_grad."""
    import dataclasses
    from maos_amd.powfs import PTYPE_MF, Powfs
    cfg0, _, layers = request.getfixturevalue(case)
    pc = dataclasses.replace(cfg0["powfs"], sigmatch=2)
    cfg = dict(cfg0, powfs=pc)
    p = Powfs(cfg["aper"], pc, cfg["sim"])
    print("case: %s sigmatch=2 %s phytype=%d nsa=%d" % (case, variant, pc.phytype, p.nsa), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    if not eng.variant_ok(VARIANTS[variant]):
        pytest.skip("%s does not apply to %s" % (variant, case))
    eng.set_variant(VARIANTS[variant])
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    if pc.phytype == PTYPE_MF:
        mtche, i0sum = _lgs_ref(p, o, wfs, dt)
        o.set_ref(mtche, i0sum)
        eng.update_ref(mtche, i0sum)
    for isim in (0, 41):
        eng.queue(isim, dt)
        for iw in range(len(wfs)):
            g, ints = eng.sync(iw)
            ri, rg, _, _ = o.step(wfs[iw], isim, dt)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d ints=%.2e grad=%.3e mas" % (isim, iw, e_int, e_g))
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16", "mfma_fp16"])
def test_two_sodium_profiles(lgs_small, variant):
    """Two sodium-profile ETFs blended per step (w1 etf + w2 etf2,
    maos/wfsints.c:211-220): the weight changes between steps through
    mx_wfsgrad_set_etf_weight; every form matches the oracle's blend."""
    import ctypes as C
    import dataclasses
    from maos_amd.powfs import Powfs, SodiumCfg
    cfg, p, layers = lgs_small
    pc2 = dataclasses.replace(cfg["powfs"], sodium=SodiumCfg(center=92e3, fwhm=6e3, hmin=70e3, hmax=110e3,
                                                             dh=20.0))
    p2 = Powfs(cfg["aper"], pc2, cfg["sim"])
    print("case: lgs order20 nsa=%d, two sodium profiles, %s" % (p.nsa, variant), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    o.set_ref(mtche, i0sum)
    eng.update_ref(mtche, i0sum)
    etf2 = np.ascontiguousarray(p2.etf.reshape(-1))
    o.keep.append(etf2)
    o.s.etf2 = etf2.ctypes.data_as(C.POINTER(C.c_double))
    for iw in range(p.nwvl):
        eng.update_etf(iw, p.etf[iw], p2.etf[iw], 0.3)
    eng.set_variant(VARIANTS[variant])
    assert eng.variant() == VARIANTS[variant]
    for isim, w2 in ((0, 0.3), (17, 0.75)):
        eng.set_etf_weight(w2)
        o.s.etf2wt = w2
        eng.queue(isim, dt)
        for iw in range(len(wfs)):
            g, ints = eng.sync(iw)
            ri, rg, _, _ = o.step(wfs[iw], isim, dt)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d w2=%.2f wfs=%d ints=%.2e grad=%.3e mas" % (isim, w2, iw, e_int, e_g))
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


def test_errors_reported():
    """Invalid calls return error codes with a message and launch nothing."""
    import ctypes as C
    from maos_amd import _lib, engine
    engine.init(0)
    L = _lib.hip()
    h = C.c_void_p()
    assert L.mx_wfsgrad_init(C.byref(h), None, None, 0) == -1
    assert b"null" in L.mx_last_error()
    cfgp = _lib.MxPowfs(nsa=10, nxsa=24, embfac=2, notf=48, pixpsax=6, pixpsay=6, nwvl=1)
    w = (_lib.MxWfs * 1)()
    assert L.mx_wfsgrad_init(C.byref(h), C.byref(cfgp), w, 1) == -1
    assert b"unsupported" in L.mx_last_error()
    assert L.mx_atm2gpu(None, 17, 1.0) == -1


@pytest.mark.parametrize("case,variant", [("lgs_small", "mfma_bf16"), ("lgs_small", "fft32"),
                                          ("lgs_small", "fft32_reg"), ("lf_mixed", "fft32"),
                                          ("ngs_small", "fft32")])
def test_dtrat_exposure(case, variant, request, monkeypatch):
    """dtrat = 3 (maos/wfsgrad.c:164-167, 411-495): an exposure's
    intensities are the sum of its three frames and its gradients those of
    the summed image (the oracle's shwfs_grad of the sum); the steps inside
    an exposure leave intensities and gradients as they were; the FP64 i0
    sum takes whole exposures and counts them.  Run as one HIP graph per
    step, so the exposure phase travels in the nodes' arguments."""
    cfg, p, layers = request.getfixturevalue(case)
    if variant == "fft32_reg":
        monkeypatch.setenv("MX_FFT_REG", "1")
    print("case: dtrat 3, %s, %s" % (case, variant), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    if not eng.variant_ok(VARIANTS[variant]):
        pytest.skip("%s does not apply to %s" % (variant, case))
    eng.set_variant(VARIANTS[variant])
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    if p.cfg.phytype == 1:  # matched filter
        mtche, i0sum = _lgs_ref(p, o, wfs, dt)
        o.set_ref(mtche, i0sum)
        eng.update_ref(mtche, i0sum)
    eng.set_graph(True)
    eng.set_dtrat(3)
    eng.i0_control(enable=True, reset=True)
    iw = len(wfs) - 1
    frames = {isim: o.step(wfs[iw], isim, dt)[0] for isim in range(6)}
    for isim in range(3):
        eng.queue(isim, dt)
    g0, ints0 = eng.sync(iw)
    want0 = frames[0] + frames[1] + frames[2]
    e_int = rel_l2(ints0, want0)
    e_g = np.abs(g0 - o.grad(want0)).max() * RAD2MAS
    print("exposure 0-2: ints=%.2e grad=%.3e mas" % (e_int, e_g), flush=True)
    assert e_int <= INTS_TOL and e_g <= GRAD_TOL_MAS
    for isim in (3, 4):
        eng.queue(isim, dt)
        g, ints = eng.sync(iw)
        assert np.array_equal(g, g0) and np.array_equal(ints, ints0)  # inside an exposure
    eng.queue(5, dt)
    g1, ints1 = eng.sync(iw)
    want1 = frames[3] + frames[4] + frames[5]
    e_int = rel_l2(ints1, want1)
    e_g = np.abs(g1 - o.grad(want1)).max() * RAD2MAS
    print("exposure 3-5: ints=%.2e grad=%.3e mas" % (e_int, e_g), flush=True)
    assert e_int <= INTS_TOL and e_g <= GRAD_TOL_MAS
    i0, nframe = eng.i0_get(iw)
    assert nframe == 2
    assert np.abs(i0 - (ints0 + ints1)).max() <= 1e-6 * np.abs(i0).max()
    eng.set_dtrat(1)
    eng.queue(6, dt)
    g, ints = eng.sync(iw)
    ri, rg, _, _ = o.step(wfs[iw], 6, dt)
    assert rel_l2(ints, ri) <= INTS_TOL and np.abs(g - rg).max() * RAD2MAS <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("case,gtype", [("lgs_small", 1), ("lgs_small", 2), ("ngs_two_wvl", 1)])
def test_pistat_matches_oracle(case, gtype, request):
    """Tip/tilt-removed PSF statistics (pistatout) over three steps against
    the oracle's restatement of MAOS's OTF-tilt path, fed the GPU's own
    traced OPD and the oracle's geometric gradients of it (Zernike tilt or
    G-tilt); the imaging step itself is unchanged by the extra kernels.
    ngs_two_wvl has two wavelengths and no LLT."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: pistat %s gtype=%d nsa=%d" % (case, gtype, p.nsa), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    o = _oracle(p, layers)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    if gtype == 2:
        eng.set_gs0(*p.gs0())
    dt = cfg["sim"].dt
    eng.queue(4, dt)
    g_off, i_off = eng.sync(0)
    eng.pistat_control(True, gtype=gtype, reset=True)
    acc = {iw: None for iw in sorted({0, len(wfs) - 1})}
    for isim in (4, 5, 6):
        eng.queue(isim, dt)
        for iw in acc:
            opd = eng.debug_opd(iw)
            g = o.ztilt(opd) if gtype == 1 else o.gtilt(opd)
            acc[iw] = o.pistat(opd, g, acc[iw])
        if isim == 4:
            g_on, i_on = eng.sync(0)
            assert np.array_equal(g_on, g_off) and np.array_equal(i_on, i_off)
    for iw, ref in acc.items():
        got, n = eng.pistat_get(iw)
        assert n == 3
        e = rel_l2(got, ref)
        c = np.unravel_index(np.argmax(got[0, np.argmax(p.saa)]), got.shape[2:])
        print("wfs=%d pistat rel L2 %.2e, peak of a full subaperture at %s" % (iw, e, c))
        assert e < INTS_TOL
    eng.pistat_control(False)
    eng.close()


@pytest.mark.parametrize("case", ["lgs_small", "ngs_small", "lgs_pad128"])
def test_psfout_matches_oracle(case, request):
    """The complex field output (psfout) of one step against the oracle's
    restatement of MAOS's transform pair, on the GPU's own traced OPD
    (FP32 phases: 1e-5 of the largest value)."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: psfout %s nsa=%d notf=%d" % (case, p.nsa, p.notf), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    o = _oracle(p, layers)
    eng.queue(7, cfg["sim"].dt)
    for iw in sorted({0, len(wfs) - 1}):
        got = eng.psfout(iw)
        ref = o.psfout(eng.debug_opd(iw))
        e = np.abs(got - ref).max() / np.abs(ref).max()
        print("wfs=%d psfout max error %.2e of the largest value" % (iw, e))
        assert got.shape == ref.shape and e < 1e-5
    eng.close()


@pytest.mark.parametrize("case,offsets,variant", [
    ("ngs_small", "random", "fft32"),          # N = 32: the register form
    ("ngs_two_wvl", "random", "fft32"),        # more than 16 sampled columns: the register form
    ("ngs_two_wvl", "samples", "mfma_bf16"),   # whole PSF samples: the union fits the MFMA form
    ("ngs_two_wvl", "samples", "fft32"),       # ... and the FP32 radix-8 form
    ("lgs_small", "random", "fft32"),          # with an LLT: si per subaperture, nominal in the ETF
    ("lf_mixed", "random", "fft32"),           # the large-field form (96^2 field in a 120^2 OTF)
])
def test_per_subaperture_tables(case, offsets, variant, request):
    """multi_nominal: an si table (and a nominal without an LLT) per
    subaperture, from per-subaperture pixel offsets and rotations
    (mkdtf's loop, lib/mkdtf.c:140-161), against the oracle fed the same
    tables."""
    cfg, p, layers = request.getfixturevalue(case)
    print("case: per-subaperture tables %s %s %s" % (case, offsets, variant), flush=True)
    rs = np.random.RandomState(17)
    if offsets == "random":
        ox, oy = rs.uniform(-0.4, 0.4, p.nsa), rs.uniform(-0.4, 0.4, p.nsa)
        rot = rs.uniform(-0.1, 0.1, p.nsa) if p.llt_nx == 0 else np.zeros(p.nsa)
    else:
        step = p.dtheta[0] / p.pixthetax  # one PSF sample, in pixels
        ox, oy = step * rs.randint(-1, 2, p.nsa), step * rs.randint(-1, 2, p.nsa)
        rot = np.zeros(p.nsa)
    tabs = p.dtf_sa(ox, oy, rot)
    eng, wfs = _engine(cfg, p, layers)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    for iw in range(p.nwvl):
        eng.set_dtf_sa(iw, tabs[0][iw], tabs[1][iw], tabs[2][iw], tabs[3][iw])
    eng.set_variant(VARIANTS[variant])
    o = _oracle(p, layers, np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    o.set_dtf_sa(*tabs)
    dt = cfg["sim"].dt
    eng.queue(9, dt)
    for iwfs in sorted({0, len(wfs) - 1}):
        g, ints = eng.sync(iwfs)
        ri, rg, _, _ = o.step(wfs[iwfs], 9, dt)
        e_int = rel_l2(ints, ri)
        e_g = np.abs(g - rg).max() * RAD2MAS
        print("wfs=%d variant=%d ints=%.2e grad=%.3e mas" % (iwfs, eng.variant(), e_int, e_g))
        assert e_int <= INTS_TOL
        assert e_g <= GRAD_TOL_MAS
    eng.close()


@pytest.mark.parametrize("variant", ["fft32", "mfma_bf16", "mfma_fp16", "fft32_reg"])
def test_wide_llt_parity(lgs_wide_llt, variant, monkeypatch):
    """A 0.6 m launch telescope on 0.5 m subapertures: the 38^2 LLT grid is
    traced in two-by-two tiles, and its uplink OTF (the 76^2 transform
    cropped to 64^2, maos/wfsints.c:160-185) is summed directly as the
    field's autocorrelation; every imaging form against the oracle."""
    cfg, p, layers = lgs_wide_llt
    if variant == "fft32_reg":
        monkeypatch.setenv("MX_FFT_REG", "1")
    print("case: wide LLT nlx=%d nxsa=%d nsa=%d, %s" % (p.llt_nx, p.nxsa, p.nsa, variant), flush=True)
    eng, wfs = _engine(cfg, p, layers)
    eng.set_variant(VARIANTS[variant])
    o = _oracle(p, layers)
    dt = cfg["sim"].dt
    mtche, i0sum = _lgs_ref(p, o, wfs, dt)
    o.set_ref(mtche, i0sum)
    eng.update_ref(mtche, i0sum)
    for isim in (0, 77):
        eng.queue(isim, dt)
        for iwfs in sorted({0, len(wfs) - 1}):
            g, ints = eng.sync(iwfs)
            ri, rg, _, rlotf = o.step(wfs[iwfs], isim, dt, want_lotf=True)
            lotf = eng.debug_lotf(iwfs)
            e_lotf = rel_l2(lotf, rlotf)
            e_int = rel_l2(ints, ri)
            e_g = np.abs(g - rg).max() * RAD2MAS
            print("isim=%d wfs=%d lotf=%.2e ints=%.2e grad=%.3e mas" % (isim, iwfs, e_lotf, e_int, e_g))
            assert e_lotf < 1e-5
            assert e_int <= INTS_TOL
            assert e_g <= GRAD_TOL_MAS
    eng.close()


def test_uplink_computed_ahead(lgs_small, monkeypatch):
    """With MX_UPLINK_PREFETCH=1 the uplink OTF of step isim + 1 is
    computed beside step isim's chain on a second stream (it needs only the
    screens).  Consecutive steps, a jump in isim, a change of imaging form
    and new screens give bit-identical intensities and uplink OTFs to an
    engine that computes every uplink in its step's chain (the default)."""
    from maos_amd import _lib, engine
    from maos_amd.powfs import genatm
    import dataclasses
    cfg, p, layers = lgs_small
    print("case: uplink computed ahead, lgs order 20", flush=True)
    monkeypatch.setenv("MX_UPLINK_PREFETCH", "1")
    eng, wfs = _engine(cfg, p, layers)
    monkeypatch.delenv("MX_UPLINK_PREFETCH")
    ref = engine.WfsGrad(p, wfs)
    for e in (eng, ref):
        e.set_graph(True)
        e.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    dt = cfg["sim"].dt
    layers2 = genatm(dataclasses.replace(cfg["atm"], seeds=[s + 100 for s in cfg["atm"].seeds]))
    for isim, what in [(10, ""), (11, ""), (12, ""), (40, "jump"), (41, "fft32"), (42, ""), (43, "screens"),
                       (44, "")]:
        if what == "fft32":
            for e in (eng, ref):
                e.set_variant(_lib.VARIANT_FFT32)
        if what == "screens":
            engine.atm2gpu(layers2)
        eng.queue(isim, dt)
        ref.queue(isim, dt)
        for iw in (0, len(wfs) - 1):
            _, a = eng.sync(iw)
            _, b = ref.sync(iw)
            assert np.array_equal(a, b), (isim, iw)
            assert np.array_equal(eng.debug_lotf(iw), ref.debug_lotf(iw)), (isim, iw)
        print("isim=%d %s: equal" % (isim, what))
    ref.close()
    eng.close()
