"""WFS frames/s of the 30 m 6x60x60 LGS step on one MI355X.

A step is one time step of all six LGS WFS: ray tracing of seven 4096^2
screens, LLT uplink OTF, per-subaperture imaging (embed, FFT, |E|^2, FFT,
x conj(uplink OTF) x ETF, inverse FFT, pixel sampling), matched-filter
gradients and FP64 i0 accumulation.  Screens and tables are resident in HBM
before the timed region.  One frame is one WFS x one step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
                    [--variant auto|fft32|mfma] [--copy-calib BYTES]
                    [--nreal N] [--hetero] [--evl] [--pywfs [--pywfs-parity]]
                    [--no-graph]

Prints one JSON line.  The CPU baseline is the FP64 oracle (a C restatement
of MAOS's CPU path) timed on this host on a bounded sample of frames.

The imaging kernel has two forms: the FP32 radix-FFT form (VALU) and the
split-BF16 MFMA form (the default where it applies).  The line reports the
selected form as `value` and the other form's frames/s beside it.
`--copy-calib BYTES` adds a plain device copy of that size after the timed
region, so that a hardware-counter run of this script can scale the
kernel's FETCH_SIZE / WRITE_SIZE against a copy of known traffic.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("WFS frames/s, 30 m 6×60×60 LGS (MFMA split-BF16 vs FP32 vs FP64), "
          "% roofline")
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_shwfs_mfma_lgs30m_traffic.txt")
EVL_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_psf_cols_evl30m_traffic.txt")
KERNELS = {0: "k_shwfs_fft", 1: "k_shwfs_mfma<false, __bf16>", 2: "k_shwfs_mfma<false, _Float16>"}
FORMS = {0: "fp32-radix8-fft", 1: "mfma-split-bf16", 2: "mfma-split-fp16"}


def flops_per_subap(N, nxsa, nps, npix, llt=True, etf=True):
    """Algorithmic FLOPs of one subaperture-wavelength (5 M log2 M per
    complex transform of M = N^2 points)."""
    M = N * N
    f = 3 * 5 * M * math.log2(M)
    f += 3 * M                       # |E|^2
    f += 6 * M if llt else 0         # x conj(lotfc)
    f += 6 * M if etf else 2 * M     # x ETF (complex) or nominal (real)
    f += 8 * nps * nxsa * nxsa       # ray trace
    f += 3 * nxsa * nxsa             # embed (2 mul + 1 sincos)
    f += 8 * npix + 4 * npix         # sampling + matched filter
    return f


def mfma_flops_per_subap():
    """BF16 MFMA FLOPs the MFMA imaging kernel executes per subaperture-
    wavelength at N=64, nxsa=32: 16x16x32 products, three per split multiply,
    four waves.  Forward transform over the 32-wide band, 96 (first axis) +
    192 (second axis, |E|^2); half-plane inverse transform with radix-2
    halved sums, 48 + 96; detector rows and columns, 48 + 12."""
    n32 = 0
    n16 = 96 + 192 + 48 + 96 + 48 + 12
    return n32 * 32 * 32 * 16 * 2 + n16 * 16 * 16 * 32 * 2


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel recorded by the counter runs
    (FETCH_SIZE + WRITE_SIZE scaled by a plain copy), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def time_variant(eng, variant, isim, dt, nsteps):
    """frames/s of one imaging form, and the next isim"""
    eng.set_variant(variant)
    for _ in range(5):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    t1 = time.perf_counter()
    for _ in range(nsteps):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    return nsteps / (time.perf_counter() - t1), isim


def copy_calibration(nbytes, reps=5):
    """A plain device-to-device copy of known size through the HIP runtime
    (its blit kernel), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        ok = ok and hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # device to device
    ok = ok and hip.hipDeviceSynchronize() == 0
    for p in (src, dst):
        if p.value:
            hip.hipFree(p)
    if not ok:
        raise RuntimeError("copy calibration failed")
    log("copy calibration: %d x %d bytes" % (reps, nbytes))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch_main(args):
    """nreal realisations of the lgs30m step in one engine: aggregate WFS
    frames/s, per-step latency p50/p99, and the error against the FP64 oracle
    on two realisations (the first and the last)."""
    import dataclasses
    from maos_amd import configs, engine
    from maos_amd.powfs import Powfs, genatm

    t0 = time.time()
    cfg = configs.lgs30m(nx=args.nx, order=args.order)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    wfs = configs.wfs_specs(cfg, p)
    nw, nreal = len(wfs), args.nreal
    dt = cfg["sim"].dt
    nps = len(cfg["atm"].ht)
    r0s = np.random.RandomState(2025).uniform(0.12, 0.25, nreal)
    engine.init(0)
    kept = {}
    for k in range(nreal):
        # realisation k: its own r0 and layer seeds 7k+1 .. 7k+7
        atm = dataclasses.replace(cfg["atm"], r0=float(r0s[k]), seeds=[nps * k + j + 1 for j in range(nps)])
        layers = genatm(atm)
        engine.atm2gpu(layers, ireal=k)
        if k in (0, nreal - 1):
            kept[k] = layers
        log("realisation %d: r0=%.3f m (%.1fs)" % (k, r0s[k], time.time() - t0))
    eng = engine.WfsGrad(p, wfs, nreal=nreal)
    if not args.no_graph:
        eng.set_graph(True)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    eng.i0_control(enable=True, reset=True)
    for isim in range(args.i0_steps):
        eng.queue(isim, dt)
    eng.wait()
    refs = {}
    for iw in range(eng.nwfs):
        i0, n = eng.i0_get(iw)
        mtche, i0sum, _ = p.mtch(i0 / n)
        eng.update_ref(mtche, i0sum, iwfs=iw)
        refs[iw] = (mtche, i0sum)
    eng.i0_control(enable=True, reset=True)
    log("matched filters of %d WFS (%.1fs)" % (eng.nwfs, time.time() - t0))
    isim = args.i0_steps
    for _ in range(args.warmup):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    sec = time.perf_counter() - t1
    fps = eng.nwfs * args.steps / sec
    # latency of one step on an idle engine
    lat = []
    for _ in range(args.latency_steps):
        ta = time.perf_counter()
        eng.queue(isim, dt)
        eng.wait()
        lat.append((time.perf_counter() - ta) * 1e3)
        isim += 1
    lat = np.array(lat)
    eng.timing_enable(True)
    for _ in range(20):
        eng.queue(isim, dt)
        isim += 1
    nl, ms_llt, ms_opd, ms_ints = eng.timing()
    eng.timing_enable(False)
    out = {"metric": METRIC, "value": round(fps, 2), "unit": "WFS frames/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
           "data": "synthetic (%d seeded von Karman atmospheres, Gaussian sodium profile)" % nreal,
           "config": {"workload": "lgs30m x %d realisations in one engine (r0 ~ U[0.12, 0.25] m, layer "
                                  "seeds 7k+1..7k+7), %d WFS x 2820 subapertures per step" % (nreal, eng.nwfs),
                      "nreal": nreal, "nwfs": eng.nwfs, "nsa": p.nsa,
                      "launch": "kernel by kernel" if args.no_graph else "one HIP graph per step"},
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3), "steps": len(lat)},
           "kernels_us": {"raytrace": round(ms_opd / nl * 1e3, 1), "uplink": round(ms_llt / nl * 1e3, 1),
                          "imaging": round(ms_ints / nl * 1e3, 1)}}
    # error against the FP64 oracle on the first and the last realisation
    import oracle
    errs = []
    eng.queue(isim, dt)
    for k, layers in kept.items():
        o = oracle.Oracle(p)
        o.atm(layers)
        iw = k * nw
        o.set_ref(*refs[iw])
        g, ints = eng.sync(iw)
        ri, rg, _, _ = o.step(wfs[0], isim, dt)
        errs.append({"realisation": k, "wfs": 0, "ints_rel_l2": float(np.linalg.norm(ints - ri) / np.linalg.norm(ri)),
                     "grad_max_abs_mas": float(np.abs(g - rg).max() * 180 / math.pi * 3600e3)})
    out["parity"] = errs
    print(json.dumps(out), flush=True)
    eng.close()


def split_main(args):
    """The six lgs30m WFS as one engine against K engines of 6/K WFS grouped
    into one graph with a branch per engine (mx_wfsgroup_*): µs per step of
    each, alternated twice."""
    from maos_amd import configs, engine
    from maos_amd.powfs import Powfs, genatm

    cfg = configs.lgs30m(nx=args.nx, order=args.order)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    wfs = configs.wfs_specs(cfg, p)
    engine.init(0)
    engine.atm2gpu(genatm(cfg["atm"]))
    dt = cfg["sim"].dt
    k = args.split
    if k not in (2, 3, 6):
        raise SystemExit("--split takes 2, 3 or 6")

    def mk(ws):
        e = engine.WfsGrad(p, ws)
        e.set_graph(True)
        e.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
        return e

    def run(queue, wait):
        isim = 100
        for _ in range(args.warmup):
            queue(isim, dt)
            isim += 1
        wait()
        t = time.perf_counter()
        for _ in range(args.steps):
            queue(isim, dt)
            isim += 1
        wait()
        return (time.perf_counter() - t) / args.steps * 1e6

    res = {"one engine": [], "%d engines" % k: []}
    for _ in range(2):
        e = mk(wfs)
        res["one engine"].append(round(run(e.queue, e.wait), 1))
        e.close()
        n = len(wfs) // k
        engs = [mk(wfs[i * n:(i + 1) * n]) for i in range(k)]
        g = engine.WfsGroup(engs)
        res["%d engines" % k].append(round(run(g.queue, g.wait), 1))
        g.close()
        for x in engs:
            x.close()
    print(json.dumps({"split": k, "us_per_step": res, "steps": args.steps, "warmup": args.warmup}), flush=True)


def hetero_main(args):
    """The heterogeneous step: the six lgs30m LGS WFS, three low-order NGS
    WFS (2x2 of 15 m, 480^2 fields in 512^2 OTFs, two wavelengths) and one
    tip-tilt NGS WFS (1x1 of 30 m, 960^2 in 1024^2) on one atmosphere,
    stepped as one group graph per step.  Each engine is autotuned first;
    per-class kernel times come from each engine stepped alone."""
    from maos_amd import configs, engine
    from maos_amd.powfs import Powfs, genatm

    t0 = time.time()
    cfgs = [configs.lgs30m(nx=args.nx, order=args.order), configs.ngs_lo(nx=args.nx),
            configs.ngs_tt(nx=args.nx)]
    layers = genatm(cfgs[0]["atm"])
    engine.init(0)
    engine.atm2gpu(layers)
    dt = cfgs[0]["sim"].dt
    classes = []
    for cfg in cfgs:
        p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
        wfs = configs.wfs_specs(cfg, p)
        eng = engine.WfsGrad(p, wfs)
        classes.append(dict(name=cfg["name"], cfg=cfg, p=p, wfs=wfs, eng=eng))
        log("%s: nsa=%d nxsa=%d notf=%d nwvl=%d nwfs=%d (%.1fs)" % (cfg["name"], p.nsa, p.nxsa, p.notf,
                                                                   p.nwvl, len(wfs), time.time() - t0))
    # LGS matched filter from the FP64 i0 (as the single-class run)
    lgs = classes[0]
    eng, p = lgs["eng"], lgs["p"]
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    eng.i0_control(enable=True, reset=True)
    for isim in range(args.i0_steps):
        eng.queue(isim, dt)
    eng.wait()
    for iw in range(len(lgs["wfs"])):
        i0, n = eng.i0_get(iw)
        mtche, i0sum, _ = p.mtch(i0 / n)
        eng.update_ref(mtche, i0sum, iwfs=iw)
        if iw == 0:
            lgs["ref"] = (mtche, i0sum)
    isim = args.i0_steps
    for c in classes:
        c["eng"].i0_control(enable=True, reset=True)
        c["tune"] = c["eng"].autotune(isim, dt, nrep=10)
        best = [t for t in c["tune"] if t["best"]][0]
        log("%s autotune: %s" % (c["name"], ", ".join("%s/%d/%d %.1f us%s" % (
            t["form"], t["opd_tile"], t["threads"], t["step_us"], "*" if t["best"] else "") for t in c["tune"])))
        c["choice"] = best
        isim += 12
    # each class alone: kernel times with events
    for c in classes:
        e = c["eng"]
        e.timing_enable(True)
        for _ in range(20):
            e.queue(isim, dt)
            isim += 1
        nl, ms_llt, ms_opd, ms_ints = e.timing()
        e.timing_enable(False)
        c["us"] = {"raytrace": round(ms_opd / nl * 1e3, 1), "uplink": round(ms_llt / nl * 1e3, 1),
                   "imaging": round(ms_ints / nl * 1e3, 1),
                   "step": round((ms_opd + ms_llt + ms_ints) / nl * 1e3, 1)}
    group = engine.WfsGroup([c["eng"] for c in classes])
    nwfs = sum(len(c["wfs"]) for c in classes)
    for _ in range(args.warmup):
        group.queue(isim, dt)
        isim += 1
    group.wait()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        group.queue(isim, dt)
        isim += 1
    group.wait()
    sec = time.perf_counter() - t1
    fps = nwfs * args.steps / sec
    lat = []
    for _ in range(args.latency_steps):
        ta = time.perf_counter()
        group.queue(isim, dt)
        group.wait()
        lat.append((time.perf_counter() - ta) * 1e3)
        isim += 1
    lat = np.array(lat)
    log("group: %d steps in %.3fs -> %.1f frames/s" % (args.steps, sec, fps))
    # parity of WFS 0 of each class against the FP64 oracle at one step
    import oracle
    group.queue(isim, dt)
    group.wait()
    parity = {}
    for c in classes:
        o = oracle.Oracle(c["p"])
        o.atm(layers)
        if "ref" in c:
            o.set_ref(*c["ref"])
        g, ints = c["eng"].sync(0)
        ri, rg, _, _ = o.step(c["wfs"][0], isim, dt)
        parity[c["name"]] = {"ints_rel_l2": float(np.linalg.norm(ints - ri) / np.linalg.norm(ri)),
                             "grad_max_abs_mas": float(np.abs(g - rg).max() * 180 / math.pi * 3600e3)}
    out = {"metric": METRIC, "value": round(fps, 2), "unit": "WFS frames/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3+fp32",
           "data": "synthetic (seeded von Karman screens, Gaussian sodium profile)",
           "config": {"workload": "heterogeneous step: lgs30m (6 LGS x %d subapertures, 64^2) + ngs_lo "
                                  "(3 NGS x 2x2 of 15 m, 480^2 in 512^2, 2 wavelengths) + ngs_tt (1 NGS, "
                                  "960^2 in 1024^2), 7 layers %d^2 shared, one group graph per step"
                                  % (classes[0]["p"].nsa, args.nx),
                      "nwfs": nwfs, "launch": "one HIP graph per step (one branch per powfs)"},
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3), "steps": len(lat)},
           "classes": {c["name"]: {"nwfs": len(c["wfs"]), "nsa": c["p"].nsa, "notf": c["p"].notf,
                                   "nwvl": c["p"].nwvl, "kernels_us_alone": c["us"],
                                   "autotune": c["tune"]} for c in classes},
           "parity": parity}
    print(json.dumps(out), flush=True)
    group.close()
    for c in classes:
        c["eng"].close()
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)


def evl_main(args):
    """Science PSF steps/s: the on-axis OPD of a 30 m aperture (3.6 m central
    obscuration) sampled at 1/64 m, traced through the seven 4096^2 screens,
    and its PSF added to the long-exposure sums at 1.25, 1.65 and 2.2 um on a
    4096^2 FFT grid with the full 4096^2 PSF kept (MAOS's evl.wvl default,
    evl.psfgridsize 4096, psfsize = the grid).  Per-pass kernel times come
    from HIP events on the handle's stream."""
    from maos_amd import configs, engine
    from maos_amd.powfs import genatm
    import oracle

    t0 = time.time()
    cfg = configs.lgs30m(nx=args.nx, order=args.order)
    layers = genatm(cfg["atm"])
    dt = cfg["sim"].dt
    D, din, dx, N = 30.0, 3.6, 1.0 / 64, 4096
    n = int(round(D / dx)) + 1
    c = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(c, c)
    r = np.hypot(X, Y)
    m = (r <= D / 2 + 1e-9) & (r >= din / 2)
    locx, locy = X[m].copy(), Y[m].copy()
    del X, Y, r, m
    nloc = len(locx)
    amp = np.full(nloc, 1.0 / nloc)
    emb, _ = engine.loc_embed(locx, locy, dx, nembed=N)
    wvls = [1.25e-6, 1.65e-6, 2.2e-6]
    P = N
    engine.init(0)
    engine.atm2gpu(layers)
    ev = engine.Evl(amp, wvls, [N] * 3, [P] * 3, np.tile(emb, 3), locx=locx, locy=locy)
    log("evl: %d points, %d^2 FFT, 3 wavelengths (%.1fs)" % (nloc, N, time.time() - t0))
    isim = 0
    for _ in range(args.warmup):
        ev.trace(isim, dt)
        isim += 1
    ev.wait()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ev.trace(isim, dt)
        isim += 1
    ev.wait()
    sec = time.perf_counter() - t1
    log("evl: %d steps in %.3fs" % (args.steps, sec))
    ev.timing_enable(True)
    nt = min(args.steps, 50)
    for _ in range(nt):
        ev.trace(isim, dt)
        isim += 1
    _, ms_trace, ms_rows, ms_cols = ev.timing()
    ev.timing_enable(False)
    us = {"k_evl_opd": ms_trace / nt * 1e3, "k_psf_rows<4096>": ms_rows / nt / 3 * 1e3,
          "k_psf_cols<4096>": ms_cols / nt / 3 * 1e3}
    # algorithmic bytes and FLOPs per launch (one wavelength)
    nband = int(emb.max() // N - emb.min() // N + 1)
    by_rows = 16.0 * nloc + 8.0 * nband * P
    by_cols = 8.0 * nband * P + 8.0 * P * P
    fl_rows = 5.0 * N * math.log2(N) * nband
    fl_cols = 5.0 * N * math.log2(N) * P
    by_trace = 20.0 * nloc + 4.0 * len(layers) * nloc  # coordinates, OPD, one screen value per point and layer
    dom = max(("k_psf_rows<4096>", by_rows, fl_rows), ("k_psf_cols<4096>", by_cols, fl_cols),
              key=lambda t: us[t[0]])
    ach = dom[1] / (us[dom[0]] * 1e-6) / 1e9
    # CPU: the FP64 oracle's trace (7 layers) and one wavelength's PSF, the
    # step counted as the trace plus three PSFs
    out_opd = np.zeros(nloc)
    import ctypes as C
    dp = C.POINTER(C.c_double)
    cpu = None
    if not args.no_cpu:
        tc0 = time.perf_counter()
        for L in layers:
            scr = np.ascontiguousarray(L.p)
            oracle.lib().orc_prop(scr.ctypes.data_as(dp), L.nx, L.ny, L.dx, L.ox, L.oy, L.ht,
                                  locx.ctypes.data_as(dp), locy.ctypes.data_as(dp), nloc, 1, L.dx,
                                  out_opd.ctypes.data_as(dp), 1.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                                  -L.vx * dt * isim, -L.vy * dt * isim, 1)
        tc1 = time.perf_counter()
        oracle.locfft_psf(out_opd, amp, emb, N, P, wvls[1])
        tc2 = time.perf_counter()
        cpu_step = (tc1 - tc0) + 3 * (tc2 - tc1)
        cpu = {"value": round(1.0 / cpu_step, 4), "unit": "PSF steps/s", "cores": 1, "kind": "port",
               "sample": "one step: the FP64 oracle's trace of 7 layers onto %d points (%.2f s) and one "
                         "wavelength's 4096^2 PSF (%.2f s, counted three times)" % (nloc, tc1 - tc0, tc2 - tc1)}
    fps = args.steps / sec
    out = {"metric": "science PSF steps/s, 30 m aperture, 3 wavelengths, 4096^2 FFT", "value": round(fps, 2),
           "unit": "PSF steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(sec / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded von Karman screens, uniform amplitude)",
           "config": {"workload": "on-axis science evaluation: 30 m aperture (3.6 m obscuration) at 1/64 m, "
                                  "%d points, 7 layers %d^2, PSF at 1.25/1.65/2.2 um on a 4096^2 grid, full "
                                  "4096^2 PSF accumulated" % (nloc, args.nx)},
           "kernels_us": {k: round(v, 2) for k, v in us.items()},
           "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(ach, 1), "peak": HBM_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_GBS, 4),
                        "traffic": read_traffic(EVL_TRAFFIC_FILE) if dom[0] == "k_psf_cols<4096>" else None,
                        "fft_tflops": round(dom[2] / (us[dom[0]] * 1e-6) / 1e12, 2)},
           "algorithmic": {"k_psf_rows_bytes": by_rows, "k_psf_cols_bytes": by_cols, "k_psf_rows_flop": fl_rows,
                           "k_psf_cols_flop": fl_cols, "k_evl_opd_bytes": by_trace, "band_rows": nband}}
    if cpu:
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(fps / cpu["value"], 1)
    # parity at the last step: the GPU PSF of the traced OPD against the
    # oracle's PSF of the same OPD (one more step, accumulators cleared)
    ev.reset()
    ev.trace(isim, dt)
    g_opd = ev.opd()[0]
    g_psf, _ = ev.psf(0, 1)
    want = oracle.locfft_psf(g_opd, amp, emb, N, P, wvls[1]) / amp.sum() ** 2
    out["parity"] = {"psf_rel_l2_1.65um": float(np.linalg.norm(g_psf - want) / np.linalg.norm(want)),
                     "strehl": float(g_psf.max())}
    print(json.dumps(out), flush=True)
    ev.close()
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)


def pywfs_main(args):
    """Pyramid WFS frames/s: the single-conjugate 30 m pyramid of MAOS's
    scao_pywfs.conf (powfs_pyngs.conf and pywfs.conf defaults): 60 x 60
    subapertures, pupil sampled at 1/32 m (715,336 points), FFT grid
    3 x the pupil (2916^2 = 2^2 3^6), 0.7 um, four sides, modulation radius
    5 lambda/D at 32 positions, sigmatch 2, traced through the seven 4096^2
    screens each step.  The GPU calibrates the sensor first (pywfs_new's
    sequence).  --pywfs-parity also compares one frame with the FP64 oracle
    (about a minute of host time)."""
    from maos_amd import configs, engine
    from maos_amd.powfs import genatm
    from maos_amd.pywfs import Pywfs, PywfsCfg

    t0 = time.time()
    cfg = configs.lgs30m(nx=args.nx, order=args.order)
    layers = genatm(cfg["atm"])
    dt = cfg["sim"].dt
    pw = Pywfs(PywfsCfg(D=30.0, din=3.6, order=60, dx=1.0 / 32))
    engine.init(0)
    engine.atm2gpu(layers)

    def gpu(p):
        def f(opd):
            e = engine.PywfsEngine(p)
            out = e.image(opd)
            e.close()
            return out
        return f
    pw.calibrate(gpu)
    eng = engine.PywfsEngine(pw)
    log("pywfs: %d points, %d^2 FFT, nsa %d, %d modulation positions (%.1fs)"
        % (pw.nloc, pw.N, pw.nsa, pw.cfg.modulpos, time.time() - t0))
    isim = 0
    for _ in range(args.warmup):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    sec = time.perf_counter() - t1
    fps = args.steps / sec
    eng.timing_enable(True)
    nt = max(3, min(20, args.steps // 5))
    for _ in range(nt):
        eng.queue(isim, dt)
        isim += 1
    n, ms = eng.timing()
    eng.timing_enable(False)
    N = pw.N
    # algorithmic work of one frame: the field's 2-D FFT, per modulation
    # position a 2-D inverse FFT, and the detector's forward and inverse
    # 2-D FFTs (5 M log2 M each, M = N^2); HBM: every line pass reads and
    # writes the N^2 complex grid once, a 2-D transform is two line passes
    # and a transpose
    M = float(N) * N
    npos = pw.cfg.modulpos * pw.cfg.modulring
    nfft = 1 + npos + 2
    flop = nfft * 5 * M * math.log2(M)
    out = {"metric": "pyramid WFS frames/s (30 m, 60x60, 2916^2 FFT, 32 modulation positions)",
           "value": round(fps, 2), "unit": "WFS frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded von Karman screens)",
           "config": {"workload": "scao pyramid: 30 m pupil at 1/32 m, 60x60 subapertures, 0.7 um, 4 sides, "
                      "modulation 5 lambda/D x 32, 7 layers 4096^2", "nloc": pw.nloc, "nembed": N, "nsa": pw.nsa},
           "kernel_ms_per_frame": round(ms / n, 4),
           "fft_tflops": round(flop / (ms / n * 1e-3) / 1e12, 2)}
    if args.pywfs_parity:
        import oracle
        eng.queue(isim, dt)
        gi, gg = eng.sync()
        opd = eng.opd()
        tc = time.perf_counter()
        ri, rg = oracle.pywfs_image(pw, opd)
        tcpu = time.perf_counter() - tc
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        out["parity"] = {"ints_rel_l2": float(np.linalg.norm(gi - ri) / np.linalg.norm(ri)),
                         "grad_max_abs_mas": float(np.abs(gg - rg).max() * 180 / math.pi * 3600e3)}
        out["cpu_baseline"] = {"value": round(1.0 / tcpu, 5), "unit": "WFS frames/s", "cores": threads,
                               "kind": "port", "sample": "one frame of the same workload, FP64 oracle"}
    print(json.dumps(out), flush=True)
    eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--i0-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--order", type=int, default=60)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--variant", choices=("auto", "fft32", "mfma", "fp16"), default="auto")
    ap.add_argument("--copy-calib", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true", help="launch kernel by kernel, not one graph per step")
    ap.add_argument("--nreal", type=int, default=1,
                    help="advance this many independent realisations together (r0 ~ U[0.12, 0.25] m)")
    ap.add_argument("--latency-steps", type=int, default=100)
    ap.add_argument("--evl", action="store_true",
                    help="science PSF steps/s (30 m aperture, 3 wavelengths, 4096^2 FFT) instead of the WFS step")
    ap.add_argument("--pywfs", action="store_true",
                    help="pyramid WFS frames/s (30 m, 60x60, 32 modulation positions) instead of the LGS step")
    ap.add_argument("--pywfs-parity", action="store_true",
                    help="with --pywfs: one frame against the FP64 oracle (about a minute of host time)")
    ap.add_argument("--hetero", action="store_true",
                    help="LGS + low-order NGS + TT NGS stepped as one group graph per step")
    ap.add_argument("--split", type=int, default=0,
                    help="lgs30m µs per step as one engine and as K grouped engines of 6/K WFS (K = 2, 3, 6)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the WFS path of this build is single-GPU "
              "(see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    if args.nreal > 1:
        batch_main(args)
        return
    if args.hetero:
        hetero_main(args)
        return
    if args.split:
        split_main(args)
        return
    if args.evl:
        evl_main(args)
        return
    if args.pywfs:
        pywfs_main(args)
        return

    from maos_amd import _lib, configs, engine
    from maos_amd.powfs import Powfs, genatm

    t0 = time.time()
    cfg = configs.lgs30m(nx=args.nx, order=args.order)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    log("powfs: nsa=%d nxsa=%d notf=%d npix=%d (%.1fs)" % (p.nsa, p.nxsa, p.notf, p.npix,
                                                         time.time() - t0))
    layers = genatm(cfg["atm"])
    log("screens: %d x %d^2 (%.1fs)" % (len(layers), args.nx, time.time() - t0))
    wfs = configs.wfs_specs(cfg, p)
    nwfs = len(wfs)
    dt = cfg["sim"].dt
    engine.init(0)
    engine.atm2gpu(layers)
    eng = engine.WfsGrad(p, wfs)
    if args.variant != "auto":
        eng.set_variant({"fft32": _lib.VARIANT_FFT32, "mfma": _lib.VARIANT_MFMA_BF16,
                         "fp16": _lib.VARIANT_MFMA_FP16}[args.variant])
    variant = eng.variant()
    log("imaging form: %s" % KERNELS[variant])
    if not args.no_graph:
        eng.set_graph(True)

    # matched filter from the FP64 long-exposure i0 (mtch_cell on the host)
    eng.update_ref(np.zeros(p.nsa * 2 * p.npix), np.ones(p.nsa))
    eng.i0_control(enable=True, reset=True)
    for isim in range(args.i0_steps):
        eng.queue(isim, dt)
    eng.wait()
    refs = []
    for iwfs in range(nwfs):
        i0, n = eng.i0_get(iwfs)
        mtche, i0sum, _ = p.mtch(i0 / n)
        eng.update_ref(mtche, i0sum, iwfs=iwfs)
        refs.append((mtche, i0sum))
    eng.i0_control(enable=True, reset=True)
    log("matched filter from %d-frame FP64 i0 (%.1fs)" % (args.i0_steps, time.time() - t0))

    isim = args.i0_steps
    for _ in range(args.warmup):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.queue(isim, dt)
        isim += 1
    eng.wait()
    t2 = time.perf_counter()
    sec = t2 - t1
    fps = nwfs * args.steps / sec
    ms_step = sec / args.steps * 1e3
    log("timed: %d steps in %.3fs -> %.1f frames/s" % (args.steps, sec, fps))

    # per-launch kernel durations with HIP events on the engine's stream
    nt = min(100, max(10, args.steps // 5))
    eng.timing_enable(True)
    for _ in range(nt):
        eng.queue(isim, dt)
        isim += 1
    nl, ms_llt, ms_opd, ms_ints = eng.timing()
    eng.timing_enable(False)
    t_ints = ms_ints / nl * 1e-3
    t_llt = ms_llt / nl * 1e-3
    t_opd = ms_opd / nl * 1e-3
    # imaging kernel only: the ray trace runs in its own kernel
    flop_launch = flops_per_subap(p.notf, p.nxsa, 0, p.npix) * p.nsa * nwfs
    achieved = flop_launch / t_ints / 1e12
    traffic = read_traffic() if variant == _lib.VARIANT_MFMA_BF16 else None
    is_mfma = variant in (_lib.VARIANT_MFMA_BF16, _lib.VARIANT_MFMA_FP16)
    # The unit that does the transforms sets the roof: the BF16/FP16 dense
    # matrix-core peak for the MFMA forms, the FP32 vector peak for the
    # radix-FFT form.  achieved is the algorithmic (FFT-convention) rate.
    peak = PEAK_BF16_TFLOPS if is_mfma else PEAK_FP32_TFLOPS
    roof = {"bound": "mfma" if is_mfma else "valu", "achieved": round(achieved, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
            "kernel": KERNELS[variant], "avg_launch_us": round(t_ints * 1e6, 2),
            "llt_kernel_us": round(t_llt * 1e6, 2), "raytrace_kernel_us": round(t_opd * 1e6, 2),
            "flop_per_launch": flop_launch,
            "peak_kind": ("bf16/fp16 dense MFMA (MI355X published)" if is_mfma
                          else "fp32 vector (MI355X published)"),
            "frac_of_fp32_vector_peak": round(achieved / PEAK_FP32_TFLOPS, 4)}
    if is_mfma:
        ex = mfma_flops_per_subap() * p.nsa * nwfs
        roof["mfma_executed_tflops"] = round(ex / t_ints / 1e12, 2)
        roof["mfma_executed_frac_of_bf16_peak"] = round(ex / t_ints / 1e12 / PEAK_BF16_TFLOPS, 4)
        if traffic is not None:
            roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)
    other = []
    # every other imaging form, and the FP32 register-FFT kernel that the
    # FP32 form falls back to: frames/s and the imaging kernel's time
    forms = [(v, FORMS[v], KERNELS[v], False) for v in (0, 1, 2) if v != variant and eng.variant_ok(v)]
    forms.append((0, "fp32-register-fft", "k_shwfs_step<64,32,1>", True))
    for v, form, kname, reg in forms:
        if reg:
            os.environ["MX_FFT_REG"] = "1"
        fps_other, isim = time_variant(eng, v, isim, dt, 100)
        eng.timing_enable(True)
        for _ in range(20):
            eng.queue(isim, dt)
            isim += 1
        nl_o, _, _, ms_o = eng.timing()
        eng.timing_enable(False)
        os.environ.pop("MX_FFT_REG", None)
        other.append({"form": form, "kernel": kname, "frames_per_s": round(fps_other * nwfs, 2),
                      "imaging_kernel_us": round(ms_o / nl_o * 1e3, 2)})
    eng.set_variant(variant)

    out = {"metric": METRIC, "value": round(fps, 2), "unit": "WFS frames/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": {0: "fp32", 1: "bf16x3", 2: "fp16x3"}[variant],
           "data": "synthetic (seeded von Karman screens, Gaussian sodium profile)",
           "config": {"workload": "lgs30m: 30 m pupil, 6 LGS WFS x 60x60 (nsa=%d), 32x32 "
                      "OPD pts -> 64x64 FFT, 6x6 px of 1\", 7 layers 4096^2 @1/64 m, LLT "
                      "uplink, ETF, matched filter, 800 Hz" % p.nsa,
                      "nwfs": nwfs, "nsa": p.nsa, "notf": p.notf, "npix": p.npix,
                      "variant": FORMS[variant],
                      "launch": "kernel by kernel" if args.no_graph else "one HIP graph per step"},
           "roofline": roof}
    if other:
        out["other_forms"] = other

    # CPU baseline: FP64 oracle on the host cores, bounded sample of frames
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        o = oracle.Oracle(p)
        o.atm(layers)
        ref_isim = isim
        o.set_ref(*refs[0])
        nf = 0
        tc0 = time.perf_counter()
        ri = rg = None
        while True:
            iw = nf % nwfs
            r = o.step(wfs[iw], ref_isim, dt, nthread=threads)
            if nf == 0:
                ri, rg = r[0], r[1]
            nf += 1
            if time.perf_counter() - tc0 > args.cpu_seconds:
                break
        tc = time.perf_counter() - tc0
        cpu_fps = nf / tc
        out["cpu_baseline"] = {"value": round(cpu_fps, 3), "unit": "WFS frames/s", "cores": threads,
                               "kind": "port",
                               "sample": "%d full WFS frames (nsa=%d) of the same workload, FP64 "
                                         "oracle with OpenMP over subapertures" % (nf, p.nsa)}
        out["gpu_over_cpu"] = round(fps / cpu_fps, 1)
        # every imaging form against the FP64 oracle on WFS 0 at one step
        parity = {}
        for v in (0, 1, 2):
            if eng.variant_ok(v):
                eng.set_variant(v)
                eng.queue(ref_isim, dt)
                g_v, ints_v = eng.sync(0)
                parity[FORMS[v]] = {
                    "ints_rel_l2": float(np.linalg.norm(ints_v - ri) / np.linalg.norm(ri)),
                    "grad_max_abs_mas": float(np.abs(g_v - rg).max() * 180 / math.pi * 3600e3)}
        eng.set_variant(variant)
        out["parity"] = {"wfs": 0, "isim": ref_isim, "forms": parity}
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
