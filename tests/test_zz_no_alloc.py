"""mx_evl_wfe_noalloc (the piston/tip/tilt wavefront error with its partial
sums in the library's static scratch array), called through the C ABI,
against mx_evl_wfe (partial sums in the handle's own buffer) on the same
OPD: two calls in a row, the second on a smaller handle (fewer points and
directions), so stale scratch contents from the first call would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip_runtime():
    return C.CDLL("libamdhip64.so")


def _noalloc():
    from maos_amd import _lib
    return _lib.hip().mx_evl_wfe_noalloc


def _annulus(D, din, dx):
    n = int(round(D / dx)) + 1
    c = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(c, c)
    r = np.hypot(X, Y)
    m = (r <= D / 2) & (r >= din / 2)
    return X[m].copy(), Y[m].copy()


def _case(D, din, dx, ndir, seed):
    """An evaluation handle with point coordinates and an OPD of ndir
    directions: a plane (piston, tip, tilt) plus a random part."""
    from maos_amd import engine
    locx, locy = _annulus(D, din, dx)
    emb, N = engine.loc_embed(locx, locy, dx, nembed=256 if D > 1 else 64)
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 1.0, len(locx))
    amp /= amp.sum()
    dirs = [(1e-6 * i, -1e-6 * i, 0.0) for i in range(ndir)]
    ev = engine.Evl(amp, [1e-6], [N], [1], emb, locx=locx, locy=locy, dirs=dirs)
    opd = np.empty((ndir, len(locx)))
    for d in range(ndir):
        p, tx, ty = rng.normal(size=3) * 1e-7
        opd[d] = p + tx * locx + ty * locy + rng.normal(size=len(locx)) * 3e-8
    return ev, opd.astype(np.float32)


def _check(ev, opd32, rt, f):
    from maos_amd import _lib
    ev.psf_add(opd32.astype(np.float64))
    rms_ref, coeff_ref = ev.wfe()
    d = C.c_void_p()
    assert rt.hipMalloc(C.byref(d), opd32.nbytes) == 0
    try:
        assert rt.hipMemcpy(d, opd32.ctypes.data_as(C.c_void_p), opd32.nbytes, 1) == 0  # host to device
        rms = np.zeros((ev.ndir, 3))
        coeff = np.zeros((ev.ndir, 3))
        rc = f(ev.h, d, rms.ctypes.data_as(_lib.c_double_p), coeff.ctypes.data_as(_lib.c_double_p))
        assert rc == 0, _lib.hip().mx_last_error()
    finally:
        ev.wait()
        rt.hipFree(d)
    print("ndir=%d nloc=%d p/t/t-removed rms %s" % (ev.ndir, ev.nloc, np.sqrt(rms[:, 2])), flush=True)
    # the same kernels add the same numbers in the same order
    np.testing.assert_array_equal(rms, rms_ref)
    np.testing.assert_array_equal(coeff, coeff_ref)


def test_wfe_noalloc_matches_wfe_twice():
    from maos_amd import _lib
    rt, f = _hip_runtime(), _noalloc()
    big, opd_big = _case(2.0, 0.3, 1 / 64, 5, seed=1)
    small, opd_small = _case(1.0, 0.2, 1 / 32, 2, seed=2)
    print("case: noalloc WFE, %d points x 5 dirs then %d points x 2 dirs" % (big.nloc, small.nloc), flush=True)
    _check(big, opd_big, rt, f)
    _check(small, opd_small, rt, f)
    # a null OPD is refused before anything is queued
    rms = np.zeros((small.ndir, 3))
    assert f(small.h, None, rms.ctypes.data_as(_lib.c_double_p), None) != 0
    big.close()
    small.close()
