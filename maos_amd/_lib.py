"""ctypes bindings of the native libraries.

libmaos_mi355x.so is the GPU path (include/maos_mi355x.h); libmaos_setup.so
holds the host FP64 table builders (include/maos_setup.h).  Loading fails
loudly when a library has not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from .build import HIP_LIB, SETUP_LIB

c_double_p = C.POINTER(C.c_double)
c_long_p = C.POINTER(C.c_long)


class MxAtmLayer(C.Structure):
    _fields_ = [("p", c_double_p), ("nx", C.c_long), ("ny", C.c_long),
                ("dx", C.c_double), ("dy", C.c_double), ("ox", C.c_double),
                ("oy", C.c_double), ("ht", C.c_double), ("vx", C.c_double),
                ("vy", C.c_double)]


class MxPowfs(C.Structure):
    _fields_ = [("nsa", C.c_long), ("nxsa", C.c_long), ("embfac", C.c_long),
                ("notf", C.c_long), ("pixpsax", C.c_long), ("pixpsay", C.c_long),
                ("nwvl", C.c_long), ("dx", C.c_double), ("ptsx", c_double_p),
                ("ptsy", c_double_p), ("amp", c_double_p), ("saa", c_double_p),
                ("saat", C.c_double), ("areascale", C.c_double), ("wvl", c_double_p),
                ("phytype", C.c_int), ("sigmatch", C.c_int), ("cogthres", C.c_double),
                ("cogoff", C.c_double), ("pixthetax", C.c_double),
                ("pixthetay", C.c_double), ("siglev_dtrat", C.c_double),
                ("saasum", C.c_double), ("llt_nx", C.c_long), ("llt_origx", C.c_double),
                ("llt_origy", C.c_double), ("llt_ox", C.c_double), ("llt_oy", C.c_double),
                ("llt_amp", c_double_p)]


class MxWfs(C.Structure):
    _fields_ = [("thetax", C.c_double), ("thetay", C.c_double), ("hs", C.c_double),
                ("sigsim", C.c_double), ("misregx", C.c_double), ("misregy", C.c_double),
                ("wvlwts", c_double_p)]


class MxsDither(C.Structure):
    """mxs_dither_t: keys, then state (zero at the start)."""
    _fields_ = [("npoint", C.c_int), ("pllrat", C.c_int), ("pllskip", C.c_int), ("dtrat", C.c_int),
                ("alfsm", C.c_int), ("detrend", C.c_int), ("amp", C.c_double), ("gpll", C.c_double),
                ("gdrift", C.c_double), ("delta", C.c_double), ("deltam", C.c_double), ("deltao", C.c_double),
                ("a2m", C.c_double), ("a2me", C.c_double)]


class MxPywfsCfg(C.Structure):
    _fields_ = [("nloc", C.c_long), ("locx", c_double_p), ("locy", c_double_p), ("amp", c_double_p),
                ("embed", c_long_p), ("nembed", C.c_long), ("dx", C.c_double), ("nwvl", C.c_int),
                ("wvl", c_double_p), ("wvlwts", c_double_p), ("mask", c_double_p), ("nominal", c_double_p),
                ("nsa", C.c_long), ("salocx", c_double_p), ("salocy", c_double_p), ("nside", C.c_int),
                ("sioff", c_double_p), ("modulate", C.c_double), ("modulpos", C.c_int), ("modulring", C.c_int),
                ("sigmatch", C.c_int), ("raw", C.c_int), ("siglev", C.c_double), ("thetax", C.c_double),
                ("thetay", C.c_double), ("hs", C.c_double)]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need contiguous float64"
    return a.ctypes.data_as(c_double_p)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous, "need contiguous int64"
    return a.ctypes.data_as(c_long_p)


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(
            "native library %s is missing: build it with `python -m maos_amd.build` "
            "(there is no fallback path)" % path)
    return C.CDLL(path)


_hip = None
_setup = None

# name -> (restype, argtypes) of every symbol in include/maos_mi355x.h
HIP_API = {
    "mx_init": (C.c_int, [C.c_int]),
    "mx_last_error": (C.c_char_p, []),
    "mx_wfsgrad_init": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(MxPowfs), C.POINTER(MxWfs), C.c_int]),
    "mx_wfsgrad_init_batch": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(MxPowfs), C.POINTER(MxWfs), C.c_int,
                                        C.c_int]),
    "mx_wfsgrad_free": (C.c_int, [C.c_void_p]),
    "mx_wfsgrad_set_dtf": (C.c_int, [C.c_void_p, C.c_int, c_double_p, C.c_long, C.c_long,
                                     c_long_p, c_long_p, c_double_p]),
    "mx_wfsgrad_set_dtf_sa": (C.c_int, [C.c_void_p, C.c_int, c_double_p, C.c_long, C.c_long, C.c_long,
                                        c_long_p, c_long_p, c_double_p]),
    "mx_wfsgrad_update_etf": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, C.c_double]),
    "mx_wfsgrad_update_ref": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p]),
    "mx_atm2gpu": (C.c_int, [C.POINTER(MxAtmLayer), C.c_int, C.c_double]),
    "mx_atm2gpu_real": (C.c_int, [C.c_int, C.POINTER(MxAtmLayer), C.c_int, C.c_double]),
    "mx_dm2gpu": (C.c_int, [C.POINTER(MxAtmLayer), C.c_int]),
    "mx_dm2gpu_real": (C.c_int, [C.c_int, C.POINTER(MxAtmLayer), C.c_int]),
    "mx_wfsgrad_queue": (C.c_int, [C.c_void_p, C.c_long, C.c_double]),
    "mx_wfsgrad_sync": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p]),
    "mx_wfsgrad_i0_control": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mx_wfsgrad_i0_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_long_p]),
    "mx_wfsgrad_pistat_control": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mx_wfsgrad_pistat_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_long_p]),
    "mx_wfsgrad_psfout": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mx_wfsgrad_set_etf_weight": (C.c_int, [C.c_void_p, C.c_double]),
    "mx_wfsgrad_set_gs0": (C.c_int, [C.c_void_p, c_long_p, c_long_p, c_double_p, C.c_long]),
    "mx_wfsgrad_dither_control": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mx_wfsgrad_set_dither": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
    "mx_wfsgrad_dither_fold": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
    "mx_wfsgrad_set_i0model": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_double_p]),
    "mx_wfsgrad_mtch_update": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int, C.c_double]),
    "mx_wfsgrad_dither_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_double_p,
                                        c_long_p]),
    "mx_wfsgrad_mtch_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_double_p,
                                      c_double_p, c_double_p, c_double_p]),
    "mx_wfsgrad_debug_opd": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mx_wfsgrad_debug_lotf": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p]),
    "mx_wfsgrad_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_wfsgrad_timing": (C.c_int, [C.c_void_p, c_long_p, c_double_p, c_double_p, c_double_p]),
    "mx_wfsgrad_wait": (C.c_int, [C.c_void_p]),
    "mx_wfsgrad_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_wfsgrad_get_variant": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mx_wfsgrad_set_dtrat": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_wfsgrad_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_wfsgrad_set_opdadd": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mx_wfsgrad_set_ptt": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mx_wfsgrad_set_geom": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_wfsgrad_set_noise": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_ulonglong]),
    "mx_wfsgrad_autotune": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_int, C.c_char_p, C.c_long]),
    "mx_wfsgroup_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]),
    "mx_wfsgroup_queue": (C.c_int, [C.c_void_p, C.c_long, C.c_double]),
    "mx_wfsgroup_free": (C.c_int, [C.c_void_p]),
    "mx_pywfs_init": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(MxPywfsCfg)]),
    "mx_pywfs_free": (C.c_int, [C.c_void_p]),
    "mx_pywfs_set_calib": (C.c_int, [C.c_void_p, C.c_double, c_double_p, c_double_p]),
    "mx_pywfs_ints_opd": (C.c_int, [C.c_void_p, c_double_p, C.c_double]),
    "mx_pywfs_queue": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_double]),
    "mx_pywfs_sync": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "mx_pywfs_opd_get": (C.c_int, [C.c_void_p, c_double_p]),
    "mx_pywfs_set_real": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_pywfs_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_pywfs_timing": (C.c_int, [C.c_void_p, c_long_p, c_double_p]),
    "mx_pywfs_wait": (C.c_int, [C.c_void_p]),
    "mx_evl_init": (C.c_int, [C.POINTER(C.c_void_p), C.c_long, c_double_p, c_double_p, c_double_p, C.c_int,
                              c_double_p, c_double_p, c_double_p, C.c_int, c_double_p, c_long_p, c_long_p,
                              c_long_p]),
    "mx_evl_free": (C.c_int, [C.c_void_p]),
    "mx_evl_psf_add": (C.c_int, [C.c_void_p, c_double_p]),
    "mx_evl_psf_add_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mx_evl_trace": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_int]),
    "mx_evl_opd_get": (C.c_int, [C.c_void_p, c_double_p]),
    "mx_evl_set_real": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_evl_wfe": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "mx_evl_wfe_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, c_double_p, c_double_p]),
    "mx_evl_psf_get": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, c_long_p]),
    "mx_evl_psf_reset": (C.c_int, [C.c_void_p]),
    "mx_evl_wait": (C.c_int, [C.c_void_p]),
    "mx_evl_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "mx_evl_timing": (C.c_int, [C.c_void_p, c_long_p, c_double_p, c_double_p, c_double_p]),
}

VARIANT_AUTO, VARIANT_FFT32, VARIANT_MFMA_BF16, VARIANT_MFMA_FP16 = -1, 0, 1, 2

# name -> (restype, argtypes) of every symbol in include/maos_setup.h
SETUP_API = {
    "mxs_bin_open": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    "mxs_bin_close": (C.c_int, [C.c_void_p]),
    "mxs_bin_write_header": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_char_p]),
    "mxs_bin_write_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "mxs_bin_read_header": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64), C.c_char_p, C.c_uint64]),
    "mxs_bin_read_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "mxs_zfarr_open": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p, C.c_long, C.c_char_p]),
    "mxs_zfarr_push": (C.c_int, [C.c_void_p, C.c_long, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p,
                                 C.c_uint64]),
    "mxs_zfarr_close": (C.c_int, [C.c_void_p]),
    "mxs_fft2": (None, [c_double_p, C.c_long, C.c_long, C.c_int]),
    "mxs_nextfftsize": (C.c_long, [C.c_long]),
    "mxs_saloc": (C.c_long, [C.c_int, C.c_double, C.c_double, C.c_double, c_double_p,
                             c_double_p, C.c_long]),
    "mxs_loc_circle_add": (None, [c_double_p, c_double_p, c_double_p, C.c_long, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double]),
    "mxs_saa": (None, [c_double_p, C.c_long, C.c_long, C.c_double, c_double_p]),
    "mxs_sa_reduce": (C.c_long, [c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                 c_double_p, C.c_long, C.c_long, C.c_double, C.c_double]),
    "mxs_dtf_nominal": (None, [c_double_p, C.c_long, C.c_long, C.c_double, C.c_double,
                               C.c_double, C.c_double]),
    "mxs_si": (C.c_long, [C.c_long, C.c_long, C.c_double, C.c_long, C.c_long, C.c_double,
                          C.c_double, C.c_double, C.c_double, C.c_double, c_long_p,
                          c_long_p, c_double_p]),
    "mxs_srot_srsa": (None, [c_double_p, c_double_p, C.c_long, C.c_double, C.c_double,
                             C.c_double, c_double_p, c_double_p]),
    "mxs_smooth": (C.c_long, [c_double_p, C.c_long, C.c_double, c_double_p, C.c_long]),
    "mxs_mketf": (C.c_int, [c_double_p, c_double_p, c_double_p, C.c_long, C.c_double,
                            c_double_p, C.c_long, c_double_p, c_double_p, C.c_long,
                            C.c_double, C.c_double, C.c_double, C.c_int]),
    "mxs_llt_amp": (C.c_long, [C.c_double, C.c_double, C.c_double, C.c_double, c_double_p,
                               C.c_long, c_double_p]),
    "mxs_genscreen": (None, [c_double_p, C.c_long, C.c_long, C.c_double, C.c_double,
                             C.c_double, c_double_p, C.c_long, C.c_ulonglong]),
    "mxs_derive_by_fft": (None, [c_double_p, c_double_p, C.c_long, C.c_long, C.c_double]),
    "mxs_mtch": (C.c_int, [c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                           c_double_p, c_double_p, C.c_long, C.c_long, C.c_long,
                           C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                           C.c_int, C.c_double]),
    "mxs_turbcov": (C.c_int, [c_double_p, C.c_long, C.c_double, C.c_double, c_double_p]),
    "mxs_mkg": (C.c_long, [c_long_p, c_long_p, c_double_p, C.c_long, c_double_p, c_double_p, C.c_long,
                           C.c_long, C.c_double, c_double_p, c_double_p, C.c_double]),
    "mxs_loc_create_embed": (C.c_long, [c_long_p, c_long_p, C.c_long, c_double_p, c_double_p, C.c_double,
                                        C.c_double, C.c_double, C.c_int]),
    "mxs_calc_dither_amp": (C.c_double, [c_double_p, c_double_p, C.c_long, C.c_long, C.c_long, C.c_long, C.c_int,
                                         C.c_int]),
    "mxs_dither_step": (C.c_int, [C.c_void_p, C.c_long, c_double_p, c_double_p, c_double_p]),
    "mxs_dither_output": (C.c_int, [C.c_void_p, c_double_p, c_double_p, C.c_long, c_double_p, c_double_p]),
    "mxs_dither_tt": (None, [C.c_void_p, C.c_double, C.c_double, c_double_p]),
    "mxs_sodium_fit": (C.c_long, [c_double_p, C.c_long, c_double_p, c_double_p, c_double_p, c_double_p, C.c_long,
                                  C.c_long, C.c_long, C.c_long, C.c_long, c_double_p, c_double_p, c_long_p, c_long_p,
                                  c_double_p, C.c_long, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                  c_double_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "mxs_pywfs_pyramid": (C.c_int, [c_double_p, C.c_long, C.c_int, C.c_long, C.c_double, C.c_double]),
    "mxs_pywfs_nominal": (None, [c_double_p, C.c_long, C.c_double, C.c_double, C.c_double]),
    "mxs_pywfs_saloc": (C.c_long, [c_double_p, c_double_p, C.c_long, C.c_long, C.c_double]),
    "mxs_pywfs_sioff": (C.c_int, [c_double_p, C.c_int]),
    "mxs_dither_position": (None, [c_double_p, c_double_p, C.c_int, C.c_int, C.c_int, C.c_long,
                                   C.c_double]),
    "mxs_genotf": (C.c_int, [c_double_p, C.c_long, C.c_double, c_double_p, c_double_p, C.c_long,
                             C.c_double, C.c_double, c_double_p, C.c_double, C.c_double, C.c_long,
                             C.c_int]),
    "mxs_gensepsf": (C.c_int, [c_double_p, c_double_p, C.c_long, C.c_long, c_double_p, C.c_long,
                               c_double_p, C.c_long]),
    "mxs_gensei": (C.c_int, [c_double_p, c_double_p, c_double_p, c_double_p, C.c_long, C.c_long,
                             c_double_p, c_double_p, C.c_long, c_long_p, c_long_p, c_double_p,
                             C.c_double, C.c_double, C.c_int, C.c_int, c_double_p]),
}


def _bind(lib, api):
    for name, (res, args) in api.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def hip():
    """libmaos_mi355x.so (loads it on first use)."""
    global _hip
    if _hip is None:
        _hip = _bind(_load(HIP_LIB), HIP_API)
    return _hip


def setup():
    """libmaos_setup.so (loads it on first use)."""
    global _setup
    if _setup is None:
        _setup = _bind(_load(SETUP_LIB), SETUP_API)
    return _setup


def check(rc):
    """Raise on a negative return code of libmaos_mi355x.so."""
    if rc != 0:
        msg = hip().mx_last_error()
        raise RuntimeError("maos_mi355x error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc
