"""MAOS .bin files: read tables and telemetry, record a run's outputs.

The C side is `mxs_bin_*` / `mxs_zfarr_*` in libmaos_setup.so
(maos_amd/host/binio.c), which follows sys/bin.c:693-800 and
math/zfarr.c:33-110.  `read_bin` returns a numpy array for a plain array and
a nested list for a cell.  `Recorder` writes per-step gradients and pixel
intensities under MAOS's names (maos/sim_utils.c:955-998):
wfs%d_gradcl_%d.bin, wfs%d_intsnf_%d.bin / wfs%d_intsny_%d.bin.
"""
import ctypes as C
import os

import numpy as np

from . import _lib

M_DBL, M_INT64, M_CMP, M_FLT, MCC_ANY = 0x6402, 0x6403, 0x6404, 0x6408, 0x6421
_DTYPE = {M_DBL: np.float64, M_INT64: np.int64, M_CMP: np.complex128, M_FLT: np.float32}


def _check(rc, what):
    if rc < 0:
        raise IOError("%s failed (%d)" % (what, rc))


class _Reader:
    def __init__(self, fn):
        self.S = _lib.setup()
        self.f = C.c_void_p()
        _check(self.S.mxs_bin_open(C.byref(self.f), os.fsencode(fn), 0), "open " + fn)

    def read(self):
        """next element: (array or list, keywords), or None at the end"""
        magic, nx, ny = C.c_uint32(), C.c_uint64(), C.c_uint64()
        kw = C.create_string_buffer(4096)
        rc = self.S.mxs_bin_read_header(self.f, C.byref(magic), C.byref(nx), C.byref(ny), kw, len(kw))
        if rc == 1:
            return None
        _check(rc, "read header")
        m, n0, n1 = magic.value, nx.value, ny.value
        if m == MCC_ANY:
            if n0 * n1 == 0:  # an empty element (NULL in MAOS)
                return None, kw.value.decode(errors="replace")
            cells = [self.read()[0] for _ in range(n0 * n1)]
            if n1 > 1:
                cells = [cells[i * n0:(i + 1) * n0] for i in range(n1)]
            return cells, kw.value.decode(errors="replace")
        dt = _DTYPE.get(m & 0xFFFF)
        if dt is None:
            raise IOError("unsupported magic %#x" % m)
        a = np.empty(n0 * n1, dtype=dt)
        _check(self.S.mxs_bin_read_data(self.f, a.ctypes.data_as(C.c_void_p), a.nbytes), "read data")
        # x fastest: [ny, nx]
        return a.reshape(n1, n0), kw.value.decode(errors="replace")

    def close(self):
        if self.f:
            self.S.mxs_bin_close(self.f)
            self.f = None


def read_bin(fn, with_keywords=False):
    """Read a MAOS .bin / .bin.gz file: an array [ny, nx] (x fastest) or a
    (nested) list for a cell."""
    r = _Reader(fn)
    try:
        out = r.read()
    finally:
        r.close()
    if out is None:
        raise IOError("empty file " + fn)
    return out if with_keywords else out[0]


def write_bin(fn, a, keywords=None):
    """Write a 1-D/2-D array ([ny, nx], x fastest) as a MAOS .bin file."""
    a = np.ascontiguousarray(a)
    magic = {np.dtype(np.float64): M_DBL, np.dtype(np.int64): M_INT64, np.dtype(np.complex128): M_CMP,
             np.dtype(np.float32): M_FLT}[a.dtype]
    ny, nx = (1, a.size) if a.ndim == 1 else a.shape
    S = _lib.setup()
    f = C.c_void_p()
    _check(S.mxs_bin_open(C.byref(f), os.fsencode(fn), 1), "open " + fn)
    try:
        _check(S.mxs_bin_write_header(f, magic, nx, ny, keywords.encode() if keywords else None), "header")
        _check(S.mxs_bin_write_data(f, a.ctypes.data_as(C.c_void_p), a.nbytes), "data")
    finally:
        _check(S.mxs_bin_close(f), "close")


class ZfArr:
    """A MAOS zfarr: nstep elements, one pushed per step (math/zfarr.c)."""

    def __init__(self, fn, nstep, keywords=None):
        self.S = _lib.setup()
        self.z = C.c_void_p()
        _check(self.S.mxs_zfarr_open(C.byref(self.z), os.fsencode(fn), nstep,
                                     keywords.encode() if keywords else None), "zfarr open " + fn)

    def push(self, istep, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        ny, nx = (1, a.size) if a.ndim == 1 else a.shape
        _check(self.S.mxs_zfarr_push(self.z, istep, M_DBL, nx, ny, a.ctypes.data_as(C.c_void_p), 8),
               "zfarr push")

    def close(self):
        if self.z:
            _check(self.S.mxs_zfarr_close(self.z), "zfarr close")
            self.z = None


class Recorder:
    """Per-step outputs of a WfsGrad engine in MAOS's telemetry files:
    gradients (2 nsa) as wfs%d_gradcl_%d.bin and, optionally, pixel
    intensities (npix x nsa) as wfs%d_intsnf_%d.bin (noise-free engine) or
    wfs%d_intsny_%d.bin (noisy engine)."""

    def __init__(self, outdir, engine, nstep, seed=1, ints=False, noisy=False, wfs_offset=0):
        self.eng, self.ints = engine, ints
        os.makedirs(outdir, exist_ok=True)
        self.grad = []
        self.im = []
        for i in range(engine.nwfs):
            iw = wfs_offset + i
            self.grad.append(ZfArr(os.path.join(outdir, "wfs%d_gradcl_%d.bin" % (iw, seed)), nstep))
            if ints:
                name = "wfs%d_intsny_%d.bin" if noisy else "wfs%d_intsnf_%d.bin"
                self.im.append(ZfArr(os.path.join(outdir, name % (iw, seed)), nstep))

    def record(self, istep):
        """sync every WFS of the last queued step and write its outputs"""
        for i in range(self.eng.nwfs):
            g, ints = self.eng.sync(i, want_ints=self.ints)
            self.grad[i].push(istep, g)
            if self.ints:
                self.im[i].push(istep, ints)

    def close(self):
        for z in self.grad + self.im:
            z.close()
