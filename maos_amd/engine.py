"""Python handle on the GPU WFS engine (libmaos_mi355x.so).

WfsGrad follows the lifecycle of MAOS's GPU plugin (cuda/sim/gpu_sim.h:
25-55): build it from a Powfs and its WFS list (gpu_wfsgrad_init), give it
screens (gpu_atm2gpu), then queue steps and sync gradients
(gpu_wfsgrad_queue / gpu_wfsgrad_sync).  Every call goes through the C ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import MxAtmLayer, MxPowfs, MxWfs, check, dptr, lptr


class WfsSpec:
    """parms->wfs[iwfs]: direction (rad), range, photons/subap/frame."""

    def __init__(self, thetax=0.0, thetay=0.0, hs=np.inf, sigsim=1.0, wvlwts=(1.0,),
                 misregx=0.0, misregy=0.0):
        self.thetax, self.thetay, self.hs, self.sigsim = thetax, thetay, hs, sigsim
        self.wvlwts = np.ascontiguousarray(wvlwts, dtype=np.float64)
        self.misregx, self.misregy = misregx, misregy


_initialised = False


def init(device=0):
    global _initialised
    check(_lib.hip().mx_init(device))
    _initialised = True


def atm2gpu(layers, atmscale=1.0, ireal=0):
    """Upload screens (list of powfs.Layer) of realisation ireal."""
    if not _initialised:
        init(0)
    arr = (MxAtmLayer * max(1, len(layers)))()
    keep = []
    for i, L in enumerate(layers):
        p = np.ascontiguousarray(L.p, dtype=np.float64)
        keep.append(p)
        arr[i] = MxAtmLayer(dptr(p), L.nx, L.ny, L.dx, L.dx, L.ox, L.oy, L.ht, L.vx, L.vy)
    check(_lib.hip().mx_atm2gpu_real(ireal, arr, len(layers), atmscale))


def dm2gpu(dms, ireal=0):
    """Upload DM surface maps: a list of (p [ny, nx] FP64, dx, ox, oy, ht)."""
    if not _initialised:
        init(0)
    arr = (MxAtmLayer * max(1, len(dms)))()
    keep = []
    for i, (p, dx, ox, oy, ht) in enumerate(dms):
        a = np.ascontiguousarray(p, dtype=np.float64)
        keep.append(a)
        arr[i] = MxAtmLayer(dptr(a), a.shape[1], a.shape[0], dx, dx, ox, oy, ht, 0.0, 0.0)
    check(_lib.hip().mx_dm2gpu_real(ireal, arr, len(dms)))


class WfsGrad:
    """One powfs and its WFS.  With nreal > 1 the engine advances nreal
    realisations together; WFS ireal * len(wfs_list) + i is realisation
    ireal's copy of wfs_list[i] and uses the screens of atm2gpu(..., ireal)."""

    def __init__(self, powfs, wfs_list, nreal=1):
        if not _initialised:
            init(0)
        self.powfs = powfs
        self.nreal = nreal
        self.nwfs = len(wfs_list) * nreal
        self.nsa, self.npix = powfs.nsa, powfs.npix
        p = powfs
        self._keep = [p.ptsx, p.ptsy, p.amp, p.saa, p.wvl]
        llt_amp = getattr(p, "llt_amp", None)
        cfg = MxPowfs(
            nsa=p.nsa, nxsa=p.nxsa, embfac=p.cfg.embfac, notf=p.notf, pixpsax=p.pixpsax,
            pixpsay=p.pixpsay, nwvl=p.nwvl, dx=p.dx, ptsx=dptr(p.ptsx), ptsy=dptr(p.ptsy),
            amp=dptr(p.amp), saa=dptr(p.saa), saat=p.cfg.saat, areascale=p.areascale,
            wvl=dptr(p.wvl), phytype=p.cfg.phytype, sigmatch=p.cfg.sigmatch,
            cogthres=p.cfg.cogthres, cogoff=p.cfg.cogoff, pixthetax=p.pixthetax,
            pixthetay=p.pixthetay, siglev_dtrat=p.cfg.siglev * p.cfg.dtrat, saasum=p.saasum,
            llt_nx=p.llt_nx, llt_origx=getattr(p, "llt_origin", 0.0),
            llt_origy=getattr(p, "llt_origin", 0.0),
            llt_ox=p.llt.ox if p.llt else 0.0, llt_oy=p.llt.oy if p.llt else 0.0,
            llt_amp=dptr(llt_amp))
        wa = (MxWfs * len(wfs_list))()
        for i, w in enumerate(wfs_list):
            self._keep.append(w.wvlwts)
            wa[i] = MxWfs(w.thetax, w.thetay, w.hs, w.sigsim, w.misregx, w.misregy, dptr(w.wvlwts))
        h = C.c_void_p()
        check(_lib.hip().mx_wfsgrad_init_batch(C.byref(h), C.byref(cfg), wa, len(wfs_list), nreal))
        self.h = h
        for iw in range(p.nwvl):
            check(_lib.hip().mx_wfsgrad_set_dtf(
                h, iw, dptr(p.nominal[iw]), p.npix, p.notf * p.notf, lptr(p.si_pp[iw]),
                lptr(p.si_pi[iw]), dptr(p.si_px[iw])))
            if p.etf is not None:
                check(_lib.hip().mx_wfsgrad_update_etf(h, iw, dptr(p.etf[iw]), None, 0.0))

    def update_ref(self, mtche=None, i0sum=None, iwfs=-1):
        m = None if mtche is None else np.ascontiguousarray(mtche, dtype=np.float64)
        s = None if i0sum is None else np.ascontiguousarray(i0sum, dtype=np.float64)
        check(_lib.hip().mx_wfsgrad_update_ref(self.h, iwfs, dptr(m), dptr(s)))

    def update_etf(self, iwvl, etf, etf2=None, etf2wt=0.0):
        check(_lib.hip().mx_wfsgrad_update_etf(self.h, iwvl, dptr(etf), dptr(etf2), etf2wt))

    def set_etf_weight(self, etf2wt):
        """Weight of the second sodium profile's ETF for the next steps."""
        check(_lib.hip().mx_wfsgrad_set_etf_weight(self.h, float(etf2wt)))

    def queue(self, isim, dt):
        check(_lib.hip().mx_wfsgrad_queue(self.h, isim, dt))

    def wait(self):
        check(_lib.hip().mx_wfsgrad_wait(self.h))

    def sync(self, iwfs, want_ints=True):
        grad = np.zeros(2 * self.nsa)
        ints = np.zeros(self.nsa * self.npix) if want_ints else None
        check(_lib.hip().mx_wfsgrad_sync(self.h, iwfs, dptr(grad), dptr(ints)))
        return grad, (ints.reshape(self.nsa, self.npix) if want_ints else None)

    def i0_control(self, enable=True, reset=False):
        check(_lib.hip().mx_wfsgrad_i0_control(self.h, int(enable), int(reset)))

    def i0_get(self, iwfs):
        out = np.zeros(self.nsa * self.npix)
        n = C.c_long(0)
        check(_lib.hip().mx_wfsgrad_i0_get(self.h, iwfs, dptr(out), C.cast(C.byref(n), _lib.c_long_p)))
        return out.reshape(self.nsa, self.npix), n.value

    def pistat_control(self, enable=True, gtype=1, reset=False):
        """Tip/tilt-removed PSF statistics (powfs.pistatout) from the next
        step on; gtype 1 shifts by the Zernike tilt, 2 by the G-tilt."""
        check(_lib.hip().mx_wfsgrad_pistat_control(self.h, int(enable), int(gtype), int(reset)))

    def pistat_get(self, iwfs):
        """(sums [nwvl, nsa, notf, notf] with the peak at the centre, steps)."""
        p = self.powfs
        out = np.zeros(p.nwvl * self.nsa * p.notf * p.notf)
        n = C.c_long(0)
        check(_lib.hip().mx_wfsgrad_pistat_get(self.h, iwfs, dptr(out), C.cast(C.byref(n), _lib.c_long_p)))
        return out.reshape(p.nwvl, self.nsa, p.notf, p.notf), n.value

    def set_dtf_sa(self, iwvl, nominal, si_pp, si_pi, si_px):
        """Per-subaperture detector tables of one wavelength: nominal
        [nsa, notf^2] (or None), si CSC [nsa, notf^2 + 1] / [nsa, nzmax]."""
        p = self.powfs
        pp = np.ascontiguousarray(si_pp, dtype=np.int64)
        pi = np.ascontiguousarray(si_pi, dtype=np.int64)
        px = np.ascontiguousarray(si_px, dtype=np.float64)
        nom = None if nominal is None else np.ascontiguousarray(nominal, dtype=np.float64)
        check(_lib.hip().mx_wfsgrad_set_dtf_sa(self.h, iwvl, dptr(nom), p.npix, p.notf * p.notf, pi.shape[-1],
                                               lptr(pp), lptr(pi), dptr(px)))

    def psfout(self, iwfs):
        """The complex field output (powfs.psfout) of WFS iwfs at the last
        queued step: [nwvl, nsa, nout, nout] complex, nout = notf/2 + 2."""
        p = self.powfs
        nout = p.notf // 2 + 2
        out = np.zeros(2 * p.nwvl * self.nsa * nout * nout)
        check(_lib.hip().mx_wfsgrad_psfout(self.h, iwfs, dptr(out)))
        return out.view(complex).reshape(p.nwvl, self.nsa, nout, nout)

    # TT-dither statistics and the on-GPU matched-filter update
    def dither_control(self, enable=True, reset=False):
        check(_lib.hip().mx_wfsgrad_dither_control(self.h, int(enable), int(reset)))

    def set_dither(self, cs, ss, iwfs=-1):
        """Dither weights of the next queued step (dither_position)."""
        check(_lib.hip().mx_wfsgrad_set_dither(self.h, iwfs, float(cs), float(ss)))

    def dither_fold(self, s0, sg, iwfs=-1):
        check(_lib.hip().mx_wfsgrad_dither_fold(self.h, iwfs, float(s0), float(sg)))

    def set_i0model(self, iwfs, i0=None, gx=None, gy=None):
        def c(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        check(_lib.hip().mx_wfsgrad_set_i0model(self.h, iwfs, dptr(c(i0)), dptr(c(gx)), dptr(c(gy))))

    def mtch_update(self, keep, acc, iwfs=-1, bkgrnd=None, bkgrndc=0.0, rne=None, mtchcr=None,
                    sigratio=1.0):
        """Blend the folded sums into the model i0/gx/gy and rebuild the
        matched filter on the GPU; bkgrnd, rne and mtchcr default to the
        powfs's (bkgrnd per frame)."""
        cfg = self.powfs.cfg
        bk = cfg.bkgrnd * cfg.dtrat if bkgrnd is None else bkgrnd
        check(_lib.hip().mx_wfsgrad_mtch_update(self.h, iwfs, float(keep), float(acc), float(bk),
                                                 float(bkgrndc), float(cfg.rne if rne is None else rne),
                                                 int(cfg.mtchcr if mtchcr is None else mtchcr),
                                                 float(sigratio)))

    def dither_get(self, iwfs):
        n = self.nsa * self.npix
        imb, imx, imy = np.zeros(n), np.zeros(n), np.zeros(n)
        nf = C.c_long(0)
        check(_lib.hip().mx_wfsgrad_dither_get(self.h, iwfs, dptr(imb), dptr(imx), dptr(imy),
                                                C.cast(C.byref(nf), _lib.c_long_p)))
        s = (self.nsa, self.npix)
        return imb.reshape(s), imx.reshape(s), imy.reshape(s), nf.value

    def mtch_get(self, iwfs, model=True):
        """(mtche [nsa, 2*npix], i0sum, sanea [3, nsa], i0, gx, gy) of the MF
        stage; the last five only after a GPU update (model=True)."""
        n = self.nsa * self.npix
        m = np.zeros(2 * n)
        if not model:
            check(_lib.hip().mx_wfsgrad_mtch_get(self.h, iwfs, dptr(m), None, None, None, None, None))
            return m.reshape(self.nsa, 2 * self.npix)
        i0sum, sanea = np.zeros(self.nsa), np.zeros(3 * self.nsa)
        i0, gx, gy = np.zeros(n), np.zeros(n), np.zeros(n)
        check(_lib.hip().mx_wfsgrad_mtch_get(self.h, iwfs, dptr(m), dptr(i0sum), dptr(sanea), dptr(i0),
                                              dptr(gx), dptr(gy)))
        s = (self.nsa, self.npix)
        return (m.reshape(self.nsa, 2 * self.npix), i0sum, sanea.reshape(3, self.nsa), i0.reshape(s),
                gx.reshape(s), gy.reshape(s))

    def debug_opd(self, iwfs):
        out = np.zeros(self.nsa * self.powfs.nxsa ** 2)
        check(_lib.hip().mx_wfsgrad_debug_opd(self.h, iwfs, dptr(out)))
        return out

    def debug_lotf(self, iwfs, iwvl=0):
        out = np.zeros(2 * self.powfs.notf ** 2)
        check(_lib.hip().mx_wfsgrad_debug_lotf(self.h, iwfs, iwvl, dptr(out)))
        return out

    def set_variant(self, variant):
        """_lib.VARIANT_AUTO, VARIANT_FFT32, VARIANT_MFMA_BF16 or VARIANT_MFMA_FP16."""
        check(_lib.hip().mx_wfsgrad_set_variant(self.h, variant))

    def variant(self):
        v = C.c_int(0)
        check(_lib.hip().mx_wfsgrad_get_variant(self.h, C.byref(v)))
        return v.value

    def set_opdadd(self, opd, iwfs=-1):
        """Static surface OPD [nsa*nxsa^2] added to the WFS OPD (None: zero)."""
        a = None if opd is None else np.ascontiguousarray(opd, dtype=np.float64).reshape(-1)
        check(_lib.hip().mx_wfsgrad_set_opdadd(self.h, iwfs, dptr(a)))

    def set_ptt(self, ptt, iwfs=-1):
        """Piston, tip, tilt added to the WFS OPD from the next step on."""
        a = None if ptt is None else np.ascontiguousarray(ptt, dtype=np.float64)
        check(_lib.hip().mx_wfsgrad_set_ptt(self.h, iwfs, dptr(a)))

    def set_gs0(self, rowp, coli, val):
        """G-tilt operator (CSR, as Powfs.gs0 returns it) for set_geom(2)."""
        r = np.ascontiguousarray(rowp, dtype=np.int64)
        c = np.ascontiguousarray(coli, dtype=np.int64)
        v = np.ascontiguousarray(val, dtype=np.float64)
        check(_lib.hip().mx_wfsgrad_set_gs0(self.h, lptr(r), lptr(c), dptr(v), len(v)))

    def set_geom(self, mode=1):
        """1: geometric Zernike-tilt gradients from the OPD; 2: G-tilt (GS0,
        set_gs0); 0: physical optics."""
        check(_lib.hip().mx_wfsgrad_set_geom(self.h, mode))

    def set_noise(self, enable=True, rne=0.0, bkgrnd=0.0, bkgrndc=0.0, seed=1):
        """Photon and read noise on every later step's intensities."""
        check(_lib.hip().mx_wfsgrad_set_noise(self.h, int(enable), rne, bkgrnd, bkgrndc, seed))

    def set_dtrat(self, dtrat):
        """Exposures of dtrat steps: intensities and gradients at the last
        step of each (mx_wfsgrad_set_dtrat)."""
        check(_lib.hip().mx_wfsgrad_set_dtrat(self.h, int(dtrat)))

    def set_graph(self, enable=True):
        """Launch each step's kernels as one HIP graph."""
        check(_lib.hip().mx_wfsgrad_set_graph(self.h, int(enable)))

    def variant_ok(self, variant):
        """True if this configuration runs `variant`; the current variant is kept."""
        cur = self.variant()
        if _lib.hip().mx_wfsgrad_set_variant(self.h, variant) != 0:
            return False
        self.set_variant(cur)
        return True

    def autotune(self, isim, dt, nrep=10):
        """Time each candidate (imaging form, ray-trace tile, large-field
        block size) and keep the fastest; returns a list of dicts, the chosen
        one with best=True."""
        buf = C.create_string_buffer(4096)
        check(_lib.hip().mx_wfsgrad_autotune(self.h, isim, dt, nrep, buf, len(buf)))
        out = []
        for line in buf.value.decode().splitlines():
            best = line.startswith("*")
            kv = dict(t.split("=") for t in line[1:].split())
            out.append({"form": kv["form"], "opd_tile": int(kv["opd_tile"]), "threads": int(kv["threads"]),
                        "step_us": float(kv["step_us"]), "best": best})
        return out

    def timing_enable(self, on=True):
        check(_lib.hip().mx_wfsgrad_timing_enable(self.h, int(on)))

    def timing(self):
        """(timed steps, ms uplink OTF, ms ray trace, ms imaging), summed."""
        n = C.c_long(0)
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        check(_lib.hip().mx_wfsgrad_timing(self.h, C.cast(C.byref(n), _lib.c_long_p),
                                           C.cast(C.byref(a), _lib.c_double_p),
                                           C.cast(C.byref(b), _lib.c_double_p),
                                           C.cast(C.byref(c), _lib.c_double_p)))
        return n.value, a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            _lib.hip().mx_wfsgrad_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WfsGroup:
    """Several WfsGrad engines stepped by one graph launch per step, each
    engine a concurrent branch (mx_wfsgroup_*).  sync/wait stay per engine."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        g = C.c_void_p()
        check(_lib.hip().mx_wfsgroup_create(C.byref(g), arr, len(self.engines)))
        self.g = g

    def queue(self, isim, dt):
        check(_lib.hip().mx_wfsgroup_queue(self.g, isim, dt))

    def wait(self):
        for e in self.engines:
            e.wait()

    def close(self):
        if getattr(self, "g", None):
            _lib.hip().mx_wfsgroup_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def loc_embed(locx, locy, dx, nembed=0, oversize=2.0, fftpad=True):
    """Embedding of aperture points in a square FFT grid (loc_create_embed,
    math/loc.c:147-179): (embed indices ix + iy * nembed, nembed)."""
    locx = np.ascontiguousarray(locx, dtype=np.float64)
    locy = np.ascontiguousarray(locy, dtype=np.float64)
    emb = np.zeros(len(locx), dtype=np.int64)
    n = C.c_long(int(nembed))
    r = _lib.setup().mxs_loc_create_embed(lptr(emb), C.cast(C.byref(n), _lib.c_long_p), len(locx), dptr(locx),
                                           dptr(locy), dx, dx, oversize, int(fftpad))
    if r < 0:
        raise ValueError("mxs_loc_create_embed failed (%d)" % r)
    return emb, int(n.value)


class Evl:
    """Science-field evaluation on the GPU (mx_evl_*): the OPD of each
    direction on the aperture points (screens and DM surfaces already
    uploaded) and the accumulated long-exposure PSF per (direction,
    wavelength), as MAOS's perfevl PSF path (gpu_perfevl_init_sim,
    psfcomp_r, gpu_perfevl_save)."""

    def __init__(self, amp, wvl, nembed, psfsize, embed, locx=None, locy=None, dirs=((0.0, 0.0, 0.0),)):
        if not _initialised:
            init(0)
        self.amp = np.ascontiguousarray(amp, dtype=np.float64)
        self.nloc = len(self.amp)
        self.ndir = len(dirs)
        self.nwvl = len(wvl)
        self.psfsize = [int(p) for p in psfsize]
        d = np.asarray(dirs, dtype=np.float64).reshape(self.ndir, 3)
        keep = [np.ascontiguousarray(d[:, i]) for i in range(3)]
        w = np.ascontiguousarray(wvl, dtype=np.float64)
        ne = np.ascontiguousarray(nembed, dtype=np.int64)
        ps = np.ascontiguousarray(psfsize, dtype=np.int64)
        em = np.ascontiguousarray(np.asarray(embed, dtype=np.int64).reshape(self.nwvl, self.nloc))
        lx = np.ascontiguousarray(locx, dtype=np.float64) if locx is not None else None
        ly = np.ascontiguousarray(locy, dtype=np.float64) if locy is not None else None
        h = C.c_void_p()
        check(_lib.hip().mx_evl_init(C.byref(h), self.nloc, dptr(lx) if lx is not None else None,
                                     dptr(ly) if ly is not None else None, dptr(self.amp), self.ndir,
                                     dptr(keep[0]), dptr(keep[1]), dptr(keep[2]), self.nwvl, dptr(w), lptr(ne),
                                     lptr(ps), lptr(em)))
        self.h = h

    def psf_add(self, opd):
        """Add the PSFs of opd [ndir, nloc] (metres)."""
        o = np.ascontiguousarray(opd, dtype=np.float64).reshape(self.ndir, self.nloc)
        check(_lib.hip().mx_evl_psf_add(self.h, dptr(o)))

    def psf_add_device(self, ptr):
        """Add the PSFs of an FP32 [ndir, nloc] OPD at device address ptr."""
        check(_lib.hip().mx_evl_psf_add_device(self.h, C.c_void_p(int(ptr))))

    def trace(self, isim, dt, add_psf=True):
        check(_lib.hip().mx_evl_trace(self.h, int(isim), float(dt), int(add_psf)))

    def set_real(self, ireal):
        """Trace the screens and DM surfaces of realisation ireal."""
        check(_lib.hip().mx_evl_set_real(self.h, int(ireal)))

    def opd(self):
        out = np.zeros((self.ndir, self.nloc))
        check(_lib.hip().mx_evl_opd_get(self.h, dptr(out)))
        return out

    def wfe(self):
        """(rms [ndir, 3]: piston removed, tip/tilt, p/t/t removed; coeff
        [ndir, 3]: piston, tip, tilt) of the last OPD (mx_evl_wfe)."""
        rms = np.zeros((self.ndir, 3))
        coeff = np.zeros((self.ndir, 3))
        check(_lib.hip().mx_evl_wfe(self.h, dptr(rms), dptr(coeff)))
        return rms, coeff

    def wfe_noalloc(self, ptr):
        """wfe() of an FP32 [ndir, nloc] OPD at device address ptr, with no
        allocation (mx_evl_wfe_noalloc)."""
        rms = np.zeros((self.ndir, 3))
        coeff = np.zeros((self.ndir, 3))
        check(_lib.hip().mx_evl_wfe_noalloc(self.h, C.c_void_p(int(ptr)), dptr(rms), dptr(coeff)))
        return rms, coeff

    def psf(self, idir=0, iwvl=0):
        """(mean PSF [psfsize, psfsize] with x fastest, number of PSFs added)."""
        P = self.psfsize[iwvl]
        out = np.zeros((P, P))
        n = C.c_long(0)
        check(_lib.hip().mx_evl_psf_get(self.h, idir, iwvl, dptr(out), C.cast(C.byref(n), _lib.c_long_p)))
        return out, n.value

    def reset(self):
        check(_lib.hip().mx_evl_psf_reset(self.h))

    def wait(self):
        check(_lib.hip().mx_evl_wait(self.h))

    def timing_enable(self, on=True):
        check(_lib.hip().mx_evl_timing_enable(self.h, int(on)))

    def timing(self):
        """(timed traces, ms trace, ms row FFT pass, ms column FFT pass), summed."""
        n = C.c_long(0)
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        check(_lib.hip().mx_evl_timing(self.h, C.cast(C.byref(n), _lib.c_long_p),
                                       C.cast(C.byref(a), _lib.c_double_p), C.cast(C.byref(b), _lib.c_double_p),
                                       C.cast(C.byref(c), _lib.c_double_p)))
        return n.value, a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            _lib.hip().mx_evl_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PywfsEngine:
    """A pyramid WFS on the GPU (mx_pywfs_*) built from a maos_amd.pywfs.Pywfs:
    mx_pywfs_init with its tables, then set_calib with its calibration."""

    def __init__(self, pw):
        if not _initialised:
            init(0)
        c = pw.cfg
        self.pw = pw
        self.nsa, self.nside, self.ng = pw.nsa, c.nside, pw.ng
        self._keep = [np.ascontiguousarray(a) for a in (pw.locx, pw.locy, pw.amp, pw.embed, pw.wvl, pw.wvlwts,
                                                         pw.mask.ravel(), pw.nominal, pw.salocx, pw.salocy,
                                                         pw.sioff)]
        k = self._keep
        cfg = _lib.MxPywfsCfg(nloc=pw.nloc, locx=dptr(k[0]), locy=dptr(k[1]), amp=dptr(k[2]), embed=lptr(k[3]),
                              nembed=pw.N, dx=c.dx, nwvl=pw.nwvl, wvl=dptr(k[4]), wvlwts=dptr(k[5]),
                              mask=dptr(k[6]), nominal=dptr(k[7]), nsa=pw.nsa, salocx=dptr(k[8]),
                              salocy=dptr(k[9]), nside=c.nside, sioff=dptr(k[10]), modulate=pw.modulate,
                              modulpos=c.modulpos, modulring=c.modulring, sigmatch=c.sigmatch, raw=c.raw,
                              siglev=c.siglev, thetax=c.thetax, thetay=c.thetay, hs=c.hs)
        h = C.c_void_p()
        check(_lib.hip().mx_pywfs_init(C.byref(h), C.byref(cfg)))
        self.h = h
        saa = np.ascontiguousarray(pw.saa, dtype=np.float64)
        goff = None if pw.gradoff is None else np.ascontiguousarray(pw.gradoff, dtype=np.float64)
        check(_lib.hip().mx_pywfs_set_calib(h, pw.gain, dptr(saa), dptr(goff)))

    def image(self, opd, siglev=None):
        """(ints [nside, nsa], grad [ng nsa]) of a pupil OPD."""
        o = np.ascontiguousarray(opd, dtype=np.float64)
        check(_lib.hip().mx_pywfs_ints_opd(self.h, dptr(o), self.pw.cfg.siglev if siglev is None else siglev))
        return self.sync()

    def queue(self, isim, dt, siglev=None):
        check(_lib.hip().mx_pywfs_queue(self.h, isim, dt, self.pw.cfg.siglev if siglev is None else siglev))

    def sync(self):
        g = np.zeros(self.ng * self.nsa)
        i = np.zeros(self.nside * self.nsa)
        check(_lib.hip().mx_pywfs_sync(self.h, dptr(g), dptr(i)))
        return i.reshape(self.nside, self.nsa), g

    def opd(self):
        o = np.zeros(self.pw.nloc)
        check(_lib.hip().mx_pywfs_opd_get(self.h, dptr(o)))
        return o

    def timing_enable(self, enable=True):
        check(_lib.hip().mx_pywfs_timing_enable(self.h, int(enable)))

    def timing(self):
        n, ms = C.c_long(0), C.c_double(0)
        check(_lib.hip().mx_pywfs_timing(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def wait(self):
        check(_lib.hip().mx_pywfs_wait(self.h))

    def close(self):
        if self.h:
            check(_lib.hip().mx_pywfs_free(self.h))
            self.h = None
