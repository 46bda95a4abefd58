"""Host-side setup of a Shack-Hartmann powfs, mirroring MAOS's setup order.

In a MAOS run these tables come from setup_shwfs_geom / setup_shwfs_phy
(maos/powfs.c:255-462, 724-1334).  This module drives the FP64 C builders in
libmaos_setup.so in the same order so the GPU path can run standalone:

  geometry   saloc, pts, amp (annulus), saa, sa_reduce   maos/powfs.c:255-462
  notf       nextfftsize(max(nxsa*embfac, FoV/dtheta))   maos/powfs.c:809-844
  DTF        nominal (treated) and the si sampler        lib/mkdtf.c:89-184
  LGS        srot/srsa, sodium smoothing, ETF            maos/powfs.c:744-758,
                                                         1077-1177, lib/mkdtf.c
  LLT        Gaussian uplink pupil                       maos/powfs.c:1182-1242

Parameter names follow MAOS's configuration keys (powfs.*, llt.*, aper.*).
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import dptr, lptr

AS2RAD = math.pi / 180.0 / 3600.0
PTYPE_MF = 1
PTYPE_COG = 2


@dataclass
class LltCfg:
    d: float = 0.4          # llt.d
    widthp: float = 0.6     # llt.widthp
    ox: float = 0.0         # llt.ox
    oy: float = 0.0         # llt.oy
    na_smooth: int = 1      # llt.na_smooth
    na_interp: int = 1      # llt.na_interp: 0 sums the profile's phasors directly


@dataclass
class SodiumCfg:
    """Gaussian sodium profile sampled on [hmin, hmax] with spacing dh, or a
    MAOS profile file (llt.fnprof, e.g. NapMean.bin): heights and column
    density as a 2 x nh .bin array."""
    center: float = 90e3
    fwhm: float = 10e3
    hmin: float = 70e3
    hmax: float = 110e3
    dh: float = 20.0
    fnprof: Optional[str] = None

    def table(self):
        if self.fnprof:
            from .telemetry import read_bin
            a = np.asarray(read_bin(self.fnprof), dtype=np.float64)
            if a.ndim != 2 or a.shape[0] != 2:
                raise ValueError("%s: expected a 2 x nh profile" % self.fnprof)
            return np.ascontiguousarray(a.reshape(-1))  # nh x 2, column-major
        return self._gaussian()

    def _gaussian(self):
        h = np.arange(self.hmin, self.hmax + 0.5 * self.dh, self.dh)
        sig = self.fwhm / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        na = np.exp(-0.5 * ((h - self.center) / sig) ** 2)
        na /= na.sum()
        return np.ascontiguousarray(np.stack([h, na], axis=1).T.reshape(-1))  # nh x 2, column-major


@dataclass
class PowfsCfg:
    order: int = 16
    dsa: float = 0.5
    dx: float = 1.0 / 32
    saat: float = 0.3
    hs: float = math.inf
    embfac: int = 2
    wvl: List[float] = field(default_factory=lambda: [0.7e-6])
    wvlwts: List[float] = field(default_factory=lambda: [1.0])
    pixpsa: int = 8
    pixtheta: float = 0.5     # arcsec
    pixblur: float = 0.3
    notf: int = 0
    phytype: int = PTYPE_COG
    sigmatch: int = 1
    cogthres: float = 0.0
    cogoff: float = 0.0
    mtchcr: int = 1
    rne: float = 3.0
    bkgrnd: float = 0.0
    siglev: float = 1000.0    # at sim.dtref
    dtrat: int = 1
    llt: Optional[LltCfg] = None
    sodium: Optional[SodiumCfg] = None


@dataclass
class AperCfg:
    d: float = 8.0
    din: float = 1.2


@dataclass
class SimCfg:
    dt: float = 1.0 / 800
    dtref: float = 1.0 / 600
    za: float = 0.0
    htel: float = 0.0


class Powfs:
    """Tables of one SH powfs (numpy, FP64, MAOS layouts)."""

    def __init__(self, aper: AperCfg, cfg: PowfsCfg, sim: SimCfg):
        S = _lib.setup()
        self.cfg, self.aper, self.sim = cfg, aper, sim
        nxsa = 2 * int(round(0.5 * cfg.dsa / cfg.dx))
        dx = cfg.dsa / nxsa
        self.nxsa, self.dx, self.dsa = nxsa, dx, cfg.dsa
        npsa = nxsa * nxsa
        # saloc on the padded grid (maos/powfs.c:302-340)
        nmax = (cfg.order + 8) ** 2
        lx = np.zeros(nmax)
        ly = np.zeros(nmax)
        n = S.mxs_saloc(cfg.order, cfg.dsa, aper.din, 0.0, dptr(lx), dptr(ly), nmax)
        lx, ly = lx[:n].copy(), ly[:n].copy()
        px, py = lx + dx * 0.5, ly + dx * 0.5
        # amplitude: annulus on every OPD point (mkamp without aper map)
        g = np.arange(nxsa) * dx
        shape = (n, nxsa, nxsa)  # [isa][iy][ix], x fastest (pts2loc order)
        locx = np.ascontiguousarray(np.broadcast_to(px[:, None, None] + g[None, None, :], shape).reshape(-1))
        locy = np.ascontiguousarray(np.broadcast_to(py[:, None, None] + g[None, :, None], shape).reshape(-1))
        amp = np.zeros(n * npsa)
        S.mxs_loc_circle_add(dptr(amp), dptr(locx), dptr(locy), n * npsa, dx, dx, 0.0, 0.0,
                             aper.d * 0.5, aper.din * 0.5, 1.0)
        self.areascale = 1.0 if cfg.order > 2 else 4.0 / math.pi
        saa = np.zeros(n)
        S.mxs_saa(dptr(amp), n, npsa, self.areascale, dptr(saa))
        if cfg.saat > 0:
            n = S.mxs_sa_reduce(dptr(lx), dptr(ly), dptr(px), dptr(py), dptr(amp), dptr(saa), n,
                                npsa, cfg.dsa, cfg.saat)
        self.nsa = n
        self.salocx, self.salocy = lx[:n].copy(), ly[:n].copy()
        self.ptsx, self.ptsy = px[:n].copy(), py[:n].copy()
        self.amp = amp[:n * npsa].copy()
        self.saa = saa[:n].copy()
        self.saasum = float(self.saa.sum())
        # pixels and notf (maos/powfs.c:800-844)
        self.pixpsax = self.pixpsay = cfg.pixpsa
        self.npix = cfg.pixpsa * cfg.pixpsa
        self.pixthetax = self.pixthetay = cfg.pixtheta * AS2RAD
        self.nwvl = len(cfg.wvl)
        self.wvl = np.array(cfg.wvl, dtype=np.float64)
        self.wvlwts = np.array(cfg.wvlwts, dtype=np.float64)
        if cfg.notf:
            notf = (cfg.notf + 1) // 2 * 2
        else:
            safov = max(self.pixpsax * self.pixthetax, self.pixpsay * self.pixthetay)
            dtheta0 = self.wvl[0] / (cfg.dsa * cfg.embfac)
            notf = S.mxs_nextfftsize(max(nxsa * cfg.embfac, int(math.ceil(safov / dtheta0))))
        self.notf = int(notf)
        N = self.notf
        # DTF: nominal and si per wavelength (lib/mkdtf.c:89-184)
        self.dtheta = self.wvl / (cfg.dsa * cfg.embfac)
        self.nominal = np.zeros((self.nwvl, N * N))
        self.si_nzmax = 4 * self.npix
        self.si_pp = np.zeros((self.nwvl, N * N + 1), dtype=np.int64)
        self.si_pi = np.zeros((self.nwvl, self.si_nzmax), dtype=np.int64)
        self.si_px = np.zeros((self.nwvl, self.si_nzmax))
        for iw in range(self.nwvl):
            dth = self.dtheta[iw]
            nom = np.zeros(N * N)
            S.mxs_dtf_nominal(dptr(nom), N, N, self.pixthetax / dth, self.pixthetay / dth, 0.0,
                              cfg.pixblur)
            self.nominal[iw] = nom
            pp = np.zeros(N * N + 1, dtype=np.int64)
            pi = np.zeros(self.si_nzmax, dtype=np.int64)
            pxs = np.zeros(self.si_nzmax)
            S.mxs_si(N, N, dth, self.pixpsax, self.pixpsay, self.pixthetax, self.pixthetay,
                     0.0, 0.0, 0.0, lptr(pp), lptr(pi), dptr(pxs))
            self.si_pp[iw], self.si_pi[iw], self.si_px[iw] = pp, pi, pxs
        # LGS: LLT, polar coordinates, sodium, ETF
        self.llt = cfg.llt
        self.etf = None
        self.llt_nx = 0
        if cfg.llt is not None:
            self.srot = np.zeros(n)
            self.srsa = np.zeros(n)
            S.mxs_srot_srsa(dptr(self.salocx), dptr(self.salocy), n, cfg.dsa, cfg.llt.ox,
                            cfg.llt.oy, dptr(self.srot), dptr(self.srsa))
            prof = (cfg.sodium or SodiumCfg()).table()
            nh = prof.size // 2
            if cfg.llt.na_smooth:
                dxnew = cfg.hs ** 2 / self.srsa.max() * self.dtheta.min()
                out = np.zeros(2 * (nh + 1))
                m = S.mxs_smooth(dptr(prof), nh, dxnew, dptr(out), nh + 1)
                prof = out[:2 * m].copy()  # m x 2, column-major
            self.sodium = prof
            nh = prof.size // 2
            self.etf = np.zeros((self.nwvl, n * N * N * 2))
            for iw in range(self.nwvl):
                rc = S.mxs_mketf(dptr(self.etf[iw]), None, dptr(self.nominal[iw]), N,
                                 self.dtheta[iw], dptr(prof), nh, dptr(self.srot),
                                 dptr(self.srsa), n, cfg.hs, sim.htel, sim.za,
                                 0 if cfg.llt.na_interp else 1)
                if rc != 0:
                    raise RuntimeError("mketf failed (%d)" % rc)
            amp = np.zeros(4096)
            org = C.c_double(0)
            nlx = S.mxs_llt_amp(cfg.llt.d, cfg.llt.widthp, dx, cfg.dsa, dptr(amp), 4096,
                                C.byref(org))
            self.llt_nx = int(nlx)
            self.llt_amp = amp[:nlx * nlx].copy()
            self.llt_origin = org.value
        self.sigsim = cfg.siglev * (sim.dt / sim.dtref) * math.cos(sim.za)

    # -------------------------------------------------------------- i0 model
    def i0_model(self, r0, L0, shift2center=None, i0scale=False):
        """Model i0, gx, gy [nsa, npix] from the turbulence-averaged
        short-exposure OTF, as MAOS builds them for the matched filter at
        setup (maos/powfs.c:1545-1622): genseotf of the subapertures and of
        the LLT pupil (piston/tip/tilt removed, sampled at nxsa*embfac),
        gensepsf with the uplink OTF, gensei with the ETF (nominal fused) or
        the nominal.  r0, L0: the powfs's (MAOS defaults them to the
        atmosphere's).  shift2center (powfs.mtchstc) defaults to on without
        an LLT and off with one, as MAOS's SH NGS and LGS configurations do.
        The images are in the units of the simulated ones (wvlwts * sigsim
        photons)."""
        S = _lib.setup()
        cfg = self.cfg
        N, nsa, npix = self.notf, self.nsa, self.npix
        if shift2center is None:
            shift2center = self.llt is None
        npsf = self.nxsa * cfg.embfac
        thres = 1.0 - 1e-10
        i0 = np.zeros(nsa * npix)
        gx = np.zeros(nsa * npix)
        gy = np.zeros(nsa * npix)
        for iw in range(self.nwvl):
            otf = np.zeros(nsa * npsf * npsf * 2)
            rc = S.mxs_genotf(dptr(otf), self.nxsa, self.dx, dptr(self.amp), dptr(self.saa), nsa, thres,
                              self.wvl[iw], None, r0, L0, npsf, 1)
            if rc != 0:
                raise RuntimeError("genotf failed (%d)" % rc)
            lotf, nl = None, 0
            if self.llt is not None:
                nl = self.llt_nx * cfg.embfac
                lotf = np.zeros(nl * nl * 2)
                rc = S.mxs_genotf(dptr(lotf), self.llt_nx, self.dx, dptr(self.llt_amp), None, 1, thres,
                                  self.wvl[iw], None, r0, L0, nl, 1)
                if rc != 0:
                    raise RuntimeError("genotf (LLT) failed (%d)" % rc)
            psf = np.zeros(nsa * N * N)
            rc = S.mxs_gensepsf(dptr(psf), dptr(otf), npsf, nsa, dptr(lotf), nl, dptr(self.saa), N)
            if rc != 0:
                raise RuntimeError("gensepsf failed (%d)" % rc)
            etf = None if self.etf is None else self.etf[iw]
            rc = S.mxs_gensei(dptr(i0), dptr(gx), dptr(gy), dptr(psf), nsa, N, dptr(etf),
                              dptr(self.nominal[iw]), npix, lptr(self.si_pp[iw]), lptr(self.si_pi[iw]),
                              dptr(self.si_px[iw]), self.dtheta[iw], self.wvlwts[iw] * self.sigsim,
                              int(shift2center), int(i0scale), dptr(self.saa))
            if rc != 0:
                raise RuntimeError("gensei failed (%d)" % rc)
        return i0.reshape(nsa, npix), gx.reshape(nsa, npix), gy.reshape(nsa, npix)

    # -------------------------------------------------------------- G-tilt
    def dtf_sa(self, pixoffx, pixoffy, pixrot=None):
        """Per-subaperture detector tables (mkdtf's loop over subapertures,
        lib/mkdtf.c:140-161): pixel offsets pixoffx, pixoffy (pixels) and
        pixel rotations pixrot (radians) per subaperture.  Returns nominal
        [nwvl, nsa, N^2] and si CSC pp [nwvl, nsa, N^2+1], pi and px
        [nwvl, nsa, 4 npix]."""
        S = _lib.setup()
        N, nsa = self.notf, self.nsa
        rot = np.zeros(nsa) if pixrot is None else np.asarray(pixrot, dtype=float)
        nz = 4 * self.npix
        nominal = np.zeros((self.nwvl, nsa, N * N))
        pp = np.zeros((self.nwvl, nsa, N * N + 1), dtype=np.int64)
        pi = np.zeros((self.nwvl, nsa, nz), dtype=np.int64)
        px = np.zeros((self.nwvl, nsa, nz))
        for iw in range(self.nwvl):
            dth = self.dtheta[iw]
            for isa in range(nsa):
                nom = np.zeros(N * N)
                S.mxs_dtf_nominal(dptr(nom), N, N, self.pixthetax / dth, self.pixthetay / dth, float(rot[isa]),
                                  self.cfg.pixblur)
                nominal[iw, isa] = nom
                a, b, c = (np.zeros(N * N + 1, dtype=np.int64), np.zeros(nz, dtype=np.int64), np.zeros(nz))
                S.mxs_si(N, N, dth, self.pixpsax, self.pixpsay, self.pixthetax, self.pixthetay,
                         float(pixoffx[isa]), float(pixoffy[isa]), float(rot[isa]), lptr(a), lptr(b), dptr(c))
                pp[iw, isa], pi[iw, isa], px[iw, isa] = a, b, c
        return nominal, pp, pi, px

    def gs0(self):
        """Average-gradient operator GS0 (mkg, lib/mkg.c:108-468) as CSR
        (rowp, coli, val): 2*nsa rows (x then y) over the OPD points."""
        S = _lib.setup()
        nz = 2 * self.nsa * (self.nxsa + 2) ** 2
        rowp = np.zeros(2 * self.nsa + 1, dtype=np.int64)
        coli = np.zeros(nz, dtype=np.int64)
        val = np.zeros(nz)
        n = S.mxs_mkg(lptr(rowp), lptr(coli), dptr(val), nz, dptr(self.ptsx), dptr(self.ptsy), self.nsa,
                      self.nxsa, self.dx, dptr(self.amp), dptr(self.saa), self.cfg.saat)
        if n < 0:
            raise RuntimeError("mkg failed (%d)" % n)
        return rowp, coli[:n].copy(), val[:n].copy()

    # -------------------------------------------------------------- MF
    def mtch(self, i0, gx=None, gy=None):
        """Matched filter from averaged images i0 [nsa, npix] (mtch_cell,
        lib/mtch.c:223-335); gx/gy None uses the FFT derivative."""
        S = _lib.setup()
        cfg = self.cfg
        i0 = np.ascontiguousarray(i0, dtype=np.float64).reshape(-1)
        mtche = np.zeros(self.nsa * 2 * self.npix)
        sanea = np.zeros(self.nsa * 3)
        i0sum = np.zeros(self.nsa)
        sumsum = C.c_double(0)
        rc = S.mxs_mtch(dptr(mtche), dptr(sanea), dptr(i0sum), C.cast(C.byref(sumsum), _lib.c_double_p),
                        dptr(i0), dptr(None if gx is None else np.ascontiguousarray(gx).reshape(-1)),
                        dptr(None if gy is None else np.ascontiguousarray(gy).reshape(-1)),
                        self.nsa, self.pixpsax, self.pixpsay, cfg.bkgrnd * cfg.dtrat, 0.0,
                        cfg.rne, self.pixthetax, self.pixthetay, cfg.mtchcr, 1.0)
        if rc != 0:
            raise RuntimeError("mtch failed")
        return mtche, i0sum, sanea.reshape(3, self.nsa)


@dataclass
class AtmCfg:
    r0: float = 0.186
    L0: float = 30.0
    ht: List[float] = field(default_factory=lambda: [0.0])
    wt: List[float] = field(default_factory=lambda: [1.0])
    ws: List[float] = field(default_factory=lambda: [10.0])
    wddeg: List[float] = field(default_factory=lambda: [0.0])
    dx: float = 1.0 / 32
    nx: int = 2048
    seeds: List[int] = field(default_factory=lambda: [1])


class Layer:
    def __init__(self, p, dx, ht, vx, vy):
        self.p = p  # [ny, nx] float64, x fastest
        self.ny, self.nx = p.shape
        self.dx = dx
        self.ox = -(self.nx // 2) * dx  # mapnew origin (math/map.c:51-63)
        self.oy = -(self.ny // 2) * dx
        self.ht, self.vx, self.vy = ht, vx, vy


def genatm(cfg: AtmCfg):
    """Seeded von Karman screens, one seed per layer (genscreen,
    lib/turbulence.c:70-171), with frozen-flow winds (maos/sim_utils.c:196-202)."""
    S = _lib.setup()
    layers = []
    for i, ht in enumerate(cfg.ht):
        scr = np.zeros(cfg.nx * cfg.nx)
        wt = np.array([cfg.wt[i]], dtype=np.float64)
        S.mxs_genscreen(dptr(scr), cfg.nx, cfg.nx, cfg.dx, cfg.r0, cfg.L0, dptr(wt), 1,
                        int(cfg.seeds[i]))
        ang = cfg.wddeg[i] * math.pi / 180
        layers.append(Layer(scr.reshape(cfg.nx, cfg.nx), cfg.dx, ht, math.cos(ang) * cfg.ws[i],
                            math.sin(ang) * cfg.ws[i]))
    return layers
