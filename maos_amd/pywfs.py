"""Host-side setup of a pyramid WFS, in the order pywfs_new builds it
(lib/pywfs.c:110-370, called from setup_pywfs, maos/powfs.c:1751-1805):

  pupil     points of the aperture grid at dx with amplitude > EPS       maos/powfs.c:1755-1758
  locfft    embedding in an FFT grid oversized 3x (PYWFS_PSIZE)           lib/pywfs.c:160
  pyramid   facet phase mask per wavelength, clear aperture, field stop  lib/pywfs.c:170-232
  nominal   detector transfer function of the pupil images              lib/pywfs.c:236-262
  saloc     order x order detector subapertures, pupil-image offsets     lib/pywfs.c:34-107, 281-296
  saa       flux per subaperture of a flat wavefront, saat pruning       lib/pywfs.c:298-325
  gradoff   gradients of a flat wavefront                                lib/pywfs.c:331-344
  gain      inverse optical gain from a 5 mas tip and tilt               lib/pywfs.c:345-365, 524-588

The tables feed the GPU engine (engine.Pywfs) and, in tests, the FP64
oracle.  The calibration steps run the imaging itself, so they take an
imaging function (the GPU engine's, or the oracle's).
"""
import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib
from ._lib import dptr

AS2RAD = math.pi / 180.0 / 3600.0
EPS = 1e-15


@dataclass
class PywfsCfg:
    """pywfs_cfg_t keys (lib/pywfs.h) with MAOS's defaults
    (config/maos/pywfs.conf, powfs_pyngs.conf) where they apply."""
    D: float = 30.0            # aper.d
    din: float = 3.6           # aper.din
    order: int = 60            # detector subapertures across
    dx: float = 1.0 / 32       # pupil sampling
    wvl: List[float] = field(default_factory=lambda: [0.7e-6])
    wvlwts: List[float] = field(default_factory=lambda: [1.0])
    nside: int = 4
    modulate: float = 5.0      # in max(wvl) / D
    modulpos: int = 32
    modulring: int = 1
    sigmatch: int = 2
    raw: int = 0
    siglev: float = 1000.0
    pixblur: float = 0.3
    fieldstop: float = 0.0     # arcsec, 0: none
    flate: float = 0.0         # mas
    flatv: float = 0.0         # mas
    saat: float = 0.5          # powfs.saat
    psize: float = 3.0         # FFT grid oversize (PYWFS_PSIZE)
    thetax: float = 0.0
    thetay: float = 0.0
    hs: float = math.inf


class Pywfs:
    def __init__(self, cfg: PywfsCfg):
        from .engine import loc_embed
        S = _lib.setup()
        self.cfg = cfg
        c = cfg
        # pupil: grid points of the square map of MAOS's mkannloc with the
        # anti-aliased annulus as amplitude, points of amplitude <= EPS dropped
        nx = max(int(c.D / c.dx + 1), 1)
        g = (np.arange(nx) - nx // 2) * c.dx
        X, Y = np.meshgrid(g, g)
        locx, locy = X.ravel().copy(), Y.ravel().copy()
        amp = np.zeros(locx.size)
        S.mxs_loc_circle_add(dptr(amp), dptr(locx), dptr(locy), locx.size, c.dx, c.dx, 0.0, 0.0, c.D / 2,
                             c.din / 2, 1.0)
        keep = amp > EPS
        self.locx, self.locy, self.amp = locx[keep].copy(), locy[keep].copy(), amp[keep].copy()
        self.nloc = self.locx.size
        self.embed, self.N = loc_embed(self.locx, self.locy, c.dx, oversize=c.psize)
        N = self.N
        self.wvl = np.array(c.wvl, dtype=np.float64)
        self.wvlwts = np.array(c.wvlwts, dtype=np.float64)
        self.nwvl = len(self.wvl)
        self.modulate = c.modulate * self.wvl.max() / c.D
        # pyramid masks per wavelength
        self.mask = np.zeros((self.nwvl, 2 * N * N))
        for iw, w in enumerate(self.wvl):
            dtheta = w / (c.dx * N)
            nstop = N
            if c.nside == 1:
                nstop = int(round(1.063 * c.dx * N / c.D))
            elif c.fieldstop > 0:
                nstop = int(math.ceil(c.fieldstop * AS2RAD / dtheta * 0.5) * 2)
            rc = S.mxs_pywfs_pyramid(dptr(self.mask[iw]), N, c.nside, nstop, c.flate * 1e-3 * AS2RAD / dtheta / 2,
                                     c.flatv * 1e-3 * AS2RAD / dtheta / 2)
            if rc:
                raise ValueError("mxs_pywfs_pyramid failed")
        self.dsa = c.D / c.order
        self.nominal = np.zeros(N * N)
        S.mxs_pywfs_nominal(dptr(self.nominal), N, self.dsa, c.dx, c.pixblur)
        sx, sy = np.zeros(c.order * c.order), np.zeros(c.order * c.order)
        S.mxs_pywfs_saloc(dptr(sx), dptr(sy), sx.size, c.order, self.dsa)
        self.salocx, self.salocy = sx, sy
        self.sioff = np.zeros(2 * c.nside)
        S.mxs_pywfs_sioff(dptr(self.sioff), c.nside)
        self.ng = c.nside if (c.raw or c.nside < 3) else 2
        self.gain = 1.0
        self.saa = np.ones(self.nsa)
        self.gradoff = None

    @property
    def nsa(self):
        return self.salocx.size

    def calibrate(self, make):
        """pywfs_new's calibration (lib/pywfs.c:298-368).  make(pywfs)
        returns an imaging function f(opd) -> (ints [nside, nsa], grad) for
        the current tables (saloc, gain, saa, gradoff)."""
        c = self.cfg
        flat = np.zeros(self.nloc)
        self.gain, self.saa, self.gradoff = 1.0, np.ones(self.nsa), None
        ints, _ = make(self)(flat)
        saa = ints.sum(axis=0)
        if c.saat > 0:
            # loc_reduce(saloc, saa, saat * mean, cont=1): per grid row, the
            # leading and trailing subapertures below the threshold go
            thres = c.saat * saa.mean()
            keep = np.ones(self.nsa, dtype=bool)
            rows = np.round(self.salocy / self.dsa * 2).astype(np.int64)
            for r in np.unique(rows):
                idx = np.where(rows == r)[0]
                idx = idx[np.argsort(self.salocx[idx])]
                for seq in (idx, idx[::-1]):
                    for i in seq:
                        if saa[i] < thres:
                            keep[i] = False
                        else:
                            break
            self.salocx, self.salocy, saa = self.salocx[keep].copy(), self.salocy[keep].copy(), saa[keep]
        self.saa = saa * (saa.size / saa.sum())
        f = make(self)
        _, goff = f(flat)
        self.gradoff = goff.copy()
        f = make(self)
        alpha = 0.005 * AS2RAD
        _, gx = f(alpha * self.locx)
        _, gy = f(alpha * self.locy)
        tt = np.concatenate([gx, gy]) / alpha
        gainscl = self.ng * self.nsa / tt.sum()
        self.gain *= gainscl
        self.gradoff = self.gradoff * gainscl
        self.GTT = tt * gainscl
        return self

    def oracle_struct(self, keep):
        """The oracle's orc_pywfs_t for these tables (test infrastructure)."""
        import oracle
        c = self.cfg
        o = oracle.OrcPywfs()
        arrs = dict(amp=self.amp, embed=self.embed, wvl=self.wvl, wvlwts=self.wvlwts, mask=self.mask.ravel(),
                    nominal=self.nominal, salocx=self.salocx, salocy=self.salocy, sioff=self.sioff, saa=self.saa,
                    gradoff=self.gradoff)
        for k, v in arrs.items():
            if v is not None:
                v = np.ascontiguousarray(v)
                keep.append(v)
                arrs[k] = v
        o.nloc, o.npsf, o.dx, o.nwvl = self.nloc, self.N, c.dx, self.nwvl
        o.amp = oracle._d(arrs["amp"])
        o.embed = oracle._l(arrs["embed"])
        o.wvl, o.wvlwts = oracle._d(arrs["wvl"]), oracle._d(arrs["wvlwts"])
        o.mask, o.nominal = oracle._d(arrs["mask"]), oracle._d(arrs["nominal"])
        o.nsa = self.nsa
        o.salocx, o.salocy = oracle._d(arrs["salocx"]), oracle._d(arrs["salocy"])
        o.nside, o.sioff = c.nside, oracle._d(arrs["sioff"])
        o.modulate, o.modulpos, o.modulring = self.modulate, c.modulpos, c.modulring
        o.sigmatch, o.raw, o.siglev, o.gain = c.sigmatch, c.raw, c.siglev, self.gain
        o.saa = oracle._d(arrs["saa"])
        o.gradoff = oracle._d(arrs["gradoff"]) if self.gradoff is not None else None
        return o
