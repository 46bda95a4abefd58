"""The workloads, expressed in MAOS configuration keys.

ngs8m:  8 m pupil (1.2 m obstruction), one NGS SH WFS 16x16 (0.5 m), dx=1/32
        m, 0.7 um, notf 32, 8x8 pixels of 0.5", pixblur 0.3, tCoG; one von
        Karman layer r0=0.186 m, L0=30 m, 2048^2 at 1/32 m, seed 1, wind
        10 m/s at 0 deg, dt=1/800 s, 100 steps.
lgs30m: 30 m pupil (3.6 m obstruction), six LGS WFS at 90 km on a 35"
        pentagon plus centre (powfs_shlgs.conf directions), 60x60 (0.5 m),
        dx=1/64 m, 0.589 um, notf 64, 6x6 pixels of 1", LLT d=0.4 m at the
        pupil centre, Gaussian sodium profile (10 km FWHM at 90 km), matched
        filter from an FP64-accumulated i0; seven layers at
        [0,0.5,1,2,4,8,16] km with Cn2 [.45,.15,.10,.10,.08,.07,.05],
        r0=0.186 m, L0=30 m, 4096^2 at 1/64 m, seeds 1-7, winds U[5,20] m/s
        in seeded directions, dt=1/800 s, 500 steps.
ngs_lo: three low-order NGS WFS on the 30 m pupil, 2x2 subapertures of 15 m,
        dx=1/16 m (240x240 points), embfac 2 (480^2 field) padded to a 512^2
        OTF, 1.25 and 1.65 um, 16x16 pixels of 8 mas, tCoG; the lgs30m
        atmosphere.
ngs_tt: one tip-tilt NGS WFS, 1x1 subaperture of 30 m, dx=1/16 m (480x480
        points, 960^2 field) padded to a 1024^2 OTF, 2.2 um, 16x16 pixels of
        8 mas, tCoG; the lgs30m atmosphere.
"""
import math

import numpy as np

from .engine import WfsSpec
from .powfs import (AS2RAD, PTYPE_COG, PTYPE_MF, AperCfg, AtmCfg, LltCfg, PowfsCfg,
                    SimCfg, SodiumCfg)

# wfs.thetax / wfs.thetay of powfs_shlgs.conf (arcsec)
LGS_THETAX = [0, 0, -33.287, -20.5725, 20.5725, 33.287]
LGS_THETAY = [0, 35, 10.8156, -28.3156, -28.3156, 10.8156]


def ngs8m(nx=2048):
    aper = AperCfg(d=8.0, din=1.2)
    pcfg = PowfsCfg(order=16, dsa=0.5, dx=1.0 / 32, saat=0.3, hs=math.inf, embfac=2,
                    wvl=[0.7e-6], wvlwts=[1.0], pixpsa=8, pixtheta=0.5, pixblur=0.3,
                    phytype=PTYPE_COG, sigmatch=1, siglev=1000.0)
    sim = SimCfg(dt=1.0 / 800)
    atm = AtmCfg(r0=0.186, L0=30.0, ht=[0.0], wt=[1.0], ws=[10.0], wddeg=[0.0],
                 dx=1.0 / 32, nx=nx, seeds=[1])
    wfs = [dict(thetax=0.0, thetay=0.0)]
    return dict(name="ngs8m", aper=aper, powfs=pcfg, sim=sim, atm=atm, wfs=wfs, nstep=100)


def lgs30m(nx=4096, nwfs=6, order=60, embfac=2, llt_d=0.4):
    """embfac=4 is the 128^2 padding variant (2x Nyquist OTF); llt_d above
    the 0.5 m subaperture makes the LLT grid wider than the subaperture's."""
    aper = AperCfg(d=30.0 * order / 60, din=3.6 * order / 60)
    pcfg = PowfsCfg(order=order, dsa=0.5, dx=1.0 / 64, saat=0.3, hs=90e3, embfac=embfac,
                    wvl=[0.589e-6], wvlwts=[1.0], pixpsa=6, pixtheta=1.0, pixblur=0.3,
                    phytype=PTYPE_MF, sigmatch=1, mtchcr=1, rne=3.0, siglev=960.0,
                    llt=LltCfg(d=llt_d, widthp=0.6, ox=0.0, oy=0.0),
                    sodium=SodiumCfg(center=90e3, fwhm=10e3, hmin=70e3, hmax=110e3, dh=20.0))
    sim = SimCfg(dt=1.0 / 800)
    rs = np.random.RandomState(2024)
    nps = 7
    atm = AtmCfg(r0=0.186, L0=30.0, ht=[0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 16000.0],
                 wt=[0.45, 0.15, 0.10, 0.10, 0.08, 0.07, 0.05],
                 ws=list(rs.uniform(5.0, 20.0, nps)), wddeg=list(rs.uniform(0.0, 360.0, nps)),
                 dx=1.0 / 64, nx=nx, seeds=list(range(1, nps + 1)))
    wfs = [dict(thetax=LGS_THETAX[i] * AS2RAD, thetay=LGS_THETAY[i] * AS2RAD)
           for i in range(nwfs)]
    return dict(name="lgs30m", aper=aper, powfs=pcfg, sim=sim, atm=atm, wfs=wfs, nstep=500)


# NGS asterism of the low-order WFS (arcsec): three stars on a 60" ring
NGS_THETAX = [60.0, -30.0, -30.0]
NGS_THETAY = [0.0, 51.9615, -51.9615]


def ngs_lo(nx=4096, nwfs=3):
    base = lgs30m(nx=nx)
    pcfg = PowfsCfg(order=2, dsa=15.0, dx=1.0 / 16, saat=0.3, hs=math.inf, embfac=2,
                    wvl=[1.25e-6, 1.65e-6], wvlwts=[0.5, 0.5], pixpsa=16, pixtheta=0.008,
                    pixblur=0.3, notf=512, phytype=PTYPE_COG, sigmatch=1, siglev=2000.0)
    wfs = [dict(thetax=NGS_THETAX[i] * AS2RAD, thetay=NGS_THETAY[i] * AS2RAD)
           for i in range(nwfs)]
    return dict(name="ngs_lo", aper=AperCfg(d=30.0, din=3.6), powfs=pcfg, sim=base["sim"],
                atm=base["atm"], wfs=wfs, nstep=1000)


def ngs_tt(nx=4096):
    base = lgs30m(nx=nx)
    pcfg = PowfsCfg(order=1, dsa=30.0, dx=1.0 / 16, saat=0.3, hs=math.inf, embfac=2,
                    wvl=[2.2e-6], wvlwts=[1.0], pixpsa=16, pixtheta=0.008, pixblur=0.3,
                    notf=1024, phytype=PTYPE_COG, sigmatch=1, siglev=5000.0)
    wfs = [dict(thetax=20.0 * AS2RAD, thetay=-15.0 * AS2RAD)]
    return dict(name="ngs_tt", aper=AperCfg(d=30.0, din=3.6), powfs=pcfg, sim=base["sim"],
                atm=base["atm"], wfs=wfs, nstep=1000)


def wfs_specs(cfg, powfs):
    """WfsSpec list (hs, sigsim, wvlwts from the powfs) for a config."""
    return [WfsSpec(thetax=w["thetax"], thetay=w["thetay"], hs=cfg["powfs"].hs,
                    sigsim=powfs.sigsim, wvlwts=powfs.wvlwts) for w in cfg["wfs"]]
