import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


RAD2MAS = 180.0 / np.pi * 3600.0 * 1000.0


def rel_l2(a, b):
    """||a-b||_2 / ||b||_2"""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def ngs_small():
    """ngs8m geometry with a 512^2 screen (sizes the oracle finishes fast)."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.ngs8m(nx=512)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    layers = genatm(cfg["atm"])
    return cfg, p, layers


@pytest.fixture(scope="session")
def lgs_small():
    """lgs30m on a 10 m pupil (order 20) with 1024^2 screens, 6 WFS."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.lgs30m(nx=1024, order=20)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    layers = genatm(cfg["atm"])
    return cfg, p, layers


@pytest.fixture(scope="session")
def ngs_two_wvl():
    """A 4 m NGS case on the N=64 geometry with no LLT and two wavelengths
    (nominal DTF only): 8x8 subapertures of 0.5 m at 1/64 m, notf 64, 6x6
    pixels of 0.3", CoG; two von Karman layers of 512^2 at 1/64 m."""
    import math
    from maos_amd.powfs import PTYPE_COG, AperCfg, AtmCfg, Powfs, PowfsCfg, SimCfg, genatm
    pcfg = PowfsCfg(order=8, dsa=0.5, dx=1.0 / 64, saat=0.3, hs=math.inf, embfac=2,
                    wvl=[0.7e-6, 0.9e-6], wvlwts=[0.6, 0.4], pixpsa=6, pixtheta=0.3,
                    pixblur=0.3, notf=64, phytype=PTYPE_COG, sigmatch=1, siglev=1000.0)
    sim = SimCfg(dt=1.0 / 800)
    atm = AtmCfg(r0=0.15, L0=30.0, ht=[0.0, 5000.0], wt=[0.7, 0.3], ws=[10.0, 15.0],
                 wddeg=[0.0, 60.0], dx=1.0 / 64, nx=512, seeds=[11, 12])
    cfg = dict(name="ngs_two_wvl", aper=AperCfg(4.0, 0.6), powfs=pcfg, sim=sim, atm=atm,
               wfs=[dict(thetax=0.0, thetay=0.0), dict(thetax=5e-5, thetay=-3e-5)])
    p = Powfs(cfg["aper"], pcfg, sim)
    return cfg, p, genatm(atm)


@pytest.fixture(scope="session")
def lf_mixed():
    """A large-field case with mixed-radix sizes: 4 m pupil, 2x2 subapertures
    of 2 m at 1/24 m (48^2 points, 96^2 field = 2^5 3) padded to a 120^2 OTF
    (2^3 3 5), 1.25 and 1.65 um, 8x8 pixels of 30 mas, CoG; screens at 1/64 m
    (finer than the OPD grid)."""
    import math
    from maos_amd.powfs import PTYPE_COG, AperCfg, AtmCfg, Powfs, PowfsCfg, SimCfg, genatm
    pcfg = PowfsCfg(order=2, dsa=2.0, dx=1.0 / 24, saat=0.3, hs=math.inf, embfac=2,
                    wvl=[1.25e-6, 1.65e-6], wvlwts=[0.7, 0.3], pixpsa=8, pixtheta=0.03,
                    pixblur=0.3, notf=120, phytype=PTYPE_COG, sigmatch=1, siglev=1000.0)
    sim = SimCfg(dt=1.0 / 800)
    atm = AtmCfg(r0=0.15, L0=30.0, ht=[0.0, 6000.0], wt=[0.6, 0.4], ws=[8.0, 12.0],
                 wddeg=[30.0, 200.0], dx=1.0 / 64, nx=512, seeds=[21, 22])
    cfg = dict(name="lf_mixed", aper=AperCfg(4.0, 0.6), powfs=pcfg, sim=sim, atm=atm,
               wfs=[dict(thetax=0.0, thetay=0.0), dict(thetax=-4e-5, thetay=6e-5)])
    p = Powfs(cfg["aper"], pcfg, sim)
    return cfg, p, genatm(atm)


@pytest.fixture(scope="session")
def ngs_lo_small():
    """ngs_lo (3 WFS, 2x2 of 15 m, 480^2 field in a 512^2 OTF, two
    wavelengths) with 1024^2 screens."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.ngs_lo(nx=1024)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    return cfg, p, genatm(cfg["atm"])


@pytest.fixture(scope="session")
def ngs_tt_small():
    """ngs_tt (1x1 of 30 m, 960^2 field in a 1024^2 OTF) with 1024^2 screens."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.ngs_tt(nx=1024)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    return cfg, p, genatm(cfg["atm"])


@pytest.fixture(scope="session")
def lgs_wide_llt():
    """lgs_small with a 0.6 m launch telescope: the LLT grid (38^2 points at
    1/64 m) is wider than the 0.5 m subaperture's 32^2, so MAOS crops the
    uplink OTF from 76^2 to 64^2 (maos/wfsints.c:160-185)."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.lgs30m(nx=1024, order=20, llt_d=0.6)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    return cfg, p, genatm(cfg["atm"])


@pytest.fixture(scope="session")
def lgs_pad128():
    """lgs_small with embfac 4: 32^2 points padded to a 128^2 field and OTF
    (the 2x Nyquist variant), LLT grid 32^2 -> nlwvf 128, ETF at 128^2."""
    from maos_amd import configs
    from maos_amd.powfs import Powfs, genatm
    cfg = configs.lgs30m(nx=1024, order=20, embfac=4)
    p = Powfs(cfg["aper"], cfg["powfs"], cfg["sim"])
    return cfg, p, genatm(cfg["atm"])
