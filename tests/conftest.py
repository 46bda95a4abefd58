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
