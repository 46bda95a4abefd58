"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

from conftest import ROOT


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:mx|mxs)_[a-z0-9_]+)\s*\(", txt)))


def test_gpu_lib_exports_header():
    from maos_amd import _lib
    from maos_amd.build import HIP_LIB
    lib = C.CDLL(HIP_LIB)
    names = _declared("maos_mi355x.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.HIP_API), "ctypes table and header disagree"


def test_setup_lib_exports_header():
    from maos_amd import _lib
    from maos_amd.build import SETUP_LIB
    lib = C.CDLL(SETUP_LIB)
    names = _declared("maos_setup.h")
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SETUP_API)


def test_gpu_lib_is_gfx950():
    from maos_amd.build import HIP_LIB
    data = open(HIP_LIB, "rb").read()
    assert b"gfx950" in data
