"""Build recipes for the native libraries (run from any directory).

    python -m maos_amd.build            # GPU library + host setup library
    python -m maos_amd.build --oracle   # also the FP64 oracle (tests, bench)

The GPU library is compiled for gfx950 only.  Packed-FP32 instruction
selection is disabled for all device code.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "maos_amd")
LIBDIR = os.path.join(PKG, "lib")

HIP_LIB = os.path.join(LIBDIR, "libmaos_mi355x.so")
SETUP_LIB = os.path.join(LIBDIR, "libmaos_setup.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-shared", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(PKG, "csrc", "shwfs.hip")
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp, src])
    os.replace(tmp, HIP_LIB)


def build_setup():
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(PKG, "host", f) for f in ("setup.c", "binio.c", "gensei.c", "mkg.c", "pywfs.c", "loop.c")]
    tmp = SETUP_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=gnu11", "-fopenmp", "-fPIC", "-shared", "-Wall",
          "-o", tmp] + srcs + ["-lz", "-lm"])
    os.replace(tmp, SETUP_LIB)


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])


def build_all(oracle=True):
    build_setup()
    build_hip()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all(oracle="--oracle" in sys.argv or "--all" in sys.argv)
