"""Builds the HIP library (gfx950) in-tree: spectral_rs_amd/lib/libspectral_rs_hip.so.

    python -m spectral_rs_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libspectral_rs_hip.so")
BINDIR = os.path.join(HERE, "bin")
HOST = os.path.join(BINDIR, "spectral_rs")
SOURCES = ["api.hip", "particles.hip", "fields.hip", "fft.hip", "diag.hip", "denoise.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # reference arithmetic: no contraction into FMA, IEEE division and sqrt
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    # hardware float atomics for the deposit
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include")] + \
        [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    shutil.move(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose=False):
    """The C++ host program (mirror of the reference binary) over the C ABI."""
    os.makedirs(BINDIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "spectral_rs.cpp"), "-o", HOST + ".tmp",
           "-L", LIBDIR, "-lspectral_rs_hip", "-Wl,-rpath,$ORIGIN/../lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    shutil.move(HOST + ".tmp", HOST)
    return HOST


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
