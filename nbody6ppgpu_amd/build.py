"""Build of the HIP engine: hipcc for gfx950, in-tree output lib/libgpunb.so.

Flags: -O3, C++17, -ffp-contract=off (the kernels spell out every fused
multiply-add they want, so the FP64 predicate and predictor stay exact), and
packed FP32 instructions disabled (-packed-fp32-ops) for all device code.
Only the extern "C" entry points of include/gpunb.h are exported.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgpunb.so")
SOURCES = ["kernels.hip", "jprep.hip", "gpunb.hip", "mlp.hip", "irr.hip", "regcor.hip",
           "energy.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPUNB_ARCH", "gfx950")


def hipcc_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
        "-fPIC", "-shared", "-fvisibility=hidden",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I" + os.path.join(ROOT, "include"),
    ]


def build_lib(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + hipcc_flags() + ["-o", tmp] + [
        os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, LIB)
    return LIB


HOSTDIR = os.path.join(HERE, "host")
BINDIR = os.path.join(HERE, "bin")
HERMITE = os.path.join(BINDIR, "hermite_ac")


def build_host(verbose: bool = False) -> str:
    """The host Hermite/Ahmad-Cohen driver (C++, OpenMP) linked to the
    library with an rpath relative to itself."""
    os.makedirs(BINDIR, exist_ok=True)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fopenmp",
           "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HOSTDIR, "hermite_ac.cpp"), "-L" + LIBDIR, "-lgpunb",
           "-Wl,-rpath,$ORIGIN/../lib", "-o", HERMITE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return HERMITE


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_host(verbose=True))
