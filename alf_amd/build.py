"""Build the gfx950 HIP library in-tree (alf_amd/libalfq.so) with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["alfq.hip", "gemm.hip", "qr.hip", "chain.hip", "rate.hip", "obs.hip", "checkerboard.hip", "tau.hip",
           "ozaki.hip"]
OUT = os.path.join(HERE, "libalfq.so")

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-result", "-Wno-unused-value",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.sep not in p or os.path.exists(p)):
            return p
    raise RuntimeError("hipcc not found")


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_DEMO = os.path.join(FORTRAN_DIR, "alfq_ring_demo")


def amdflang() -> str | None:
    for p in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang"):
        if p and os.path.exists(p):
            return p
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING module and the Fortran host demo against libalfq.so."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran binding not built")
        return None
    cmd = [fc, "-O2", "-J", FORTRAN_DIR, os.path.join(FORTRAN_DIR, "alfq_mod.F90"),
           os.path.join(FORTRAN_DIR, "alfq_ring_demo.F90"), "-o", FORTRAN_DEMO + ".tmp",
           "-L" + HERE, "-lalfq", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=FORTRAN_DIR)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("amdflang failed building the Fortran binding")
    os.replace(FORTRAN_DEMO + ".tmp", FORTRAN_DEMO)
    if verbose:
        print("built", FORTRAN_DEMO)
    return FORTRAN_DEMO


def build_library(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", OUT + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libalfq.so")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build_library(verbose=True)
    build_fortran(verbose=True)
