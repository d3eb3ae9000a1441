"""Build recipe of the engine library (hipcc, gfx950) and the Fortran binding.

    python -m ggi_tlm_amd.build

compiles csrc/tlmscn.hip into ggi_tlm_amd/libtlmscn.so in-tree.  The file is
compiled with -ffp-contract=off: the plain FP64 form must not fuse multiply
and add, so that it reproduces the reference's source-order arithmetic bit
for bit.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "libtlmscn.so")
SRC = [os.path.join(HERE, "csrc", "tlmscn.hip"), os.path.join(HERE, "csrc", "batch.hip")]
ARCH = os.environ.get("TLM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_engine(verbose: bool = False) -> str:
    tmp = SO + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtlmscn.so")
    os.replace(tmp, SO)
    return SO


COPY_SRC = os.path.join(HERE, "csrc", "copy_probe.hip")
COPY_BIN = os.path.join(HERE, "copy_probe")


def build_copy_probe(verbose: bool = False) -> str:
    """HBM calibration program (plain copy of known size) used when scaling
    rocprof FETCH_SIZE / WRITE_SIZE and as the practical bandwidth ceiling."""
    cmd = [hipcc(), "-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", COPY_SRC, "-o", COPY_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building copy_probe")
    return COPY_BIN


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BIN = os.path.join(FORTRAN_DIR, "tlm_gpu_cavity")


def amdflang() -> str:
    for c in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("amdflang not found")


def build_fortran(verbose: bool = False) -> str:
    """The iso_c_binding module and the Fortran host driver, linked against
    the engine library (rpath to this directory)."""
    mod_dir = os.path.join(FORTRAN_DIR, "_build")
    os.makedirs(mod_dir, exist_ok=True)
    cmd = [amdflang(), "-O2", "-cpp", "-J", mod_dir,
           os.path.join(FORTRAN_DIR, "tlm_gpu_iface.F90"), os.path.join(FORTRAN_DIR, "tlm_gpu_cavity.F90"),
           "-L", HERE, "-ltlmscn", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=mod_dir)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("amdflang failed building the Fortran driver")
    return FORTRAN_BIN


if __name__ == "__main__":
    print(build_engine(verbose=True))
    print(build_copy_probe(verbose=True))
    print(build_fortran(verbose=True))
