"""Build step for the HIP library (bsr_amd/libbsr_gpu.so) -- gfx950 only.

Run as ``python -m bsr_amd.build``; ``__graft_entry__.build()`` calls it too.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbsr_gpu.so")
SOURCES = ["host_tables.cpp", "lists.cpp", "kernels.hip", "capi.hip", "lists_dev.hip", "ozaki.hip", "dbs.hip", "dirac.hip", "eigen.hip", "xasm.hip", "dirac_diag.hip", "state_res.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *FLAGS, *lang, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs,
           "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_fortran(verbose)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "bsr_mat_gpu")


def build_fortran(verbose: bool = False) -> str:
    """The Fortran host: module bsr_gpu (ISO_C_BINDING), the driver
    program bsr_mat_gpu and the drop-in I_data/L_data/O_data it calls,
    linked against libbsr_gpu.so."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    mod = os.path.join(FORTRAN, "_mod")
    os.makedirs(mod, exist_ok=True)
    cmd = [fc, "-O2", "-module-dir", mod, os.path.join(FORTRAN, "bsr_gpu.f90"),
           os.path.join(FORTRAN, "bsr_mat_gpu.f90"), os.path.join(FORTRAN, "bsr_gpu_dropin.f90"),
           "-o", DRIVER + ".tmp", "-L" + HERE, "-lbsr_gpu",
           "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
