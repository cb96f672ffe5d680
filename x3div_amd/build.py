"""Build libx3dhip.so in-tree for gfx950 with hipcc.

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions
disabled, then linked against hipFFT.  Usage: ``python -m x3div_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libx3dhip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = ["x3d_api.hip", "x3d_batch.hip", "x3d_compact.hip", "x3d_fused.hip", "x3d_mfma.hip", "x3d_poisson.hip"]

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep hand-written FP32 code off the packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN = os.path.join(HERE, "fortran")
FC = os.path.join(ROCM, "bin", "amdflang")
# module order: each file uses modules of the files before it
FSOURCES = ["x3d_hip.f90", "x3d_hip_step.f90", "xcompact3d.f90"]
FDRIVER = os.path.join(HERE, "xcompact3d_hip")


def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran modules and the xcompact3d driver against libx3dhip."""
    fbuild = os.path.join(BUILD, "fortran")
    os.makedirs(fbuild, exist_ok=True)
    objs = []
    for src in FSOURCES:
        obj = os.path.join(fbuild, os.path.splitext(src)[0] + ".o")
        cmd = [FC, "-O2", "-c", os.path.join(FORTRAN, src), "-module-dir", fbuild,
               "-I", fbuild, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = FDRIVER + ".tmp"
    cmd = [FC, *objs, "-o", tmp, f"-L{HERE}", "-lx3dhip", "-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FDRIVER)
    if verbose:
        print(f"built {FDRIVER}")
    return FDRIVER


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp,
           f"-L{ROCM}/lib", "-lhipfft", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
    sys.exit(0)
