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

SOURCES = ["x3d_api.hip", "x3d_batch.hip", "x3d_bfp.hip", "x3d_compact.hip", "x3d_fused.hip", "x3d_mfma.hip",
           "x3d_poisson.hip"]

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
FSOURCES = ["x3d_hip.f90", "decomp_2d.f90", "x3d_hip_ops.f90", "x3d_hip_step.f90", "x3d_tools.f90",
            "xcompact3d.f90"]
FDRIVER = os.path.join(HERE, "xcompact3d_hip")
# the same driver built with -DDEBUG, which prints the reference's DEBUG-only
# progress lines ('# parameter start', '# x3d_operator_1d_init start/done')
FDRIVER_DEBUG = os.path.join(HERE, "xcompact3d_hip_debug")
FMODDIR = os.path.join(BUILD, "fortran")


def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran modules and the xcompact3d driver against libx3dhip
    (and the -DDEBUG variant of the driver)."""
    for exe, defs, sub in ((FDRIVER, [], "fortran"), (FDRIVER_DEBUG, ["-DDEBUG"], "fortran_debug")):
        fbuild = os.path.join(BUILD, sub)
        os.makedirs(fbuild, exist_ok=True)
        objs = []
        for src in FSOURCES:
            obj = os.path.join(fbuild, os.path.splitext(src)[0] + ".o")
            cmd = [FC, "-O2", "-cpp", *defs, "-c", os.path.join(FORTRAN, src), "-module-dir", fbuild,
                   "-I", fbuild, "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"amdflang failed for {src}:\n{r.stdout}\n{r.stderr}")
            objs.append(obj)
        tmp = exe + ".tmp"
        cmd = [FC, *objs, "-o", tmp, f"-L{HERE}", "-lx3dhip", "-Wl,-rpath,$ORIGIN",
               f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran link failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, exe)
        if verbose:
            print(f"built {exe}")
    return FDRIVER


REF_CALLS = os.path.join(os.path.dirname(HERE), "tests", "fortran", "ref_calls.f90")
REF_CALLS_EXE = os.path.join(os.path.dirname(HERE), "tests", "fortran", "ref_calls_hip")


def build_ref_calls(verbose: bool = False) -> str:
    """Compile and link the test program of reference call sites against the
    Fortran modules (build_fortran first) and libx3dhip."""
    fbuild = os.path.join(BUILD, "fortran")
    objs = [os.path.join(fbuild, os.path.splitext(src)[0] + ".o") for src in FSOURCES
            if src not in ("xcompact3d.f90", "x3d_tools.f90")]
    objs.append(os.path.join(fbuild, "x3d_tools.o"))
    obj = os.path.join(fbuild, "ref_calls.o")
    cmd = [FC, "-O2", "-cpp", "-c", REF_CALLS, "-I", fbuild, "-module-dir", os.path.join(fbuild, "ref_calls_mod"),
           "-o", obj]
    os.makedirs(os.path.join(fbuild, "ref_calls_mod"), exist_ok=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for ref_calls.f90:\n{r.stdout}\n{r.stderr}")
    tmp = REF_CALLS_EXE + ".tmp"
    cmd = [FC, *objs, obj, "-o", tmp, f"-L{HERE}", "-lx3dhip", "-Wl,-rpath,$ORIGIN/../../x3div_amd", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ref_calls link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, REF_CALLS_EXE)
    if verbose:
        print(f"built {REF_CALLS_EXE}")
    return REF_CALLS_EXE


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
    build_ref_calls(verbose=True)
    sys.exit(0)
