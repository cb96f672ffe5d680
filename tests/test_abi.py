"""The C-ABI library loads on the CPU and exports every symbol include/x3dhip.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "x3dhip.h")
LIB = os.path.join(ROOT, "x3div_amd", "libx3dhip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(x3d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("x3d_create", "x3d_derx", "x3d_interzpv", "x3d_poisson", "x3d_step",
              "x3d_xthomas", "x3d_transpose_y_to_x"):
        assert s in syms
    assert len(syms) >= 49


@pytest.mark.skipif(not os.path.exists(LIB), reason="libx3dhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libx3dhip.so not built")
def test_binding_covers_header():
    from x3div_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared_symbols())
    lib = _lib.load()
    assert lib.x3d_version() >= 100
    assert lib.x3d_strerror(-2).decode().startswith("array sizes")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libx3dhip.so not built")
def test_library_is_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf" if os.path.exists("/opt/rocm/llvm/bin/llvm-readelf") else "readelf",
                          "-S", LIB], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out


@pytest.mark.skipif(not os.path.exists(LIB), reason="libx3dhip.so not built")
def test_create_rejects_bad_sizes_without_gpu_work():
    # x3d_create validates sizes before touching the device
    lib = ctypes.CDLL(LIB)
    h = ctypes.c_void_p()
    lib.x3d_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_void_p)]
    assert lib.x3d_create(4, 16, 16, 1.0, 1.0, 1.0, ctypes.byref(h)) == -2
    assert lib.x3d_create(16, 16, 16, -1.0, 1.0, 1.0, ctypes.byref(h)) == -1


def fortran_bindings():
    names = set()
    fdir = os.path.join(ROOT, "x3div_amd", "fortran")
    for f in os.listdir(fdir):
        if f.endswith(".f90"):
            names |= set(re.findall(r'bind\(C,\s*name="(x3d_[a-z0-9_]+)"\)', open(os.path.join(fdir, f)).read()))
    return names


def test_fortran_binds_only_declared_symbols():
    """Every bind(C) name of the Fortran modules is a header entry point, and the
    hot-path entries the reference's modules call are all bound."""
    bound = fortran_bindings()
    assert bound <= set(declared_symbols()), bound - set(declared_symbols())
    hot = {"x3d_xthomas", "x3d_ythomas", "x3d_zthomas", "x3d_derx", "x3d_dery", "x3d_derz",
           "x3d_poisson", "x3d_momentum_rhs", "x3d_int_time", "x3d_divergence", "x3d_gradp",
           "x3d_cor_vel", "x3d_solve_poisson"}
    hot |= {f"x3d_{k}{d}{s}" for k in ("der", "inter") for d in "xyz" for s in ("vp", "pv")}
    hot |= {f"x3d_transpose_{a}" for a in ("x_to_y", "y_to_z", "z_to_y", "y_to_x")}
    assert hot <= bound, hot - bound
