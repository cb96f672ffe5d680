"""The Fortran boundary: the reference's own call statements against the
x3div_amd modules.

tests/fortran/ref_calls.f90 is made of call statements of the reference
(each cited file:line, written as the reference writes them) on arrays of the
reference's shapes.  On the CPU the modules and the program are compiled in a
fresh directory: the modules accept the reference's callers.  On the GPU the
prebuilt program runs on a 32x24x16 grid (L = 1, the reference's vel field)
and every result it writes is compared with the oracle.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORTRAN = os.path.join(ROOT, "x3div_amd", "fortran")
REF_CALLS = os.path.join(ROOT, "tests", "fortran", "ref_calls.f90")
REF_CALLS_EXE = os.path.join(ROOT, "tests", "fortran", "ref_calls_hip")
FC = "/opt/rocm/bin/amdflang"
RTOL = 1e-12


def _amdflang():
    fc = FC if os.path.exists(FC) else shutil.which("amdflang")
    if fc is None:
        pytest.skip("amdflang is not installed")
    return fc


def test_reference_call_sites_compile(tmp_path):
    from x3div_amd.build import FSOURCES
    fc = _amdflang()
    for src in [s for s in FSOURCES if s != "xcompact3d.f90"] + [REF_CALLS]:
        path = src if os.path.isabs(src) else os.path.join(FORTRAN, src)
        obj = tmp_path / (os.path.basename(path) + ".o")
        r = subprocess.run([fc, "-cpp", "-c", path, "-module-dir", str(tmp_path), "-I", str(tmp_path),
                            "-o", str(obj)], capture_output=True, text=True)
        assert r.returncode == 0, f"{os.path.basename(path)}:\n{r.stdout}\n{r.stderr}"
    # every call site carries its reference citation
    text = open(REF_CALLS).read()
    cites = re.findall(r"! ref (src/\w+\.f90:[\d-]+)", text)
    assert len(cites) >= 25, cites


def _read(path, shape, count):
    data = np.fromfile(path, dtype=np.float64)
    n = int(np.prod(shape))
    assert data.size == n * count, (data.size, n, count)
    return [data[i * n:(i + 1) * n].reshape(shape) for i in range(count)]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.gpu
def test_reference_call_sites_run(torch_cuda, tmp_path):
    if not os.path.exists(REF_CALLS_EXE):
        pytest.fail("tests/fortran/ref_calls_hip is not built (x3div_amd.build.build_ref_calls)")
    nx, ny, nz = 32, 24, 16
    print(f"\n[ref_calls {nx}x{ny}x{nz}]", flush=True)
    r = subprocess.run([REF_CALLS_EXE, str(nx), str(ny), str(nz)], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0 and "ref_calls done" in r.stdout
    out = _read(tmp_path / "ref_calls.bin", (nz, ny, nx), 27)

    orc = Oracle(nx, ny, nz, 1.0, 1.0, 1.0)
    ux, uy, uz = orc.fields(0)
    d = orc.divergence
    expect = [
        orc.apply("derx", ux),
        orc.apply("deryp", ux),
        orc.apply("derzp", ux),
        orc.apply("derxvp", ux),
        orc.apply("intyvp", orc.apply("derxvp", ux)),
        orc.apply("derzvp", orc.apply("intyvp", orc.apply("derxvp", ux))),
        orc.apply("intxpv", orc.apply("derypv", orc.apply("intzpv", uy))),
        orc.thomas("derx", uy),
        orc.thomas12("derx", uy),
        orc.thomas("dery", uz),
        orc.apply("derz", ux),
        orc.apply("derzz", ux),
    ]
    for i, (a, b) in enumerate(zip(out[:12], expect)):
        assert _rel(a, b) < RTOL, (i + 1, _rel(a, b))
    # test_du's printed RMS error (src/mom.f90:59-243): the derivative is held
    # to 1e-12 of its max, and |rms(a - f) - rms(b - f)| <= max|a - b|
    m = re.search(r"RMS error in dudx:\s*([-+0-9.EeDd]+)", r.stdout)
    assert m, r.stdout
    e_orc = orc.test_err(expect[0], 0)
    assert abs(float(m.group(1).replace("D", "E")) - e_orc) <= RTOL * np.abs(expect[0]).max()

    # one step (dt = 0) of the reference's solenoidal vel field: the pressure
    # and its gradient are rounding noise, held in absolute terms
    ru, rdu, rpp, rg = orc.step(ux, uy, uz, gdt=0.0)
    umax = max(np.abs(a).max() for a in ru)
    for i, (a, b) in enumerate(zip(out[12:15] + out[15:18], ru + rdu)):
        assert _rel(a, b) < RTOL, ("step", i, _rel(a, b))
    for i, (a, b) in enumerate(zip(out[18:22], [rpp] + rg)):
        assert np.abs(a - b).max() < RTOL * umax, ("pressure", i, np.abs(a - b).max())
    # divergence, poisson and gradp of the corrected field: noise as well
    div = d(*ru)
    assert np.abs(out[22] - div).max() < RTOL * umax
    pp, g = orc.solve_poisson(*ru)
    for i, (a, b) in enumerate(zip(out[23:27], [pp] + g)):
        assert np.abs(a - b).max() < RTOL * umax, ("poisson", i, np.abs(a - b).max())
