"""The Fortran driver (x3div's xcompact3d program on libx3dhip) against the oracle.

The driver keeps the reference's command line and prints; in test mode it
reports the RMS errors of derx(ux), dery(uy), derz(uz) against the analytic
derivatives of the initial field (src/mom.f90:59-243) and the divergence
statistics of the predicted velocity (src/navier.f90:213-231).  The RMS
errors of both steps are compared with the oracle's on the same fields.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "x3div_amd", "xcompact3d_hip")
FLOAT = r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?"


def _driver():
    if not os.path.exists(DRIVER):
        if shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"):
            pytest.skip("amdflang not available to build the Fortran driver")
        from x3div_amd.build import build_fortran
        build_fortran()
    return DRIVER


def test_driver_reports_library_errors():
    """Without a usable device x3d_create fails and the driver stops with the
    status, the decomp_2d_abort path; with one it runs (checked on the GPU)."""
    exe = _driver()
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([exe, "8", "8", "8", "0", "0", "1", "0"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode != 0
    assert "x3d_create failed with status" in r.stdout + r.stderr


def _floats(lines, key):
    out = []
    for ln in lines:
        if key in ln:
            out.append([float(v.replace("D", "E").replace("d", "e"))
                        for v in re.findall(FLOAT, ln.split(key, 1)[1])])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [(64, 2), (16, 3)], ids=["config1-64cube", "16cube"])
def test_driver_test_mode_matches_oracle(torch_cuda, n, steps):
    """x3div in test mode (the 64^3, test=1 configuration): the start-up
    banner, then per step the RMS errors of derx(ux), dery(uy), derz(uz) of
    the current field and the divergence statistics of u*; every RMS line is
    compared with the oracle's on the same step."""
    exe = _driver()
    print(f"\n[fortran driver {n}^3, {steps} steps, test mode]", flush=True)
    r = subprocess.run([exe, str(n), str(n), str(n), "0", "0", str(steps), "1"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    # listing() (src/parameters.f90:141-228)
    for key in ("===============mini app of Xcompact3D======================", "Temporal scheme",
                "Forwards Euler", "Numerical precision: Double"):
        assert any(key in ln for ln in lines), key
    assert [int(v) for v in re.findall(r"n[xyz]\s+:\s+(\d+)", r.stdout)] == [n, n, n]
    assert len(_floats(lines, "Elapse time min")) == steps
    got = {k: [v[0] for v in _floats(lines, f"RMS error in {k}:")] for k in ("dudx", "dvdy", "dwdz")}
    divs = _floats(lines, "DIV U* max mean=")
    assert len(divs) == steps and all(len(got[k]) == steps for k in got)

    orc = Oracle(n, n, n, 1.0, 1.0, 1.0)
    u = orc.fields(0, 0, 0.0)
    for s in range(steps):
        for which, (k, op) in enumerate((("dudx", "derx"), ("dvdy", "dery"), ("dwdz", "derz"))):
            d = orc.apply(op, u[which])
            w = orc.test_err(d, which)
            # |rms(a - f) - rms(b - f)| <= max|a - b| <= 1e-12 max|b|, plus the
            # printed digits
            assert abs(got[k][s] - w) <= 1e-12 * np.abs(d).max() + 1e-15 * abs(w), (s, k, got[k][s], w)
        pp = orc.divergence(*u)
        mx, mean = orc.div_stats(pp)
        # the reference's field is solenoidal: both sides print rounding noise
        assert divs[s][0] < 1e-9 and mx < 1e-9, (s, divs[s], mx)
        u = list(orc.step(*u, gdt=0.0)[0])


@pytest.mark.gpu
def test_driver_replays_one_graph_per_step(torch_cuda):
    """Outside test mode the driver records its four step calls once and
    replays the graph each step (x3d_graph_begin/end/launch)."""
    exe = _driver()
    n, steps = 64, 4
    print(f"\n[fortran driver {n}^3, {steps} steps, graph replay]", flush=True)
    r = subprocess.run([exe, str(n), str(n), str(n), "0", "0", str(steps), "0"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    el = [v[0] for v in _floats(lines, "Elapse time min")]
    assert len(el) == steps and all(b >= a for a, b in zip(el, el[1:]))
    done = re.search(r"Timesteps completed:\s+(\d+)", r.stdout)
    assert done and int(done.group(1)) == steps + 1
    assert not _floats(lines, "RMS error in")


@pytest.mark.gpu
def test_driver_debug_build_prints_init_lines(torch_cuda):
    """The -DDEBUG build prints the reference's DEBUG-only progress lines
    (src/parameters.f90:32, src/x3d_operator_1d.f90:73,192)."""
    from x3div_amd.build import FDRIVER_DEBUG
    if not os.path.exists(FDRIVER_DEBUG):
        pytest.fail("the -DDEBUG driver is not built (x3div_amd.build.build_fortran)")
    r = subprocess.run([FDRIVER_DEBUG, "16", "16", "16", "0", "0", "1", "0"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    for key in ("# parameter start", "# x3d_operator_1d_init start", "# x3d_operator_1d_init done",
                "Compile flag DEBUG detected"):
        assert key in out, key
    assert out.index("# parameter start") < out.index("mini app of Xcompact3D") < \
        out.index("# x3d_operator_1d_init start") < out.index("# x3d_operator_1d_init done")
