"""Grouped Poisson of the batching form: one launch per stage and line length
over every grid the line kernels take, the other grids solved one by one.

* Against the per-grid solves (X3D_BATCH_PGROUP=0) on the same fields: in FP32
  both run the same line kernels with contraction off, so the steps agree bit
  for bit; in FP64 the per-grid inverse x transform is hipFFT's, the grouped
  one the line kernel, so they agree to rounding from the Poisson solve on:
  the right-hand sides du (formed before it) bit for bit, the pressure, its
  gradient and the corrected velocities to 1e-12 relative (the oracle bar
  for the pressure is 1e-11), over one step (a second
  step's pressure comes from the O(gdt) divergence left after the first
  projection, about 1e3 times smaller, so the same rounding reads 1e3 times
  larger relative to it).
* The launches of one pass or Poisson stage side by side on the side streams
  (the default) against the same launches in order (X3D_BATCH_OVERLAP=0):
  bit for bit, eager and captured.
* A batch mixing grouped grids with ones solved alone (an x length the x line
  kernel is not built for, a 512-point z line) against the oracle at the FP64
  bar, eager and through a captured graph (replays equal to eager steps bit
  for bit).
"""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu

DIMS = [(32, 48, 64), (48, 32, 32), (64, 64, 48), (32, 48, 64), (96, 32, 40), (40, 160, 32), (16, 24, 8)]
MIXED = [(32, 48, 64), (16, 16, 512), (48, 32, 32), (20, 24, 16), (32, 48, 64)]
# longer grouped lines (1024-thread tiles): x 192..384 (r2c / c2r on 96..192
# points), y and z 192, 320 and 384
LONG = [(384, 16, 16), (192, 320, 8), (16, 24, 384), (32, 192, 320), (320, 16, 192)]
SEED = 97
FIELDS = ("ux", "uy", "uz", "dux", "duy", "duz", "pp3", "px", "py", "pz")


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _batch(dims, prec, grouped, monkeypatch):
    from x3div_amd import X3DBatch
    if grouped:
        monkeypatch.delenv("X3D_BATCH_PGROUP", raising=False)
    else:
        monkeypatch.setenv("X3D_BATCH_PGROUP", "0")
    b = X3DBatch(dims, precision=prec)
    monkeypatch.delenv("X3D_BATCH_PGROUP", raising=False)
    b.init_fields(kind=2, seed=SEED)
    return b


@pytest.mark.parametrize("dims", [DIMS, LONG], ids=["short", "long"])
@pytest.mark.parametrize("prec", [8, 4])
def test_grouped_poisson_equals_per_grid(torch_cuda, prec, dims, monkeypatch):
    print(f"\n[grouped vs per-grid Poisson, {dims}, precision {prec}]", flush=True)
    a = _batch(dims, prec, True, monkeypatch)
    b = _batch(dims, prec, False, monkeypatch)
    try:
        for step in range(2 if prec == 4 else 1):
            a.step(gdt=1e-3)
            b.step(gdt=1e-3)
            for i in range(len(dims)):
                for name in FIELDS:
                    x, y = a.get(i, name), b.get(i, name)
                    e = relerr(x, y)
                    if prec == 4:
                        assert np.array_equal(x, y), (step, i, dims[i], name, e)
                    elif name.startswith("du"):
                        assert np.array_equal(x, y), (step, i, dims[i], name, e)
                    else:
                        assert e < 1e-12, (step, i, dims[i], name, e)
    finally:
        a.close()
        b.close()


def test_long_grouped_lines_match_oracle(torch_cuda, monkeypatch):
    print(f"\n[grouped long lines vs oracle {LONG}]", flush=True)
    b = _batch(LONG, 8, True, monkeypatch)
    try:
        b.step(gdt=1e-3)
        for i, (nx, ny, nz) in enumerate(LONG):
            o = Oracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi)
            u, du, pp, g = o.step(*o.fields(2, SEED + i, 0.0), gdt=1e-3)
            for name, r in zip(FIELDS[:6], u + du):
                assert relerr(b.get(i, name), r) < 1e-12, (i, LONG[i], name)
            assert relerr(b.get(i, "pp3"), pp) < 1e-11, (i, LONG[i])
            for name, r in zip(FIELDS[7:], g):
                assert relerr(b.get(i, name), r) < 1e-11, (i, LONG[i], name)
    finally:
        b.close()


def test_no_side_streams(torch_cuda, monkeypatch):
    """X3D_BATCH_PSTREAMS=0: every launch on the batch stream, eager and
    captured, equal to the default batch bit for bit."""
    from x3div_amd import X3DBatch
    print("\n[batch without side streams]", flush=True)
    monkeypatch.setenv("X3D_BATCH_PSTREAMS", "0")
    plain = X3DBatch(MIXED, precision=8)
    captured = X3DBatch(MIXED, precision=8)
    monkeypatch.delenv("X3D_BATCH_PSTREAMS")
    ref = X3DBatch(MIXED, precision=8)
    try:
        for b in (plain, captured, ref):
            b.init_fields(kind=2, seed=SEED)
        captured.graph_capture(gdt=1e-3)
        for step in range(2):
            plain.step(gdt=1e-3)
            captured.graph_launch()
            ref.step(gdt=1e-3)
            for i in range(len(MIXED)):
                for name in FIELDS:
                    r = ref.get(i, name)
                    assert np.array_equal(plain.get(i, name), r), (step, i, name)
                    assert np.array_equal(captured.get(i, name), r), (step, i, name)
    finally:
        plain.close()
        captured.close()
        ref.close()


def test_mixed_batch_matches_oracle_eager_and_captured(torch_cuda, monkeypatch):
    print(f"\n[mixed grouped / alone batch {MIXED}]", flush=True)
    eager = _batch(MIXED, 8, True, monkeypatch)
    graph = _batch(MIXED, 8, True, monkeypatch)
    try:
        eager.step(gdt=1e-3)
        for i, (nx, ny, nz) in enumerate(MIXED):
            o = Oracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi)
            u, du, pp, g = o.step(*o.fields(2, SEED + i, 0.0), gdt=1e-3)
            for name, r in zip(FIELDS[:6], u + du):
                assert relerr(eager.get(i, name), r) < 1e-12, (i, MIXED[i], name)
            assert relerr(eager.get(i, "pp3"), pp) < 1e-11, (i, MIXED[i])
            for name, r in zip(FIELDS[7:], g):
                assert relerr(eager.get(i, name), r) < 1e-11, (i, MIXED[i], name)
        graph.graph_capture(gdt=1e-3)
        for rep in range(3):
            if rep:
                eager.step(gdt=1e-3)
            graph.graph_launch()
            for i in range(len(MIXED)):
                for name in FIELDS:
                    assert np.array_equal(graph.get(i, name), eager.get(i, name)), (rep, i, name)
    finally:
        eager.close()
        graph.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_overlapped_launches_equal_serial(torch_cuda, prec, monkeypatch):
    from x3div_amd import X3DBatch
    print(f"\n[overlapped vs serial launches, precision {prec}]", flush=True)
    monkeypatch.setenv("X3D_BATCH_OVERLAP", "0")
    serial = X3DBatch(DIMS, precision=prec)
    monkeypatch.delenv("X3D_BATCH_OVERLAP")
    over = X3DBatch(DIMS, precision=prec)
    captured = X3DBatch(DIMS, precision=prec)
    try:
        for b in (serial, over, captured):
            b.init_fields(kind=2, seed=SEED)
        captured.graph_capture(gdt=1e-3)
        for step in range(2):
            serial.step(gdt=1e-3)
            over.step(gdt=1e-3)
            captured.graph_launch()
            for i in range(len(DIMS)):
                for name in FIELDS:
                    ref = serial.get(i, name)
                    assert np.array_equal(over.get(i, name), ref), (step, i, name)
                    assert np.array_equal(captured.get(i, name), ref), (step, i, name)
    finally:
        serial.close()
        over.close()
        captured.close()
