"""The batching form: many uneven grids stepped by grouped persistent launches.

Every grid of a batch must come out as the single-grid step would: each is
checked against the CPU oracle's step on the same seeded fields (FP64 at the
1e-12 bar, FP32 at the precision form's bounds).  The batches mix line
lengths so that several tiling groups share each launch, repeat a grid shape,
and include lengths that need different chunk lengths.  Graph replays of the
batched step must equal eager batched steps bit for bit, and the autotuner
may change tilings but not results beyond rounding.
"""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu

DIMS = [(32, 48, 64), (48, 32, 32), (64, 64, 48), (32, 48, 64), (96, 32, 40), (40, 160, 32)]
SEED = 4321


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_steps(dims, gdt, steps=1):
    out = []
    for i, (nx, ny, nz) in enumerate(dims):
        o = Oracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi)
        u = o.fields(2, SEED + i, 0.0)
        for _ in range(steps):
            u, du, pp, g = o.step(*u, gdt=gdt)
        out.append((u, du, pp, g))
    return out


@pytest.mark.parametrize("prec", [8, 4])
def test_batch_step_matches_oracle(torch_cuda, prec):
    from x3div_amd import X3DBatch
    print(f"\n[batch {len(DIMS)} grids, precision {prec}]", flush=True)
    b = X3DBatch(DIMS, precision=prec)
    b.init_fields(kind=2, seed=SEED)
    b.step(gdt=1e-3)
    ref = oracle_steps(DIMS, 1e-3)
    tol_u, tol_pp = (1e-12, 1e-11) if prec == 8 else (2e-4, 2e-3)
    for i, (u, du, pp, g) in enumerate(ref):
        for name, r in zip(("ux", "uy", "uz"), u):
            assert relerr(b.get(i, name), r) < tol_u, (i, name)
        for name, r in zip(("dux", "duy", "duz"), du):
            assert relerr(b.get(i, name), r) < (1e-12 if prec == 8 else 2e-6), (i, name)
        assert relerr(b.get(i, "pp3"), pp) < tol_pp, i
    b.close()


def test_batch_graph_replays_match_eager(torch_cuda):
    from x3div_amd import X3DBatch
    eager = X3DBatch(DIMS, precision=8)
    graph = X3DBatch(DIMS, precision=8)
    for b in (eager, graph):
        b.init_fields(kind=2, seed=SEED)
    graph.graph_capture(gdt=1e-3)
    for rep in range(3):
        eager.step(gdt=1e-3)
        graph.graph_launch()
        for i in range(len(DIMS)):
            for name in ("ux", "uy", "uz", "pp3", "px"):
                assert np.array_equal(graph.get(i, name), eager.get(i, name)), (rep, i, name)
    eager.close()
    graph.close()


def test_batch_autotune_keeps_results(torch_cuda):
    from x3div_amd import X3DBatch
    b = X3DBatch(DIMS[:4], precision=8)
    report = b.autotune(reps=1)
    print(report)
    assert "ms/step" in report
    b.init_fields(kind=2, seed=SEED)
    b.step(gdt=0.0)
    ref = oracle_steps(DIMS[:4], 0.0)
    for i, (u, du, pp, g) in enumerate(ref):
        for name, r in zip(("ux", "uy", "uz"), u):
            assert relerr(b.get(i, name), r) < 1e-12, (i, name)
    b.close()


def test_benchmark_batch_subset_tuned_and_captured(torch_cuda):
    """The bench's own batch: the first 24 of its 96 seeded grids (all six
    sizes 32..160 appear), autotuned, then stepped through the captured graph,
    each grid against the oracle's step."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import BATCH_SIZES, batch_dims
    from x3div_amd import X3DBatch
    dims = batch_dims()[:24]
    assert set(BATCH_SIZES) <= {v for d in dims for v in d}, dims
    print(f"\n[bench batch subset: {len(dims)} grids]", flush=True)
    b = X3DBatch(dims, precision=8)
    print(b.autotune(reps=1))
    b.init_fields(kind=2, seed=SEED)
    b.graph_capture(gdt=1e-3)
    b.graph_launch()
    ref = oracle_steps(dims, 1e-3)
    for i, (u, du, pp, g) in enumerate(ref):
        for name, r in zip(("ux", "uy", "uz", "dux", "duy", "duz"), u + du):
            assert relerr(b.get(i, name), r) < 1e-12, (i, dims[i], name)
        assert relerr(b.get(i, "pp3"), pp) < 1e-11, (i, dims[i])
        for name, r in zip(("px", "py", "pz"), g):
            assert relerr(b.get(i, name), r) < 1e-11, (i, dims[i], name)
    b.close()
