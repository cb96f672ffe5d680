"""The captured step: graph replays against eager steps on the same inputs.

The graph records the fused passes and the hipFFT calls of one step.  Each
replay is checked against an eager step, with other eager work and host
synchronisations between the replays, so that a graph node that misbehaves
after its first replay is caught.  Same kernels, same order: bitwise equal.
"""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_graph_replays_match_eager(torch_cuda, prec):
    import torch
    from x3div_amd import X3D
    nx, ny, nz = 64, 48, 32
    print(f"\n[graph {prec} {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, 1.0, 2.0, 3.0)
    orc = Oracle(nx, ny, nz, 1.0, 2.0, 3.0)
    dt = torch.float64 if prec == "f64" else torch.float32
    u0 = [torch.from_numpy(a).to(dt).cuda() for a in orc.fields(1, 99, 1e-3)]
    mk = ctx.field if prec == "f64" else ctx.field32
    eager = [t.clone() for t in u0] + [mk() for _ in range(7)]
    graph = [t.clone() for t in u0] + [mk() for _ in range(7)]
    step = ctx.step if prec == "f64" else ctx.step_f32
    if prec == "f64":
        ctx.graph_capture_step(*graph, gdt=1e-3)
    else:
        ctx.graph_capture_step_f32(*graph, gdt=1e-3)
    torch.cuda.synchronize()
    # capture records only: the fields are untouched
    for a, b in zip(graph[:3], u0):
        assert torch.equal(a, b)
    for rep in range(3):
        step(*eager, gdt=1e-3)
        ctx.graph_launch()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(graph, eager)):
            assert torch.equal(a, b), (rep, i)
        junk = mk()
        ctx.derx(junk, eager[0]) if prec == "f64" else ctx.derx_f32(junk, eager[0])
        torch.cuda.synchronize()
    ctx.close()


def test_operator_set_releases_captured_step(torch_cuda):
    """Replacing an operator's tables replaces the device tables a captured
    step reads: the graph is released and its launch refused.  Steps on
    caller tables run the unfused composition, which is not captured; the
    eager step with the caller's tables still matches the oracle.  The
    library's own tables handed back (x3doperator1d%init with them) change
    nothing."""
    import torch
    from x3div_amd import X3D, X3DError
    nx, ny, nz = 32, 24, 16
    ctx = X3D(nx, ny, nz, 1.0, 2.0, 3.0)
    orc = Oracle(nx, ny, nz, 1.0, 2.0, 3.0)
    u0 = orc.fields(1, 5, 1e-3)
    f = [torch.from_numpy(a).cuda() for a in u0] + [ctx.field() for _ in range(7)]
    g = [t.clone() for t in f]
    ctx.graph_capture_step(*g, gdt=1e-3)
    ctx.graph_launch()
    torch.cuda.synchronize()
    fs, s, w, p, alfa, _ = orc.op("derx")
    # the library's own tables, bit for bit: nothing changes, the graph stays
    ctx.set_operator("derx", fs, s, w, p, alfa)
    ctx.graph_launch()
    torch.cuda.synchronize()
    # a caller's tables (one ulp off): the graph is released
    ctx.set_operator("derx", np.nextafter(fs, 2 * fs), s, w, p, alfa)
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.graph_launch()
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.graph_capture_step(*g, gdt=1e-3)
    ctx.step(*f, gdt=1e-3)
    torch.cuda.synchronize()
    ru, rdu, _, _ = orc.step(*u0, gdt=1e-3)
    for a, b in zip(f[:6], ru + rdu):
        a = a.cpu().numpy()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    ctx.close()


def test_graph_begin_end_records_the_reference_calls(torch_cuda):
    """The calls a reference-shaped driver makes for one step
    (src/xcompact3d.f90:83-99), recorded between x3d_graph_begin and
    x3d_graph_end, replay equal to the same calls made eagerly, bit for bit."""
    import torch
    from x3div_amd import X3D
    nx, ny, nz = 48, 32, 40
    ctx = X3D(nx, ny, nz, 1.0, 2.0, 3.0)
    orc = Oracle(nx, ny, nz, 1.0, 2.0, 3.0)
    u0 = [torch.from_numpy(a).cuda() for a in orc.fields(1, 17, 1e-3)]
    g = [t.clone() for t in u0] + [ctx.field() for _ in range(7)]
    e = [t.clone() for t in u0] + [ctx.field() for _ in range(7)]

    def calls(f):
        ux, uy, uz, dux, duy, duz, pp3, px, py, pz = f
        ctx.calculate_transeq_rhs(dux, duy, duz, ux, uy, uz)
        ctx.int_time(ux, uy, uz, dux, duy, duz, gdt=1e-3)
        ctx.solve_poisson(pp3, px, py, pz, ux, uy, uz)
        ctx.cor_vel(ux, uy, uz, px, py, pz)

    ctx.graph_begin()
    calls(g)
    ctx.graph_end()
    torch.cuda.synchronize()
    for a, b in zip(g[:3], u0):
        assert torch.equal(a, b)  # recorded, not run
    for rep in range(3):
        calls(e)
        ctx.graph_launch()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(g, e)):
            assert torch.equal(a, b), (rep, i)
    ctx.close()


def test_graph_begin_refuses_a_synchronising_call(torch_cuda):
    """A reduction needs a host value: recorded, it fails, and x3d_graph_end
    then fails and keeps no graph."""
    from x3div_amd import X3D, X3DError
    ctx = X3D(16, 16, 16)
    f = ctx.field(fill=1.0)
    ctx.graph_begin()
    with pytest.raises(X3DError):
        ctx.div_stats(f)
    with pytest.raises(X3DError):
        ctx.graph_end()
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.graph_launch()
    # the context works eagerly afterwards
    mx, mean = ctx.div_stats(f)
    assert mx == 1.0 and mean == 1.0
    ctx.close()


def test_fill_kernel(torch_cuda):
    import torch
    from x3div_amd import X3D
    ctx = X3D(8, 8, 8)
    t = ctx.field(fill=np.nan)
    ctx.fill(t, 2.5)
    torch.cuda.synchronize()
    assert bool((t == 2.5).all())
    ctx.close()
