"""GPU parity of y-stretching (x3d_set_stretching: istret /= 0 with the
vectors ppy and ppyi, src/module_param.f90:70-71,104) against the oracle,
whose stretching is pinned in tests/test_oracle_stretch.py.

The reference multiplies dery's result by ppy(j) (src/x3d_derive.f90:331-337,
427-433), deryvp's by ppyi(j) and derypv's by ppy(j)
(src/x3d_staggered.f90:628-634, 874-880).  Here:
  * every operator, FP64, at 1e-12 of the result's maximum (x, z and the
    interpolations unchanged), and the explicit vector argument of the
    reference's routines;
  * the FP32 and BFP32 operators at their forms' bars;
  * the non-periodic dery_ij (a free-slip/Dirichlet y direction);
  * the fused step pieces and the step in all three forms (FP64 at 1e-12;
    FP32 and BFP32 at the bars of tests/test_gpu_precision.py and
    tests/test_gpu_bfp.py), and a captured step's replays;
  * what a stretched context refuses: the y second derivative and a step
    graph captured before the change.
"""
import numpy as np
import pytest

import bfp_model as bm
import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
OPS = ["derx", "derxp", "dery", "deryp", "derz", "derzp",
       "derxvp", "deryvp", "derzvp", "derxpv", "derypv", "derzpv",
       "intxvp", "intyvp", "intzvp", "intxpv", "intypv", "intzpv"]
Y_SCALED = ("dery", "deryp", "deryvp", "derypv")

# nx a multiple of 16 (BFP32); y lines of 48, 64 and 512 points
GRIDS = [
    (32, 48, 24, 1.0, 2.0, 1.5),
    (64, 64, 64, 2 * np.pi, 2 * np.pi, 2 * np.pi),
    (16, 512, 32, 1.0, 2 * np.pi, 1.0),
]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vectors(ny, seed):
    """A smooth stretching and its values at the staggered points."""
    r = np.random.default_rng(seed)
    a, ph = 0.3 + 0.2 * r.random(), 2 * np.pi * r.random()
    ppy = 1.0 / (1.0 + a * np.cos(2 * np.pi * np.arange(ny) / ny + ph))
    ppyi = 1.0 / (1.0 + a * np.cos(2 * np.pi * (np.arange(ny) + 0.5) / ny + ph))
    return ppy, ppyi


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
def grid(request, torch_cuda):
    from x3div_amd import X3D
    nx, ny, nz, lx, ly, lz = request.param
    print(f"\n[stretch grid {nx}x{ny}x{nz}]", flush=True)
    ppy, ppyi = vectors(ny, nx + ny + nz)
    ctx = X3D(nx, ny, nz, lx, ly, lz)
    ctx.set_stretching(1, ppy, ppyi)
    orc = oracle.Oracle(nx, ny, nz, lx, ly, lz)
    orc.set_stretching(1, ppy, ppyi)
    yield ctx, orc, ppy, ppyi
    ctx.close()


def test_operators_fp64(grid):
    import torch
    ctx, orc, ppy, ppyi = grid
    u = np.random.default_rng(5).standard_normal(orc.shape)
    ud = dev(u)
    for name in OPS:
        t = ctx.field(fill=np.nan)
        ctx.apply(name, t, ud)
        torch.cuda.synchronize()
        assert rel(host(t), orc.apply(name, u)) < RTOL, name


def test_explicit_vector_argument(grid):
    """dery/deryvp/derypv scale by the vector they are given, as the
    reference's routines do; a different one replaces the context's."""
    import torch
    from x3div_amd import X3D
    ctx, orc, ppy, ppyi = grid
    u = np.random.default_rng(6).standard_normal(orc.shape)
    other = X3D(orc.nx, orc.ny, orc.nz, *orc.L)
    plain = oracle.Oracle(orc.nx, orc.ny, orc.nz, *orc.L)
    try:
        q, qi = vectors(orc.ny, 99)
        other.set_stretching(1, ppy, ppyi)
        for name, routine, vec in (("dery", other.dery, q), ("deryvp", other.deryvp, qi),
                                   ("derypv", other.derypv, q)):
            t = other.field(fill=np.nan)
            if name == "dery":
                routine(t, dev(u), name, ppy=vec)
            elif name == "deryvp":
                routine(t, dev(u), name, ppyi=vec)
            else:
                routine(t, dev(u), name, ppy=vec)
            torch.cuda.synchronize()
            ref = plain.apply(name, u) * vec[None, :, None]
            assert rel(host(t), ref) < RTOL, name
        # istret = 0: the vectors are ignored
        other.set_stretching(0)
        t = other.field(fill=np.nan)
        other.dery(t, dev(u), "dery", ppy=q)
        torch.cuda.synchronize()
        assert rel(host(t), plain.apply("dery", u)) < RTOL
    finally:
        other.close()


def test_operators_fp32(grid):
    """Against the oracle on the FP32-rounded input (the FP32 form's operator
    bars, tests/test_gpu_precision.py): the scaling is one more rounding."""
    import torch
    ctx, orc, ppy, ppyi = grid
    u32 = np.random.default_rng(7).standard_normal(orc.shape).astype(np.float32)
    for name in OPS:
        t = ctx.field32()
        ctx.apply_f32(name, t, dev(u32))
        torch.cuda.synchronize()
        bar = 5e-6 if name.startswith("int") else 1e-6
        assert rel(host(t).astype(np.float64), orc.apply(name, u32.astype(np.float64))) < bar, name
    t = ctx.field32()
    ctx.dery_f32(t, dev(u32), "dery")
    torch.cuda.synchronize()
    assert rel(host(t).astype(np.float64), orc.apply("dery", u32.astype(np.float64))) < 1e-6


def test_operators_bfp(grid):
    """BFP32: the line kernel's word (2^-29 of its segment's largest value)
    is decoded, scaled and stored again (another 2^-29 of the scaled
    segment's; a segment is 16 x-points of one y-row, one factor): 2^-28 of
    the segment maximum, plus 1e-12 of the field's for the FP64 solve."""
    import torch
    ctx, orc, ppy, ppyi = grid
    w = bm.encode(np.random.default_rng(8).standard_normal(orc.shape))
    u = bm.decode(w)
    for name in OPS:
        t = ctx.bfp_empty()
        ctx.apply_bfp(name, t, dev(w))
        torch.cuda.synchronize()
        ref = orc.apply(name, u)
        bound = np.ldexp(bm.segment_max(ref), -28) + 1e-12 * np.abs(ref).max()
        err = np.abs(bm.decode(host(t)) - ref)
        assert np.all(err <= bound), (name, float((err / bound).max()))


def test_refusals(grid):
    import torch
    from x3div_amd._lib import X3DError
    ctx, orc, ppy, ppyi = grid
    u = dev(np.random.default_rng(9).standard_normal(orc.shape))
    with pytest.raises(X3DError):
        ctx.deryy(ctx.field(), u)
    t = ctx.field()
    ctx.derxx(t, u)  # x and z second derivatives are unaffected
    torch.cuda.synchronize()


def test_step_pieces_fp64(grid):
    import torch
    ctx, orc, ppy, ppyi = grid
    u = orc.fields(1, 99, 1e-3)
    d = [ctx.field(fill=np.nan) for _ in range(3)]
    ctx.calculate_transeq_rhs(*d, *(dev(a) for a in u))
    pp = ctx.field(fill=np.nan)
    ctx.divergence(pp, *(dev(a) for a in u))
    p = np.random.default_rng(13).standard_normal(orc.shape)
    g = [ctx.field(fill=np.nan) for _ in range(3)]
    ctx.gradp(*g, dev(p))
    torch.cuda.synchronize()
    for a, b in zip(d, orc.momentum_rhs(*u)):
        assert rel(host(a), b) < RTOL
    assert rel(host(pp), orc.divergence(*u)) < RTOL
    for a, b in zip(g, orc.gradp(p)):
        assert rel(host(a), b) < RTOL


@pytest.mark.parametrize("gdt", [0.0, 1e-3])
def test_step_fp64(grid, gdt):
    """Bars as the unstretched step parity (tests/test_gpu_parity.py)."""
    import torch
    ctx, orc, ppy, ppyi = grid
    u0 = orc.fields(1, 2024, 1e-3)
    ru, rdu, rpp, rg = orc.step(*u0, gdt=gdt)
    u = [dev(a) for a in u0]
    du = [ctx.field() for _ in range(3)]
    pp = ctx.field(); g = [ctx.field() for _ in range(3)]
    ctx.step(*u, *du, pp, *g, gdt=gdt)
    torch.cuda.synchronize()
    for a, b in zip(u + du, ru + rdu):
        assert rel(host(a), b) < RTOL
    assert rel2(host(pp), rpp) < RTOL
    assert rel(host(pp), rpp) < 10 * RTOL
    for a, b in zip(g, rg):
        assert rel(host(a), b) < 10 * RTOL


def test_step_fp32(grid):
    import torch
    ctx, orc, ppy, ppyi = grid
    u0 = [np.asarray(a, dtype=np.float32) for a in orc.fields(1, 1234, 1e-3)]
    ref_u, ref_du, ref_pp, ref_g = orc.step(*(a.astype(np.float64) for a in u0), gdt=1e-3)
    u = [dev(a) for a in u0]
    du = [ctx.field32() for _ in range(3)]
    pp = ctx.field32()
    g = [ctx.field32() for _ in range(3)]
    ctx.step_f32(*u, *du, pp, *g, gdt=1e-3)
    torch.cuda.synchronize()
    for a, b in zip(u, ref_u):
        assert rel(host(a), b) < 2e-4
    for a, b in zip(du, ref_du):
        assert rel(host(a), b) < 2e-6
    assert rel(host(pp), ref_pp) < 2e-3
    for a, b in zip(g, ref_g):
        assert rel(host(a), b) < 3e-3


def test_step_bfp(grid):
    import torch
    ctx, orc, ppy, ppyi = grid
    u0 = [bm.encode(a) for a in orc.fields(1, 1234, 1e-3)]
    ref_u, ref_du, ref_pp, ref_g = orc.step(*(bm.decode(a) for a in u0), gdt=1e-3)
    u = [dev(a) for a in u0]
    rest = [ctx.bfp_empty() for _ in range(7)]
    ctx.step_bfp(*u, *rest, gdt=1e-3)
    torch.cuda.synchronize()
    du, pp, g = rest[:3], rest[3], rest[4:]

    def r(a, b):
        return rel(bm.decode(host(a)), b)
    err = {"u": max(r(a, b) for a, b in zip(u, ref_u)),
           "du": max(r(a, b) for a, b in zip(du, ref_du)),
           "pp": r(pp, ref_pp),
           "gradp": max(r(a, b) for a, b in zip(g, ref_g))}
    print("stretched bfp step errors", err)
    assert err["du"] < 1e-8 and err["gradp"] < 2e-6, err
    assert err["u"] < 5e-7 and err["pp"] < 1e-6, err


def test_captured_step_and_release(grid):
    """A step captured on a stretched context replays to the eager result;
    changing the stretching releases the graph (its launch is refused)."""
    import torch
    from x3div_amd._lib import X3DError
    ctx, orc, ppy, ppyi = grid
    u0 = orc.fields(1, 77, 1e-3)
    a = [dev(x) for x in u0]
    da = [ctx.field() for _ in range(3)]; pa = ctx.field(); ga = [ctx.field() for _ in range(3)]
    ctx.step(*a, *da, pa, *ga, gdt=1e-3)
    b = [dev(x) for x in u0]
    db = [ctx.field() for _ in range(3)]; pb = ctx.field(); gb = [ctx.field() for _ in range(3)]
    ctx.graph_capture_step(*b, *db, pb, *gb, gdt=1e-3)
    for rep in range(3):
        for x, y in zip(b, u0):
            x.copy_(dev(y))
        ctx.graph_launch()
        torch.cuda.synchronize()
        for x, y in zip(a + da + [pa] + ga, b + db + [pb] + gb):
            assert torch.equal(x, y), rep
    ctx.set_stretching(1, ppy, ppyi)
    with pytest.raises(X3DError):
        ctx.graph_launch()


def test_non_periodic_y(torch_cuda):
    """dery_ij on a free-slip / Dirichlet y direction, then ppy."""
    import torch
    from x3div_amd import X3D
    nx, ny, nz = 16, 33, 12
    ppy, ppyi = vectors(ny, 3)
    for bc in ((1, 2), (2, 2)):
        ctx = X3D(nx, ny, nz, 1.0, 2.0, 1.0, ncl1=(0, bc[0], 0), ncln=(0, bc[1], 0))
        try:
            ctx.set_stretching(1, ppy, ppyi)
            u = np.random.default_rng(11).standard_normal((nz, ny, nx))
            for npaire, name in ((0, "dery"), (1, "deryp")):
                t = ctx.field(fill=np.nan)
                ctx.dery(t, dev(u), name)
                torch.cuda.synchronize()
                ref = oracle.der_np(u, "y", 2.0, bc[0], bc[1], npaire) * ppy[None, :, None]
                assert rel(host(t), ref) < RTOL, (bc, name)
        finally:
            ctx.close()


def test_batch_refuses_a_stretched_grid(torch_cuda):
    """The batched passes carry no stretching vectors: a batch whose grid
    context was given stretching refuses to step (and steps again once the
    stretching is turned off)."""
    import ctypes as C
    import torch
    from x3div_amd import X3DBatch
    from x3div_amd._lib import X3DError
    b = X3DBatch([(16, 16, 16), (32, 16, 8)])
    try:
        b.step(0.0)
        torch.cuda.synchronize()
        h = b._ctx(1)
        ppy, ppyi = vectors(16, 5)
        a, ai = np.ascontiguousarray(ppy), np.ascontiguousarray(ppyi)
        assert b.lib.x3d_set_stretching(h, 1, a.ctypes.data, ai.ctypes.data) == 0
        with pytest.raises(X3DError):
            b.step(0.0)
        assert b.lib.x3d_set_stretching(h, 0, None, None) == 0
        b.step(0.0)
        torch.cuda.synchronize()
    finally:
        b.close()


def test_batch_refuses_caller_tables(torch_cuda):
    """The batched passes run the circulant solvers of the library's own
    tables: a batch whose grid context was given tables of the caller's own
    (x3d_operator_set) refuses to step and to replay its graph, and steps
    again once the library's tables are set back."""
    import ctypes as C
    import torch
    from x3div_amd import X3DBatch
    from x3div_amd._lib import X3DError
    from x3div_amd.context import _op_id
    b = X3DBatch([(16, 16, 16), (32, 16, 8)])
    try:
        b.step(0.0)
        torch.cuda.synchronize()
        b.graph_capture(0.0)
        b.graph_launch()
        torch.cuda.synchronize()
        h, op = b._ctx(0), _op_id("dery")
        hp = C.POINTER(C.c_double)
        f, s, w, p = (np.zeros(16) for _ in range(4))
        alfa, n, npaire = C.c_double(), C.c_int(), C.c_int()
        assert b.lib.x3d_operator_get(h, op, *(a.ctypes.data_as(hp) for a in (f, s, w, p)),
                                      C.byref(alfa), C.byref(n), C.byref(npaire)) == 0
        assert n.value == 16
        f2 = f * (1 + 1e-3)
        assert b.lib.x3d_operator_set(h, op, *(a.ctypes.data_as(hp) for a in (f2, s, w, p)), alfa.value) == 0
        with pytest.raises(X3DError):
            b.step(0.0)
        with pytest.raises(X3DError):
            b.graph_launch()
        assert b.lib.x3d_operator_set(h, op, *(a.ctypes.data_as(hp) for a in (f, s, w, p)), alfa.value) == 0
        b.step(0.0)
        b.graph_capture(0.0)
        b.graph_launch()
        torch.cuda.synchronize()
    finally:
        b.close()
