"""GPU parity of the precision form (FP32 fields and line solves, FP64 spectral
division) against the FP64 CPU oracle.

Bounds, relative to the max norm of the FP64 result (measured worst cases on
these grids in brackets):
  * one derivative (D, staggered D) on a smooth field or white noise: 1e-6
    (4e-7); the FP32 round-off of the solve and the output;
  * one mid-point interpolation: 5e-6 (3e-6): alfa = 0.49 makes its
    tridiagonal system ~100 times worse conditioned than the derivative's;
  * Poisson alone: 2e-6 (8e-7);
  * one step: velocity 2e-4 (5e-5) and its right-hand side 2e-6 (8e-7);
    pressure 2e-3 in the max norm (7e-4) and 1e-3 in L2 (5e-4), its gradient
    3e-3 (1e-3).  The pressure comes from the divergence of a nearly
    solenoidal field (Taylor-Green + 1e-3 noise), a sum of O(1) terms that
    cancel to O(1e-3), so FP32 keeps about three fewer digits there.  (With
    equal x and y lengths the Taylor-Green part is discretely solenoidal and
    the pressure, from the noise alone, is 1e3 times smaller and that much
    worse conditioned; the grids avoid it.)
The grids cover every FP32 chunk length the plans use at these sizes,
including 24- and 32-point chunks (lines of 768 and 512 points, 512 in each
direction) and even x chunks (two points per access).
"""
import numpy as np
import pytest

from oracle import OPS, Oracle

pytestmark = pytest.mark.gpu

GRIDS = [
    (64, 48, 32, 1.0, 2.0, 3.0),
    (40, 96, 160, 2 * np.pi, 1.0, 0.5),
    (16, 8, 24, 1.0, 1.0, 1.0),
    (24, 512, 8, 1.0, 2 * np.pi, 1.0),
    (16, 8, 768, 1.0, 1.0, 2 * np.pi),
    (128, 24, 16, 1.0, 1.0, 2.0),
    (512, 8, 16, 2 * np.pi, 1.0, 1.0),
    (8, 16, 512, 1.0, 2.0, 2 * np.pi),
]


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def relerr2(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
def grid(request, torch_cuda):
    from x3div_amd import X3D
    nx, ny, nz, lx, ly, lz = request.param
    print(f"\n[fp32 grid {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, lx, ly, lz)
    orc = Oracle(nx, ny, nz, lx, ly, lz)
    yield ctx, orc
    ctx.close()


def dev32(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    return t.detach().cpu().numpy().astype(np.float64)


def smooth(orc, seed):
    return orc.fields(2, seed, 0.0)[0]


@pytest.mark.parametrize("name", OPS)
def test_operator_fp32(grid, name):
    ctx, orc = grid
    import torch
    bound = 5e-6 if name.startswith("int") else 1e-6
    for kind, inp in (("smooth", smooth(orc, 11)),
                      ("noise", np.random.default_rng(5).standard_normal(orc.shape))):
        u32 = np.asarray(inp, dtype=np.float32)
        ref = orc.apply(name, u32.astype(np.float64))
        out = ctx.field32()
        ctx.apply_f32(name, out, dev32(u32))
        torch.cuda.synchronize()
        e = relerr(host(out), ref)
        assert e < bound, (name, kind, e)


@pytest.mark.parametrize("name", OPS)
def test_operator_fp32_vs_fp64_input(grid, name):
    """The FP32 form against the FP64 operator on the unrounded FP64 field,
    input rounding included.  Rounding moves each input by at most 2^-24 of
    its magnitude, so the output moves by at most 2^-24 max|u| ||Op||_inf,
    where ||Op||_inf is the sum of |c_j| over the operator's circulant column
    (the oracle applied to a unit vector), plus the FP32 arithmetic of the
    test above.  On 512-point lines that bound is 5-9e-6 of the result's
    maximum, above 1e-6; the BFP32 form (tests/test_gpu_bfp.py) is held to
    1e-6 at 512^3 with the same reference."""
    ctx, orc = grid
    import torch
    n = orc.shape[::-1]["xyz".index(name[3])]
    # the response to a unit impulse is one circulant column along the line
    e = np.zeros(orc.shape)
    e[0, 0, 0] = 1.0
    norm = np.abs(orc.apply(name, e)).sum()
    arith = 5e-6 if name.startswith("int") else 1e-6
    for kind, inp in (("smooth", smooth(orc, 11)),
                      ("noise", np.random.default_rng(5).standard_normal(orc.shape))):
        ref = orc.apply(name, inp)
        out = ctx.field32()
        ctx.apply_f32(name, out, dev32(inp))
        torch.cuda.synchronize()
        err = np.abs(host(out) - ref).max()
        bound = 2.0 ** -24 * np.abs(inp).max() * norm + arith * np.abs(ref).max()
        assert err <= bound, (name, kind, n, err / np.abs(ref).max(), bound / np.abs(ref).max())


def test_derivative_entry_points_match_apply(grid):
    ctx, orc = grid
    import torch
    u = dev32(smooth(orc, 3))
    for fn, name in ((ctx.derx_f32, "derx"), (ctx.dery_f32, "dery"), (ctx.derz_f32, "derz")):
        a, b = ctx.field32(), ctx.field32()
        fn(a, u)
        ctx.apply_f32(name, b, u)
        torch.cuda.synchronize()
        assert torch.equal(a, b), name


def test_poisson_fp32(grid):
    ctx, orc = grid
    import torch
    rhs = smooth(orc, 21)
    rhs32 = np.asarray(rhs, dtype=np.float32)
    ref = orc.poisson(rhs32.astype(np.float64))
    t = dev32(rhs32)
    ctx.poisson_f32(t)
    torch.cuda.synchronize()
    assert relerr(host(t), ref) < 2e-6


@pytest.mark.parametrize("gdt", [0.0, 1e-3])
def test_step_fp32(grid, gdt):
    ctx, orc = grid
    import torch
    u0 = [np.asarray(a, dtype=np.float32) for a in orc.fields(1, 1234, 1e-3)]
    ref_u, ref_du, ref_pp, ref_g = orc.step(*(a.astype(np.float64) for a in u0), gdt=gdt)
    u = [dev32(a) for a in u0]
    du = [ctx.field32() for _ in range(3)]
    pp = ctx.field32()
    g = [ctx.field32() for _ in range(3)]
    ctx.step_f32(*u, *du, pp, *g, gdt=gdt)
    torch.cuda.synchronize()
    for a, b in zip(u, ref_u):
        assert relerr(host(a), b) < 2e-4
    for a, b in zip(du, ref_du):
        assert relerr(host(a), b) < 2e-6
    assert relerr(host(pp), ref_pp) < 2e-3
    assert relerr2(host(pp), ref_pp) < 1e-3
    for a, b in zip(g, ref_g):
        assert relerr(host(a), b) < 3e-3


def test_step_timed_fp32_matches_step(grid):
    ctx, orc = grid
    import torch
    u0 = [np.asarray(a, dtype=np.float32) for a in orc.fields(1, 77, 1e-3)]
    outs = []
    for timed in (False, True):
        f = [dev32(a) for a in u0] + [ctx.field32() for _ in range(7)]
        if timed:
            passes = ctx.step_timed_f32(*f, gdt=1e-3)
            assert len(passes) == 10 and all(ms >= 0 for ms, _ in passes.values())
        else:
            ctx.step_f32(*f, gdt=1e-3)
        torch.cuda.synchronize()
        outs.append(f)
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_large_grid_fp32_only(torch_cuda):
    """1024 x 768 x 768: FP32 fields fit the kernels' 32-bit offsets, FP64 ones
    (4.8 GB) do not, so the FP64 entry points refuse and FP32 runs.  Checked
    by the closed form: derx of sin(2 pi x / L) is (2 pi / L) lambda cos, with
    lambda the scheme's modified-wavenumber ratio, here within FP32 round-off
    of the exact derivative."""
    import torch
    from x3div_amd import X3D, X3DError
    nx, ny, nz = 1024, 768, 768
    print(f"\n[fp32 large {nx}x{ny}x{nz}]", flush=True)
    L = 2 * np.pi
    ctx = X3D(nx, ny, nz, L, L, L)
    x = torch.arange(nx, dtype=torch.float64, device="cuda") * (L / nx)
    u = torch.sin(x).float().expand(nz, ny, nx).contiguous()
    t = ctx.field32()
    ctx.derx_f32(t, u)
    torch.cuda.synchronize()
    ref = torch.cos(x).float()
    err = float((t - ref).abs().max())
    assert err < 2e-5, err
    u64 = torch.zeros((nz, ny, nx), dtype=torch.float64, device="cuda")
    t64 = torch.zeros_like(u64)
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.derx(t64, u64)
    del u64, t64
    ctx.close()
    torch.cuda.empty_cache()
