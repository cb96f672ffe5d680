"""GPU parity of the BFP32 form (block-floating-point int32 fields, FP64
stencils and line solves).

* encode and decode are bit-exact against the numpy model of the format
  (tests/bfp_model.py, pinned by test_bfp_model.py);
* each operator against the FP64 oracle applied to the decoded input: the
  kernels compute in FP64, so the only difference is the rounding of the
  output word, at most 2^-29 of its segment's largest value (plus 1e-12 of
  the field's maximum for the FP64 solve);
* the step against the oracle's step on the decoded fields: the stored
  intermediates (du, tmp, the divergence and gradient stages, the pressure)
  are rounded the same way; bars stated per field below;
* at 512^3, each operator on the BFP32 words of an FP64 field against the FP64
  operator on the unrounded field: this includes the input's rounding, and is
  held to 1e-6 of the result's maximum (FP32 storage gives 5-9e-6 there).
"""
import numpy as np
import pytest

import bfp_model as bm
from oracle import Oracle

pytestmark = pytest.mark.gpu

OPS_ALL = ["derx", "derxp", "dery", "deryp", "derz", "derzp",
           "derxvp", "deryvp", "derzvp", "derxpv", "derypv", "derzpv",
           "intxvp", "intyvp", "intzvp", "intxpv", "intypv", "intzpv",
           "derxx", "deryy", "derzz"]

# nx multiple of 16; chunk lengths 1-16 along x, y and z
GRIDS = [
    (64, 48, 32, 1.0, 2.0, 3.0),
    (32, 96, 160, 2 * np.pi, 1.0, 0.5),
    (16, 8, 24, 1.0, 1.0, 1.0),
    (128, 40, 16, 1.0, 2 * np.pi, 2.0),
    (512, 16, 8, 2 * np.pi, 1.0, 1.0),
    (16, 512, 64, 1.0, 2 * np.pi, 1.0),
]


def dev(a, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
def grid(request, torch_cuda):
    from x3div_amd import X3D
    nx, ny, nz, lx, ly, lz = request.param
    print(f"\n[bfp grid {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, lx, ly, lz)
    orc = Oracle(nx, ny, nz, lx, ly, lz)
    yield ctx, orc
    ctx.close()


def _inputs(orc):
    rng = np.random.default_rng(9)
    smooth = orc.fields(2, 11, 0.0)[0]
    noise = rng.standard_normal(orc.shape) * np.ldexp(1.0, rng.integers(-8, 8, orc.shape[:2] + (1,)))
    return {"smooth": smooth, "noise": noise}


def test_encode_decode_bit_exact(grid):
    ctx, orc = grid
    import torch
    for kind, u in _inputs(orc).items():
        u = u.copy()
        u[0, 0, :16] = 0.0
        q = ctx.bfp_empty()
        ctx.bfp_encode(q, dev(u))
        v = ctx.field()
        ctx.bfp_decode(v, q)
        torch.cuda.synchronize()
        w = bm.encode(u)
        assert np.array_equal(host(q), w), kind
        assert np.array_equal(host(v), bm.decode(w)), kind


@pytest.mark.parametrize("name", OPS_ALL)
def test_operator_bfp(grid, name):
    ctx, orc = grid
    import torch
    for kind, u in _inputs(orc).items():
        w = bm.encode(u)
        ref = orc.apply(name, bm.decode(w))
        out = ctx.bfp_empty()
        ctx.apply_bfp(name, out, dev(w))
        torch.cuda.synchronize()
        got = bm.decode(host(out))
        bound = np.ldexp(bm.segment_max(ref), -29) + 1e-12 * np.abs(ref).max()
        err = np.abs(got - ref)
        assert np.all(err <= bound), (name, kind, float((err / bound).max()))


@pytest.mark.parametrize("gdt", [0.0, 1e-3])
def test_step_bfp(grid, gdt):
    """Bars, relative to each field's max (measured worst case on these grids
    in brackets).  The step stores its intermediates as BFP32 (tmp, the
    divergence and gradient stages, the pressure), each rounded to 2^-29 of
    its segment's largest value, and later derivatives amplify that by up to
    about n / pi (160 on the 512-point lines):
      * du 1e-8 [2.4e-9]: its own rounding and that of tmp, not amplified;
      * pressure gradient 2e-6 [4.8e-7]: rounding of the interpolated
        pressure stages, amplified by the last derivative;
      * u 5e-7 [1.6e-7]: its rounding plus the gradient's error carried by
        the velocity correction;
      * pressure 1e-6 [2.1e-7, on the unit cube where the Taylor-Green part
        is discretely solenoidal and the pressure comes from the noise alone;
        1e-9 elsewhere]: the divergence cancels its O(1) terms to O(1e-3) or
        less, so the rounding of its stages weighs that much more.
    The FP32 form's bars on the same fields are 2e-6, 3e-3, 2e-4 and 2e-3."""
    ctx, orc = grid
    import torch
    u0 = [bm.encode(a) for a in orc.fields(1, 1234, 1e-3)]
    ref_u, ref_du, ref_pp, ref_g = orc.step(*(bm.decode(a) for a in u0), gdt=gdt)
    u = [dev(a) for a in u0]
    rest = [ctx.bfp_empty() for _ in range(7)]
    ctx.step_bfp(*u, *rest, gdt=gdt)
    torch.cuda.synchronize()
    du, pp, g = rest[:3], rest[3], rest[4:]

    def rel(a, b):
        return float(np.abs(bm.decode(host(a)) - b).max() / max(np.abs(b).max(), 1e-300))
    err = {"u": max(rel(a, b) for a, b in zip(u, ref_u)),
           "du": max(rel(a, b) for a, b in zip(du, ref_du)),
           "pp": rel(pp, ref_pp),
           "gradp": max(rel(a, b) for a, b in zip(g, ref_g))}
    print("step errors", ctx.nx, ctx.ny, ctx.nz, gdt, err)
    assert err["du"] < 1e-8 and err["gradp"] < 2e-6, err
    assert err["u"] < 5e-7 and err["pp"] < 1e-6, err


def test_step_timed_bfp_matches_step(grid):
    ctx, orc = grid
    import torch
    u0 = [bm.encode(a) for a in orc.fields(1, 77, 1e-3)]
    outs = []
    for timed in (False, True):
        f = [dev(a) for a in u0] + [ctx.bfp_empty() for _ in range(7)]
        if timed:
            passes = ctx.step_timed_bfp(*f, gdt=1e-3)
            assert len(passes) == 10 and all(ms >= 0 for ms, _ in passes.values())
        else:
            ctx.step_bfp(*f, gdt=1e-3)
        torch.cuda.synchronize()
        outs.append(f)
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_bfp_refuses_unaligned_x(torch_cuda):
    from x3div_amd import X3D, X3DError
    ctx = X3D(24, 16, 16)
    q = ctx.bfp_empty()
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.bfp_encode(q, ctx.field())
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.apply_bfp("derx", ctx.bfp_empty(), q)
    ctx.close()


def test_precision_512_cubed_vs_fp64(torch_cuda):
    """Every operator at 512^3 on the BFP32 words of a Taylor-Green + 1e-3
    noise field (2 pi cube), against the FP64 operator on the unrounded FP64
    field: within 1e-6 of the result's maximum."""
    import torch
    from x3div_amd import X3D
    n = 512
    print(f"\n[bfp precision {n}^3]", flush=True)
    L = 2 * np.pi
    ctx = X3D(n, n, n, L, L, L)
    u = ctx.field()
    ctx.init_fields(u, ctx.field(), ctx.field(), kind=1, seed=5, amp=1e-3)
    q = ctx.bfp_empty()
    ctx.bfp_encode(q, u)
    ref, out, dec = ctx.field(), ctx.bfp_empty(), ctx.field()
    worst = {}
    for name in OPS_ALL:
        ctx.apply(name, ref, u)
        ctx.apply_bfp(name, out, q)
        ctx.bfp_decode(dec, out)
        torch.cuda.synchronize()
        worst[name] = float((dec - ref).abs().max() / ref.abs().max())
    print(worst)
    assert max(worst.values()) < 1e-6, worst
    ctx.close()
    torch.cuda.empty_cache()
