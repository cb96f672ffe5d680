"""The dense-circulant matrix-core variant against the FP64 oracle.

Every periodic operator is applied as C = A^-1 B on v_mfma_f32_32x32x16_f16
with three split-precision (fp16 hi/lo) terms and FP32 accumulation.  The
dropped lo*lo term and the fp16 rounding of the lo halves are each about
2^-22 relative per product; a derivative of a smooth field cancels terms of
size |C| |u| down to its own size, which costs up to n / (2 pi) of that, so
the bound is 2e-5 of the output's max norm (measured: 9e-6 for derivatives,
7e-7 for interpolations) against the FP64 operator on the same FP32 input.
A wrong fragment layout or circulant offset gives errors of order one.  Line
lengths 128 and 384 cover one and three 32-row blocks per wavefront.
"""
import numpy as np
import pytest

from oracle import OPS, Oracle

pytestmark = pytest.mark.gpu

GRIDS = [(128, 128, 128, 2 * np.pi, 2 * np.pi, 2 * np.pi), (256, 128, 384, 1.0, 2.0, 3.0)]


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
def grid(request, torch_cuda):
    from x3div_amd import X3D
    nx, ny, nz, lx, ly, lz = request.param
    print(f"\n[mfma grid {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, lx, ly, lz)
    orc = Oracle(nx, ny, nz, lx, ly, lz)
    u = np.asarray(orc.fields(2, 5, 0.0)[0], dtype=np.float32)
    yield ctx, orc, u
    ctx.close()


@pytest.mark.parametrize("name", OPS)
def test_circulant_mfma_operator(grid, name):
    import torch
    ctx, orc, u = grid
    ref = orc.apply(name, u.astype(np.float64))
    out = ctx.field32()
    ctx.apply_circ_mfma(name, out, torch.from_numpy(u).cuda())
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{name}: {err:.2e}")
    assert err < 2e-5, (name, err)


def test_circulant_mfma_rejects_untileable(torch_cuda):
    from x3div_amd import X3D, X3DError
    ctx = X3D(96, 64, 64)
    a, b = ctx.field32(), ctx.field32()
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.apply_circ_mfma("derx", a, b)
    ctx.close()
