"""The dense-circulant matrix-core variant against the FP64 oracle.

Every periodic operator is applied as C = A^-1 B on v_mfma_f32_32x32x16_f16
with three split-precision (fp16 hi/lo) terms and FP32 accumulation, on the
precision form's FP32 fields; the reference is the FP64 operator on the same
FP32 input.

Error bound.  Each product C[i][k] u[k] is formed from 22-bit halves of both
operands with the lo*lo term dropped: at most 2^-21 of |C[i][k]| |u[k]|.  The
FP32 accumulator adds 2^-24 of the running sum per step, and the running
sums stay within sum_k |C[i][k]| |u[k]|.  Summed over a row, both stay below
2^-20 ||C||_1 max|u|, where ||C||_1 = sum_d |c_d| is the operator's row sum
(its impulse response, taken from the oracle).  The test holds every
operator to that bound relative to max|out|.  A derivative's ||C||_1 grows
as n / L (about 160 at n = 512 on 2 pi), so the relative error does too:
measured about 1e-5 at n = 384-512, against bounds of 3e-5 and more.  A
wrong fragment layout or circulant offset gives errors of order one.  Line
lengths 128, 384 and 512 cover one, three and four 32-row blocks per
wavefront.
"""
import numpy as np
import pytest

from oracle import OPS, Oracle

pytestmark = pytest.mark.gpu

GRIDS = [(128, 128, 128, 2 * np.pi, 2 * np.pi, 2 * np.pi), (256, 128, 384, 1.0, 2.0, 3.0),
         (512, 128, 512, 2 * np.pi, 1.0, 2 * np.pi), (128, 512, 128, 1.0, 2 * np.pi, 1.0)]


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


def row_norm(orc, name):
    """||C||_1: the sum of |c_d| over the operator's impulse response, from
    the oracle on a line grid of the same length and spacing."""
    axis = "xyz".index(name[3])
    dims = [8, 8, 8]
    dims[axis] = (orc.nx, orc.ny, orc.nz)[axis]
    line = Oracle(*dims, *orc.L)
    d = np.zeros(line.shape)
    d[0, 0, 0] = 1.0
    idx = [0, 0, 0]
    idx[2 - axis] = slice(None)
    return float(np.abs(line.apply(name, d)[tuple(idx)]).sum())


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
    bound = 2.0 ** -20 * row_norm(orc, name) * np.abs(u).max() / np.abs(ref).max()
    print(f"{name}: {err:.2e} (bound {bound:.2e})")
    assert err < bound, (name, err, bound)


def test_circulant_mfma_rejects_untileable(torch_cuda):
    from x3div_amd import X3D, X3DError
    ctx = X3D(96, 64, 64)
    a, b = ctx.field32(), ctx.field32()
    with pytest.raises(X3DError, match="UNSUPPORTED"):
        ctx.apply_circ_mfma("derx", a, b)
    ctx.close()
