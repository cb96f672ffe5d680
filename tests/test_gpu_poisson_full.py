"""Poisson at the benchmark size (512^3), checked by an identity that does not
depend on the grid size or on the FFT: the staggered compact operators give
div(grad p) the Fourier symbol -K, the same K the spectral division uses
(src/poisson.f90:673-695 with src/navier.f90:98-359), so for a field p0
without null modes (zero mean, no Nyquist content)

    poisson(divergence(gradp(p0))) == p0

to rounding.  The line passes and the FFT stages are independent code, so a
wrong transform stage at full size (hipFFT's batched x c2r, the y- and z-line
kernels) shows up as an O(1) error.  This direction divides the line
operators' rounding by K; the other one (div grad of a Poisson solution)
would multiply white rounding by K, up to ~1e5 at 512^3.  Bounds: FP64
through the drop-in entry points 1e-12 (measured 4e-15), FP32 through the
single operators composed in the same way 2e-5 (measured 2e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 512
MODES = [  # (amplitude, kx, ky, kz, phase) on a 2 pi periodic box
    (1.0, 1, 2, 3, 0.3),
    (0.5, 5, 1, 2, 1.1),
    (0.25, 3, 7, 1, 2.0),
    (0.125, 17, 4, 9, 0.7),
]


def smooth_field(dtype):
    import torch
    x = torch.arange(N, dtype=torch.float64, device="cuda") * (2 * np.pi / N)
    f = torch.zeros((N, N, N), dtype=torch.float64, device="cuda")
    for a, kx, ky, kz, ph in MODES:
        # cos(kx x + ky y + kz z + ph), built from separable factors
        cx, sx = torch.cos(kx * x + ph), torch.sin(kx * x + ph)
        cy, sy = torch.cos(ky * x), torch.sin(ky * x)
        cz, sz = torch.cos(kz * x), torch.sin(kz * x)
        cxy = cx[None, None, :] * cy[None, :, None] - sx[None, None, :] * sy[None, :, None]
        sxy = sx[None, None, :] * cy[None, :, None] + cx[None, None, :] * sy[None, :, None]
        f += a * (cxy * cz[:, None, None] - sxy * sz[:, None, None])
    return f.to(dtype).contiguous()


def relerr(a, b):
    return float((a.double() - b.double()).abs().max() / b.double().abs().max())


def test_poisson_inverts_div_grad_fp64_512(torch_cuda):
    import torch
    from x3div_amd import X3D
    print(f"\n[poisson identity fp64 {N}^3]", flush=True)
    ctx = X3D(N, N, N, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    p0 = smooth_field(torch.float64)
    px, py, pz, r = (ctx.field() for _ in range(4))
    ctx.gradp(px, py, pz, p0)
    ctx.divergence(r, px, py, pz)
    ctx.poisson(r)
    torch.cuda.synchronize()
    e = relerr(r, p0)
    print(f"fp64 poisson(div(grad(p0))) vs p0: {e:.2e}")
    del p0, px, py, pz, r
    ctx.close()
    torch.cuda.empty_cache()
    assert e < 1e-12, e


def test_poisson_inverts_div_grad_fp32_512(torch_cuda):
    import torch
    from x3div_amd import X3D
    print(f"\n[poisson identity fp32 {N}^3]", flush=True)
    ctx = X3D(N, N, N, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    p0 = smooth_field(torch.float32)
    t1, t2 = ctx.field32(), ctx.field32()

    def chain(names, src, out):
        # apply the operators right to left: names[-1] first
        cur = src
        for k, name in enumerate(reversed(names)):
            dst = out if k == len(names) - 1 else (t1 if cur is not t1 else t2)
            ctx.apply_f32(name, dst, cur)
            cur = dst
        return out

    g = [ctx.field32() for _ in range(3)]
    chain(["derxpv", "intypv", "intzpv"], p0, g[0])
    chain(["intxpv", "derypv", "intzpv"], p0, g[1])
    chain(["intxpv", "intypv", "derzpv"], p0, g[2])
    d = ctx.field32()
    r = ctx.field32()
    chain(["derxvp", "intyvp", "intzvp"], g[0], d)
    r.copy_(d)
    chain(["intxvp", "deryvp", "intzvp"], g[1], d)
    r += d
    chain(["intxvp", "intyvp", "derzvp"], g[2], d)
    r += d
    ctx.poisson_f32(r)
    torch.cuda.synchronize()
    e = relerr(r, p0)
    print(f"fp32 poisson(div(grad(p0))) vs p0: {e:.2e}")
    del p0, g, d, r, t1, t2
    ctx.close()
    torch.cuda.empty_cache()
    assert e < 2e-5, e
