"""x3d_step_noalloc: the step with its scratch fields in the library's static
device array, called through the C ABI and compared with x3d_step.

The same kernels run with the same arguments in the same order, so the results
must be equal bit for bit.  Two calls in a row share the array, the second on a
smaller grid, which then starts from what the first left there.  On a fresh
context the call must not take device memory (x3d_step's first call allocates
its 8 work fields, 128 MiB at 128^3, which the same measurement sees), and a
grid above X3D_NOALLOC_MAX_POINTS is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21  # X3D_NOALLOC_MAX_POINTS in include/x3dhip.h


def _fields(ctx, seed):
    u = [ctx.field() for _ in range(3)]
    ctx.init_fields(*u, kind=1, seed=seed, amp=1e-3)
    return u + [ctx.field(fill=np.nan) for _ in range(7)]


@pytest.mark.parametrize("dims", [(128, 128, 128), (48, 40, 32)], ids=["128cube", "48x40x32"])
def test_noalloc_step_equals_step(torch_cuda, dims):
    import torch
    from x3div_amd import X3D
    nx, ny, nz = dims
    print(f"\n[noalloc step {nx}x{ny}x{nz}]", flush=True)
    assert nx * ny * nz <= MAX_POINTS
    ctx = X3D(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    a = _fields(ctx, 31)
    b = [t.clone() for t in a]
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(2):
        ctx.step_noalloc(*a, gdt=1e-3)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    for _ in range(2):
        ctx.step(*b, gdt=1e-3)
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    for i, (x, y) in enumerate(zip(a, b)):
        assert torch.equal(x, y), i
    work = 8 * 8 * nx * ny * nz
    print(f"  free memory: noalloc step {free0 - free1} B, first x3d_step {free1 - free2} B "
          f"(its work fields {work} B)", flush=True)
    assert free0 - free1 < (8 << 20), free0 - free1
    if work >= (64 << 20):
        assert free1 - free2 >= work // 2, free1 - free2
    ctx.close()


def test_noalloc_step_refuses_larger_grids(torch_cuda):
    from x3div_amd import X3D, X3DError
    ctx = X3D(160, 128, 128, 1.0, 1.0, 1.0)
    f = [ctx.field() for _ in range(10)]
    with pytest.raises(X3DError, match="SIZE"):
        ctx.step_noalloc(*f, gdt=0.0)
    ctx.close()
