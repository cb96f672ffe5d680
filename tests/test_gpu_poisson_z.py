"""Poisson with the fused z stage (the x r2c and forward y DFT as line
kernels, or hipFFT's 2-D plan, then one kernel for the forward z DFT, the
division and the inverse z DFT) against the oracle, for
every z length the kernel is built for (radix-8 passes with a final radix-4
or radix-2 pass, then radix-3 and radix-5 passes for 3- and 5-smooth
lengths), in both precisions.  FP64 at the 1e-12 bar,
FP32 at 2e-6 (the precision form's Poisson bound).

Lines of up to 512 points run the persistent, pipelined z kernel (each
work-group walks several tiles and prefetches the next one); X3D_PZ_PLAIN
selects the one-tile-per-work-group kernel.  Both evaluate the same
expressions in the same order; the compiler may fuse different products into
multiply-adds in the two, so on grids with many tiles per work-group their
results must agree to a few units of rounding (1e-14 FP64, 1e-6 FP32,
relative to the max norm)."""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu

# powers of two (radix-8 passes, then radix-4 or radix-2) and 3- and 5-smooth
# lengths (radix-3 and radix-5 passes after those)
NZ = [8, 16, 32, 64, 128, 256, 512, 1024, 24, 40, 48, 96, 160, 384, 768]


@pytest.mark.parametrize("nz", NZ)
@pytest.mark.parametrize("xline", [False, True], ids=["x-hipfft", "x-line"])
def test_poisson_fused_z(torch_cuda, nz, xline):
    """xline: an x length the x r2c line kernel takes (then the forward y
    stage is a line kernel too); else hipFFT's 2-D forward plan."""
    import torch
    from x3div_amd import X3D
    nx, ny = (24, 16) if nz <= 256 else (12, 8)
    if xline:
        nx = 32 if nz <= 256 else 16
    print(f"\n[poisson z {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, 1.0, 2.0, 3.0)
    orc = Oracle(nx, ny, nz, 1.0, 2.0, 3.0)
    rhs = np.random.default_rng(nz).standard_normal(orc.shape)
    ref = orc.poisson(rhs)
    t = torch.from_numpy(rhs.copy()).cuda()
    ctx.poisson(t)
    torch.cuda.synchronize()
    e64 = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
    r32 = np.asarray(rhs, dtype=np.float32)
    ref32 = orc.poisson(r32.astype(np.float64))
    t32 = torch.from_numpy(r32.copy()).cuda()
    ctx.poisson_f32(t32)
    torch.cuda.synchronize()
    e32 = np.abs(t32.cpu().numpy().astype(np.float64) - ref32).max() / np.abs(ref32).max()
    print(f"nz={nz}: fp64 {e64:.2e} fp32 {e32:.2e}")
    ctx.close()
    assert e64 < 1e-12, e64
    assert e32 < 2e-6, e32


@pytest.mark.parametrize("ny", [8, 32, 128, 512, 1024, 24, 48, 80, 96, 160, 192, 320, 768])
@pytest.mark.parametrize("xline", [False, True], ids=["x-hipfft", "x-line"])
def test_poisson_fused_z_y_lengths(torch_cuda, ny, xline):
    """The fused z stage and the y stages for several y lengths: the forward
    y stage as hipFFT's 2-D plan (with its x r2c) or as the y-line kernel
    after the x r2c line kernel (xline), the inverse y stage as the y-line
    kernel where the length allows."""
    import torch
    from x3div_amd import X3D
    nx, nz = (20, 16) if ny <= 128 else (10, 8)
    if xline:
        nx = 32 if ny <= 128 else 16
    print(f"\n[poisson y {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, 2.0, 1.0, 3.0)
    orc = Oracle(nx, ny, nz, 2.0, 1.0, 3.0)
    rhs = np.random.default_rng(ny).standard_normal(orc.shape)
    ref = orc.poisson(rhs)
    t = torch.from_numpy(rhs.copy()).cuda()
    ctx.poisson(t)
    r32 = np.asarray(rhs, dtype=np.float32)
    ref32 = orc.poisson(r32.astype(np.float64))
    t32 = torch.from_numpy(r32.copy()).cuda()
    ctx.poisson_f32(t32)
    torch.cuda.synchronize()
    e64 = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
    e32 = np.abs(t32.cpu().numpy().astype(np.float64) - ref32).max() / np.abs(ref32).max()
    print(f"ny={ny}: fp64 {e64:.2e} fp32 {e32:.2e}")
    ctx.close()
    assert e64 < 1e-12, e64
    assert e32 < 2e-6, e32


# nx = 2M with M a line-kernel length: the forward x r2c runs as the x-line
# kernel (one length-M complex DFT per line and the split into the half
# spectrum), then the forward y-line kernel; other nx keep hipFFT's 2-D plan
NX = [16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 768, 1024, 24, 20]


@pytest.mark.parametrize("nx", NX)
@pytest.mark.parametrize("xsplit", ["1", "0"])
def test_poisson_x_lengths(torch_cuda, nx, xsplit, monkeypatch):
    """Every x length of the x r2c line kernel, with and without it
    (X3D_XSPLIT=0: hipFFT's 2-D forward plan), both precisions."""
    import torch
    from x3div_amd import X3D
    monkeypatch.setenv("X3D_XSPLIT", xsplit)
    ny, nz = (24, 16) if nx <= 256 else (16, 8)
    print(f"\n[poisson x {nx}x{ny}x{nz} xsplit={xsplit}]", flush=True)
    ctx = X3D(nx, ny, nz, 3.0, 1.0, 2.0)
    orc = Oracle(nx, ny, nz, 3.0, 1.0, 2.0)
    rhs = np.random.default_rng(nx).standard_normal(orc.shape)
    ref = orc.poisson(rhs)
    t = torch.from_numpy(rhs.copy()).cuda()
    ctx.poisson(t)
    r32 = np.asarray(rhs, dtype=np.float32)
    ref32 = orc.poisson(r32.astype(np.float64))
    t32 = torch.from_numpy(r32.copy()).cuda()
    ctx.poisson_f32(t32)
    torch.cuda.synchronize()
    e64 = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
    e32 = np.abs(t32.cpu().numpy().astype(np.float64) - ref32).max() / np.abs(ref32).max()
    print(f"nx={nx}: fp64 {e64:.2e} fp32 {e32:.2e}")
    ctx.close()
    assert e64 < 1e-12, e64
    assert e32 < 2e-6, e32


@pytest.mark.parametrize("shape", [(512, 256, 64), (64, 96, 512)])
def test_poisson_pipelined_z_many_tiles(torch_cuda, shape, monkeypatch):
    """Grids with more tiles than resident work-groups: the pipelined kernel
    against the oracle (FP64) and against the plain kernel."""
    import torch
    from x3div_amd import X3D
    nx, ny, nz = shape
    print(f"\n[poisson pipelined {nx}x{ny}x{nz}]", flush=True)
    ctx = X3D(nx, ny, nz, 1.0, 2.0, 3.0)
    rhs = np.random.default_rng(nx + nz).standard_normal((nz, ny, nx))
    outs = {}
    for plain in (False, True):
        if plain:
            monkeypatch.setenv("X3D_PZ_PLAIN", "1")
        else:
            monkeypatch.delenv("X3D_PZ_PLAIN", raising=False)
        t = torch.from_numpy(rhs.copy()).cuda()
        ctx.poisson(t)
        t32 = torch.from_numpy(rhs.astype(np.float32)).cuda()
        ctx.poisson_f32(t32)
        torch.cuda.synchronize()
        outs[plain] = (t.cpu().numpy(), t32.cpu().numpy())
    monkeypatch.delenv("X3D_PZ_PLAIN", raising=False)
    ctx.close()
    d64 = np.abs(outs[False][0] - outs[True][0]).max() / np.abs(outs[True][0]).max()
    d32 = np.abs(outs[False][1].astype(np.float64) - outs[True][1]).max() / np.abs(outs[True][1]).max()
    print(f"{shape}: pipelined vs plain fp64 {d64:.2e} fp32 {d32:.2e}")
    assert d64 < 1e-14, d64
    assert d32 < 1e-6, d32
    orc = Oracle(nx, ny, nz, 1.0, 2.0, 3.0)
    ref = orc.poisson(rhs)
    e64 = np.abs(outs[False][0] - ref).max() / np.abs(ref).max()
    print(f"{shape}: fp64 {e64:.2e}")
    assert e64 < 1e-12, e64
