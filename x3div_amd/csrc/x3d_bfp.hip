// BFP32 form: block-floating-point fields (4 bytes per point, the word format
// of Store<bfp32> in x3d_line.h), FP64 stencils and line solves.
//
// Rounding a 512^3 field to FP32 moves its compact derivatives by 5-9e-6 of
// their maximum: the operators amplify the rounding (2^-24 of each value) by
// up to ~n/pi.  A BFP32 word keeps 31 bits against the largest value of its
// 16-point x-segment, 2^-30 (at most 2^-29) of it, so the same derivatives
// stay near 1e-7 of FP64 while every pass moves as many bytes as FP32.
//
// Entry points: conversion from and to FP64, every operator, and the step
// (the fused passes on BFP32 fields; the Poisson solve runs in FP64, its x
// line kernels reading and writing the BFP32 words, or on a decoded copy of
// the right-hand side whose result is encoded back where those kernels do
// not apply).
#include "x3d_line.h"

namespace x3d {

// Elementwise over the field in 256-thread blocks: a 16-lane row holds 16
// consecutive points, one segment (the point count is a multiple of 16, so
// every row leaves the loop together).
__global__ __launch_bounds__(256) void k_bfp_encode(const double* __restrict__ in, int* __restrict__ out, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = bfp_word(in[i]);
}

__global__ __launch_bounds__(256) void k_bfp_decode(const int* __restrict__ in, double* __restrict__ out, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = bfp_value(in[i]);
}

static int bfp_check(const x3d_ctx* c)
{
  if (c->nx % 16) return X3D_ERR_UNSUPPORTED;
  // 32-bit byte offsets of the line kernels
  if ((unsigned long long)c->nx * c->ny * c->nz * 4ull > 0xFFFFFFFFull) return X3D_ERR_UNSUPPORTED;
  return X3D_OK;
}

static unsigned ew_blocks(long long n)
{
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? b : 65536);
}

static int bfp_encode(x3d_ctx* c, const double* in, bfp32* out)
{
  const long long n = (long long)c->nx * c->ny * c->nz;
  hipLaunchKernelGGL(k_bfp_encode, dim3(ew_blocks(n)), dim3(256), 0, c->stream, in, reinterpret_cast<int*>(out), n);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

static int bfp_decode(x3d_ctx* c, const bfp32* in, double* out)
{
  const long long n = (long long)c->nx * c->ny * c->nz;
  hipLaunchKernelGGL(k_bfp_decode, dim3(ew_blocks(n)), dim3(256), 0, c->stream, reinterpret_cast<const int*>(in),
                     out, n);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

}  // namespace x3d

using namespace x3d;

#define X3D_TRY(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

int x3d_bfp_encode(x3d_ctx* c, int32_t* q, const double* u)
{
  if (!c || !q || !u) return X3D_ERR_ARG;
  X3D_TRY(bfp_check(c));
  return bfp_encode(c, u, reinterpret_cast<bfp32*>(q));
}

int x3d_bfp_decode(x3d_ctx* c, double* u, const int32_t* q)
{
  if (!c || !q || !u) return X3D_ERR_ARG;
  X3D_TRY(bfp_check(c));
  return bfp_decode(c, reinterpret_cast<const bfp32*>(q), u);
}

int x3d_apply_bfp(x3d_ctx* c, int op, int32_t* out, const int32_t* in)
{
  if (!c || op < 0 || op >= X3D_NOPS_ALL || !out || !in || out == in) return X3D_ERR_ARG;
  X3D_TRY(bfp_check(c));
  // operator ids as x3d_apply_f32
  const int dir = op < 6 ? op / 2 : (op - 6) % 3;
  static const int kinds[5] = {K_DVP, K_DPV, K_IVP, K_IPV, K_DD};
  const int kind = op < 6 ? K_D : kinds[(op - 6) / 3];
  return launch_line_bfp(c, dir, kind, op, reinterpret_cast<const bfp32*>(in), reinterpret_cast<bfp32*>(out));
}

int x3d_step_bfp(x3d_ctx* c, double gdt, int32_t* ux, int32_t* uy, int32_t* uz, int32_t* dux, int32_t* duy,
                 int32_t* duz, int32_t* pp3, int32_t* px, int32_t* py, int32_t* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  X3D_TRY(bfp_check(c));
  if (!fused_ok(c) || !all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  auto B = [](int32_t* p) { return reinterpret_cast<bfp32*>(p); };
  X3D_TRY(fused_momentum<bfp32>(c, gdt, 1, B(ux), B(uy), B(uz), B(dux), B(duy), B(duz)));
  X3D_TRY(fused_divergence<bfp32>(c, B(pp3), B(ux), B(uy), B(uz)));
  if (c->xsplit) {
    // the x line kernels of Poisson decode and encode the words themselves
    X3D_TRY(poisson_solve_bfp(c, pp3));
  } else {
    // the work arrays are free between the divergence and the gradient passes
    X3D_TRY(ensure_work(c));
    double* rhs = c->work[0];
    X3D_TRY(bfp_decode(c, B(pp3), rhs));
    X3D_TRY(poisson_solve(c, rhs));
    X3D_TRY(bfp_encode(c, rhs, B(pp3)));
  }
  X3D_TRY(mark_pass(c, 7));
  X3D_TRY(fused_gradp<bfp32>(c, B(px), B(py), B(pz), B(pp3), B(ux), B(uy), B(uz), 1));
  return X3D_OK;
}

int x3d_step_timed_bfp(x3d_ctx* c, double gdt, int32_t* ux, int32_t* uy, int32_t* uz, int32_t* dux, int32_t* duy,
                       int32_t* duz, int32_t* pp3, int32_t* px, int32_t* py, int32_t* pz, float* pass_ms)
{
  if (!c || !pass_ms) return X3D_ERR_ARG;
  for (int i = 0; i <= X3D_NPASS; ++i)
    if (!c->pass_ev[i]) X3D_CHECK_HIP(hipEventCreate(&c->pass_ev[i]));
  X3D_CHECK_HIP(hipEventRecord(c->pass_ev[0], c->stream));
  c->timing = true;
  const int rc = x3d_step_bfp(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz);
  c->timing = false;
  if (rc != X3D_OK) return rc;
  X3D_CHECK_HIP(hipEventSynchronize(c->pass_ev[X3D_NPASS]));
  for (int i = 0; i < X3D_NPASS; ++i)
    X3D_CHECK_HIP(hipEventElapsedTime(&pass_ms[i], c->pass_ev[i], c->pass_ev[i + 1]));
  return X3D_OK;
}
