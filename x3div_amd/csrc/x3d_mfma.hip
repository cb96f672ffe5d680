// Dense-circulant variant of the periodic operators on the matrix cores.
//
// With periodic boundaries every operator of src/x3d_operator_1d.f90 is a
// circulant matrix C = A^-1 B (A the tridiagonal left-hand side, B the
// stencil); column j of C is column 0 shifted by j, so C[i][j] = c[(i-j) mod n]
// with c = A^-1 (B e_0).  Applying it to all lines of a direction is a GEMM
// with K = n:
//   y/z lines:  OUT[i][col] = sum_k C[i][k] IN[k][col]   (IN viewed as n x N)
//   x lines:    OUT[l][i]   = sum_k IN[l][k] C[i][k]     (IN viewed as L x n)
// computed with v_mfma_f32_32x32x16_f16 in three split-precision terms,
// x = x_hi + x_lo (fp16 each, together ~22 bits):
//   C_hi IN_hi + C_hi IN_lo + C_lo IN_hi, accumulated in FP32.
// fp16 keeps three more bits per half than bf16 but has a narrow exponent
// range, so C is scaled by a power of two on the host and each work-group
// scales its input block by the power of two that brings its largest value
// just under 2^14 (exact scalings, undone on the FP32 result).
// Inputs and outputs are FP32 fields of the precision form.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "x3d_line.h"

namespace x3d {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// hi/lo fp16 halves of eight FP32 values
__device__ __forceinline__ void split8(const float (&x)[8], f16x8& hi, f16x8& lo)
{
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const _Float16 h = (_Float16)x[v];
    hi[v] = h;
    lo[v] = (_Float16)(x[v] - (float)h);
  }
}

// One fragment of C for the 32x32x16 operand layout (lane l: row l % 32,
// k = 8 (l / 32) + v): hi and lo halves of 8 values, indexed by the block
// offset d = (row0 - k0) mod n, d a multiple of 16.
struct CircFrag {
  f16x8 hi, lo;
};

// y/z: OUT = C IN, IN rows k at stride rs, WG columns [col0, col0 + 32) of
// slab blockIdx.y (slab stride ss); each wavefront owns RB blocks of 32 rows.
// x: OUT = IN C^T, IN lines [32 bx, 32 bx + 32) of n contiguous points; each
// wavefront owns RB blocks of 32 output points.
// STAGE (n <= 512): the work-group's 32 lines of input are read once, with
// 16-byte loads along memory order (x: the 32 lines are one contiguous run;
// y/z: 32-column rows of 128 B), into an LDS block of 32 rows of n + 4
// floats, keeping the block maximum on the way; the MFMA operands then come
// from LDS.  Otherwise the input is read twice from memory, once for the
// maximum and once per MFMA step.
template <int RB, bool XDIR, bool STAGE>
__global__ __launch_bounds__(256) void k_circ_mfma(const float* __restrict__ in, float* __restrict__ out,
                                                   const CircFrag* __restrict__ tab, int n, long long rs,
                                                   long long ss, float inv_cscale)
{
  __shared__ float wmax[4];
  // the operator's fragments (n / 16 x 64 lanes x 32 B) are read by every
  // wavefront at every step: staged once per work-group in LDS
  extern __shared__ unsigned char smem_raw[];
  CircFrag* ltab = reinterpret_cast<CircFrag*>(smem_raw);
  float* blk = reinterpret_cast<float*>(smem_raw + (size_t)(n >> 4) * 64 * sizeof(CircFrag));
  const int P = n + 4;  // block row pitch in floats (16-byte multiple)
  for (int q = threadIdx.x; q < (n >> 4) * 64; q += 256) ltab[q] = tab[q];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, lk8 = 8 * (lane >> 5), lk = 4 * (lane >> 5);
  long long base;
  if constexpr (XDIR) base = ((long long)blockIdx.x * 32 + lc) * n;  // this lane's line
  else base = (long long)blockIdx.y * ss + (long long)blockIdx.x * 32 + lc;  // this lane's column
  // the block's largest magnitude
  float m = 0.f;
  if constexpr (STAGE) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    if constexpr (XDIR) {
      const f32x4* src = reinterpret_cast<const f32x4*>(in + (long long)blockIdx.x * 32 * n);
      for (int q = threadIdx.x; q < 8 * n; q += 256) {
        const f32x4 x = src[q];
        const int l = (4 * q) / n, k = 4 * q - l * n;
        *reinterpret_cast<f32x4*>(blk + l * P + k) = x;
        m = fmaxf(m, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
      }
    } else {
      const float* slab = in + (long long)blockIdx.y * ss + (long long)blockIdx.x * 32;
      for (int q = threadIdx.x; q < 8 * n; q += 256) {
        const int k = q >> 3, c4 = 4 * (q & 7);
        const f32x4 x = *reinterpret_cast<const f32x4*>(slab + (long long)k * rs + c4);
        blk[(c4 + 0) * P + k] = x.x;
        blk[(c4 + 1) * P + k] = x.y;
        blk[(c4 + 2) * P + k] = x.z;
        blk[(c4 + 3) * P + k] = x.w;
        m = fmaxf(m, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
      }
    }
  } else {
    // thread (lc, g) scans points g, g + 8, ...
    const int g = threadIdx.x >> 5;
    for (int k = g; k < n; k += 8) {
      const float x = XDIR ? in[base + k] : in[base + (long long)k * rs];
      m = fmaxf(m, fabsf(x));
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
  if (lane == 0) wmax[wave] = m;
  __syncthreads();
  m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  int e = 0;
  if (m > 0.f) (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
  const float scale = ldexpf(1.f, 14 - e), unscale = ldexpf(inv_cscale, e - 14);
  f32x16 acc[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[r][v] = 0.f;
  for (int k0 = 0; k0 < n; k0 += 16) {
    float x[8];
    if constexpr (STAGE) {
      const float* row = blk + lc * P + k0 + lk8;
#pragma unroll
      for (int v = 0; v < 8; ++v) x[v] = row[v] * scale;
    } else if constexpr (XDIR) {
#pragma unroll
      for (int v = 0; v < 8; ++v) x[v] = in[base + k0 + lk8 + v] * scale;
    } else {
#pragma unroll
      for (int v = 0; v < 8; ++v) x[v] = in[base + (long long)(k0 + lk8 + v) * rs] * scale;
    }
    f16x8 fh, fl;
    split8(x, fh, fl);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int b0 = 32 * (wave * RB + r);
      int d = b0 - k0;
      d = d < 0 ? d + n : d;
      const CircFrag cf = ltab[(d >> 4) * 64 + lane];
      if constexpr (XDIR) {
        // D[line][i] = sum_k IN[line][k] C^T[k][i]
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fh, cf.hi, acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fl, cf.hi, acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fh, cf.lo, acc[r], 0, 0, 0);
      } else {
        // D[i][col] = sum_k C[i][k] IN[k][col]
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cf.hi, fh, acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cf.hi, fl, acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cf.lo, fh, acc[r], 0, 0, 0);
      }
    }
  }
  // D layout: column lc; row 8 (v / 4) + 4 (lane / 32) + v % 4
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const int b0 = 32 * (wave * RB + r);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = 8 * (v >> 2) + lk + (v & 3);
      if constexpr (XDIR) {
        // row = line within the block, column = output point b0 + lc
        out[((long long)blockIdx.x * 32 + row) * n + b0 + lc] = acc[r][v] * unscale;
      } else {
        out[base + (long long)(b0 + row) * rs] = acc[r][v] * unscale;
      }
    }
  }
}

// ---- host ----------------------------------------------------------------------
// First column of C = A^-1 B: the stencil applied to the periodic delta at 0,
// then the reference's periodic Thomas sweeps and Sherman-Morrison correction
// with the operator's own tables (src/thomas.f90:34-84).
static std::vector<double> circ_column(const x3d_ctx* c, int op, int dir, int kind)
{
  const OpHost& o = c->op[op];
  const int n = o.n;
  const Coef<double> k = stencil_coef(c->sc[dir], kind);
  auto e = [n](int i) { i %= n; if (i < 0) i += n; return i == 0 ? 1.0 : 0.0; };
  std::vector<double> t(n);
  for (int i = 0; i < n; ++i) {
    switch (kind) {
      case K_D: t[i] = k.a * (e(i + 1) - e(i - 1)) + k.b * (e(i + 2) - e(i - 2)); break;
      case K_DVP: t[i] = k.a * (e(i + 1) - e(i)) + k.b * (e(i + 2) - e(i - 1)); break;
      case K_DPV: t[i] = k.a * (e(i) - e(i - 1)) + k.b * (e(i + 1) - e(i - 2)); break;
      case K_IVP:
        t[i] = k.a * (e(i + 1) + e(i)) + k.b * (e(i + 2) + e(i - 1)) + k.c * (e(i + 3) + e(i - 2)) +
               k.d * (e(i + 4) + e(i - 3));
        break;
      default:
        t[i] = k.a * (e(i) + e(i - 1)) + k.b * (e(i + 1) + e(i - 2)) + k.c * (e(i + 2) + e(i - 3)) +
               k.d * (e(i + 3) + e(i - 4));
        break;
    }
  }
  for (int i = 1; i < n; ++i) t[i] = t[i] - t[i - 1] * o.s[i];
  t[n - 1] = t[n - 1] * o.w[n - 1];
  for (int i = n - 2; i >= 0; --i) t[i] = (t[i] - o.f[i] * t[i + 1]) * o.w[i];
  const double ss = (t[0] - o.alfa * t[n - 1]) / (1.0 + o.p[0] - o.alfa * o.p[n - 1]);
  for (int i = 0; i < n; ++i) t[i] = t[i] - ss * o.p[i];
  return t;
}

}  // namespace x3d

using namespace x3d;

extern "C" int x3d_apply_circ_mfma(x3d_ctx* c, int op, float* out, const float* in)
{
  if (!c || op < 0 || op >= X3D_NOPS || !out || !in || out == in) return X3D_ERR_ARG;
  if (!c->op[op].circulant) return X3D_ERR_UNSUPPORTED;
  const int dir = op < 6 ? op / 2 : (op - 6) % 3;
  static const int kinds[4] = {K_DVP, K_DPV, K_IVP, K_IPV};
  const int kind = op < 6 ? K_D : kinds[(op - 6) / 3];
  if (stretch_which(c, dir, kind) >= 0) return X3D_ERR_UNSUPPORTED;  // stretched y-derivatives: the line solvers
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  // the kernel's tiles: 4 wavefronts x RB blocks of 32 along the line, and
  // 32 columns (y/z) or 32 lines (x) per work-group
  const int cols = dir == 0 ? c->ny * c->nz : dir == 1 ? c->nx : c->nx * c->ny;
  if (n % 128 || n > 1024 || cols % 32) return X3D_ERR_UNSUPPORTED;
  if (!c->circ_tab[op]) {
    const std::vector<double> col = circ_column(c, op, dir, kind);
    // scale C by the power of two that brings its largest entry under 2^10
    double cmax = 0.0;
    for (double v : col) cmax = std::fmax(cmax, std::fabs(v));
    int ce = 0;
    if (cmax > 0.0) (void)std::frexp(cmax, &ce);
    const double cs = std::ldexp(1.0, 10 - ce);
    c->circ_scale[op] = (float)std::ldexp(1.0, ce - 10);
    std::vector<CircFrag> h((size_t)(n / 16) * 64);
    for (int d = 0; d < n; d += 16)
      for (int l = 0; l < 64; ++l) {
        CircFrag f;
        for (int v = 0; v < 8; ++v) {
          int idx = (d + (l & 31) - 8 * (l >> 5) - v) % n;
          if (idx < 0) idx += n;
          const double x = col[idx] * cs;
          const _Float16 hi = (_Float16)x;
          f.hi[v] = hi;
          f.lo[v] = (_Float16)(x - (double)hi);
        }
        h[(size_t)(d / 16) * 64 + l] = f;
      }
    void* p = nullptr;
    X3D_CHECK_HIP(hipMalloc(&p, h.size() * sizeof(CircFrag)));
    const hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(CircFrag), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_status(e);
    }
    c->circ_tab[op] = p;
  }
  const CircFrag* tab = static_cast<const CircFrag*>(c->circ_tab[op]);
  const long long nxy = (long long)c->nx * c->ny;
  dim3 grid;
  long long rs = 0, ss = 0;
  if (dir == 0) {
    grid = dim3((unsigned)(cols / 32));
  } else if (dir == 1) {
    grid = dim3((unsigned)(c->nx / 32), (unsigned)c->nz);
    rs = c->nx;
    ss = nxy;
  } else {
    grid = dim3((unsigned)(nxy / 32));
    rs = nxy;
  }
  const int RB = n / 128;
  const float inv = c->circ_scale[op];  // undoes the host scaling of C
  const size_t tsmem = (size_t)(n / 16) * 64 * sizeof(CircFrag);  // 128 n bytes
  // lines of up to 512 points: the input block too (32 rows of n + 4 floats)
  const bool stage = n <= 512;
  const size_t smem = tsmem + (stage ? (size_t)32 * (n + 4) * sizeof(float) : 0);
#define X3D_RB(R)                                                                                         \
  case R:                                                                                                 \
    if (stage) {                                                                                          \
      (void)hipFuncSetAttribute((const void*)k_circ_mfma<R, true, true>,                                  \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);                   \
      (void)hipFuncSetAttribute((const void*)k_circ_mfma<R, false, true>,                                 \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);                   \
      if (dir == 0)                                                                                       \
        hipLaunchKernelGGL((k_circ_mfma<R, true, true>), grid, dim3(256), smem, c->stream, in, out, tab, n, rs, ss, inv); \
      else                                                                                                \
        hipLaunchKernelGGL((k_circ_mfma<R, false, true>), grid, dim3(256), smem, c->stream, in, out, tab, n, rs, ss, inv); \
    } else {                                                                                              \
      if (smem > 65536) {                                                                                 \
        (void)hipFuncSetAttribute((const void*)k_circ_mfma<R, true, false>,                               \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);                 \
        (void)hipFuncSetAttribute((const void*)k_circ_mfma<R, false, false>,                              \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);                 \
      }                                                                                                   \
      if (dir == 0)                                                                                       \
        hipLaunchKernelGGL((k_circ_mfma<R, true, false>), grid, dim3(256), smem, c->stream, in, out, tab, n, rs, ss, inv); \
      else                                                                                                \
        hipLaunchKernelGGL((k_circ_mfma<R, false, false>), grid, dim3(256), smem, c->stream, in, out, tab, n, rs, ss, inv); \
    }                                                                                                     \
    break;
  switch (RB) {
    X3D_RB(1) X3D_RB(2) X3D_RB(3) X3D_RB(4) X3D_RB(5) X3D_RB(6) X3D_RB(7) X3D_RB(8)
    default: return X3D_ERR_UNSUPPORTED;
  }
#undef X3D_RB
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}
