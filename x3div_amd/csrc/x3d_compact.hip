// Single-operator periodic line kernels: stencil right-hand side, periodic
// tridiagonal solve and store, one pass over each line (one read of the
// input window, one write).  Drop-ins for derx_00/dery_00/derz_00, the twelve
// staggered routines and xthomas_0/ythomas_0/zthomas_0.
#include <cmath>

#include <cstdio>
#include <cstdlib>

#include "x3d_line.h"

namespace x3d {

double circ_gamma(double alfa) { return 0.5 * (1.0 + std::sqrt(1.0 - 4.0 * alfa * alfa)); }

template <typename T>
Circ<T> make_circ(double alfa, const LinePlan& P)
{
  const double gam = circ_gamma(alfa);
  const double beta = -alfa / gam;
  Circ<T> c{};
  c.beta = (T)beta;
  double pw = 1.0;
  for (int k = 0; k <= 32; ++k) {
    c.pw[k] = (T)(k <= P.M ? pw : 0.0);
    pw *= beta;
  }
  double bM = std::pow(beta, P.M);
  for (int k = 0; k < 7; ++k) {
    c.pwM2[k] = (T)bM;
    bM = bM * bM;
  }
  // beta^(M c) multiplied out of the binary powers in T, lowest bit first
  // (the device's pow_Mc, so the table holds the same values)
  for (int i = 0; i < 64; ++i) {
    T r = T(1);
    for (int k = 0; k < 7; ++k)
      if ((i >> k) & 1) r = r * c.pwM2[k];
    c.pMc[i] = r;
  }
  c.pwMG = (T)std::pow(beta, (double)P.M * P.G);
  c.close = (T)(1.0 / (1.0 - std::pow(beta, (double)P.n)));
  return c;
}
template Circ<double> make_circ<double>(double, const LinePlan&);
template Circ<float> make_circ<float>(double, const LinePlan&);

Coef<double> stencil_coef(const Scheme& s, int kind)
{
  switch (kind) {
    case K_D: return {s.afi, s.bfi, 0.0, 0.0};
    case K_DVP: case K_DPV: return {s.aci, s.bci, 0.0, 0.0};
    case K_IVP: case K_IPV: return {s.aici, s.bici, s.cici, s.dici};
    case K_DD: return {s.asi, s.bsi, 0.0, 0.0};
    default: return {0.0, 0.0, 0.0, 0.0};
  }
}

// The y/z circulant kernels of 16-point chunks are held to 80 VGPRs (they
// need 78-80, no spills), so that three 512-thread work-groups share a
// compute unit: 2-4% faster at 512^3 than two.  BFP32 needs more (the
// exponent of each loaded and stored row): 128.
// S: the fields' storage (FP64, FP32 or BFP32, computed in FP64).
template <typename S, int KIND, int M, bool CIRC, bool LF>
__global__ __launch_bounds__(LF ? 1024 : 256) __attribute__((amdgpu_waves_per_eu(LF && CIRC && M == 16 ? (is_bfp<S> ? 4 : 6) : 1)))
void k_line(const S* in, S* out, OpDev<typename Store<S>::T> op, Coef<typename Store<S>::T> cf,
            Circ<typename Store<S>::T> cc, Geom geo, Tiling tl)
{
  using T = typename Store<S>::T;
  constexpr int HL = Halo<KIND>::L, HR = Halo<KIND>::R;
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<S>(geo, tl, M, tl.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x);
  T v[M];
  if constexpr (LF) {
    T win[M + HL + HR];
    load_window<S, M, HL, HR, LF>(in, t, geo.n, win);
    rhs_chunk<KIND, T, M, HL>(win, cf, v);
    if constexpr (CIRC) solve_circ<T, M, LF>(v, t, cc, lds);
    else solve_table<T, M, LF>(v, t, op, lds, geo.n, tl.fix != 0);
    store_chunk<S, M>(out, t, v);
  } else {
    // x-lines: the solvers use no LDS (a line is inside one wavefront), so
    // the LDS holds the wavefronts' staging tiles
    const XBlock<S, M> xb = x_block<S, M>(geo, t, lds);
    {
      T win[M + HL + HR];
      x_load_window<S, M, HL, HR>(in, xb, t, win);
      rhs_chunk<KIND, T, M, HL>(win, cf, v);
    }
    if constexpr (CIRC) solve_circ<T, M, LF>(v, t, cc, lds);
    else solve_table<T, M, LF>(v, t, op, lds, geo.n, tl.fix != 0);
    x_store<S, M>(out, xb, v);
  }
}

// Right-hand side of the non-periodic first derivative at point j of n from
// a window whose entry q is point j (der?_ij, src/x3d_derive.f90:172-210):
// the interior sixth-order stencil, and at j = 0, 1, n-2, n-1 the free-slip
// rows (npaire: even or odd extension) or the one-sided Dirichlet closures.
// Every boundary row reads points inside [j - 2, j + 2], so the window holds
// them whatever the chunk.
__device__ __forceinline__ double np_rhs(const double* w, int q, int j, int n, const NpRows& r)
{
  if (j >= 2 && j <= n - 3) return r.afi * (w[q + 1] - w[q - 1]) + r.bfi * (w[q + 2] - w[q - 2]);
  if (j == 0) {
    if (r.ncl1 == 1) return r.npaire == 1 ? 0.0 : r.afi * (w[q + 1] + w[q + 1]) + r.bfi * (w[q + 2] + w[q + 2]);
    return r.af1 * w[q] + r.bf1 * w[q + 1] + r.cf1 * w[q + 2];
  }
  if (j == 1) {
    if (r.ncl1 == 1)
      return r.npaire == 1 ? r.afi * (w[q + 1] - w[q - 1]) + r.bfi * (w[q + 2] - w[q])
                           : r.afi * (w[q + 1] - w[q - 1]) + r.bfi * (w[q + 2] + w[q]);
    return r.af2 * (w[q + 1] - w[q - 1]);
  }
  if (j == n - 2) {
    if (r.ncln == 1)
      return r.npaire == 1 ? r.afi * (w[q + 1] - w[q - 1]) + r.bfi * (w[q] - w[q - 2])
                           : r.afi * (w[q + 1] - w[q - 1]) + r.bfi * ((-w[q]) - w[q - 2]);
    return r.afm * (w[q + 1] - w[q - 1]);
  }
  if (r.ncln == 1) return r.npaire == 1 ? 0.0 : r.afi * ((-w[q - 1]) - w[q - 1]) + r.bfi * ((-w[q - 2]) - w[q - 2]);
  return -r.afn * w[q] - r.bfn * w[q - 1] - r.cfn * w[q - 2];
}

// One non-periodic first derivative along x, y or z (RHS), or the Thomas
// sweeps alone: the right-hand side, then the reference's sweeps with the
// instance's tables and no periodic correction (xthomas_12,
// src/thomas.f90:61-84).  Lines have any length n: they are cut into C
// chunks of M points with C M >= n, and the points past n are identity rows
// of the padded tables (f = s = 0, w = 1) with a zero right-hand side, which
// the sweeps carry through without touching the real points (f(n-1) = 0).
// Every load is clamped into the line (the clamped values feed no real row)
// and only real points are stored.  The Thomas form reads its own chunk
// alone, so it may run in place.
template <int M, bool LF, bool RHS>
__global__ __launch_bounds__(LF ? 1024 : 256) void k_der_np(const double* in, double* out, OpDev<double> op, NpRows r,
                                                            Geom geo, Tiling tl)
{
  constexpr int H = RHS ? 2 : 0;
  extern __shared__ unsigned char smem_raw[];
  double* lds = reinterpret_cast<double*>(smem_raw);
  const LineThread t = line_thread<double>(geo, tl, M, blockIdx.x);
  const int n = r.n;
  double v[M];
  {
    double win[M + 2 * H];
#pragma unroll
    for (int q = 0; q < M + 2 * H; ++q) {
      int j = t.j0 - H + q;
      j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
      win[q] = ld(in, t.at(j));
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int j = t.j0 + k;
      double x = 0.0;
      if constexpr (RHS) {
        if (j < n) x = np_rhs(win, k + H, j, n, r);
      } else {
        if (j < n) x = win[k];
      }
      v[k] = x;
    }
  }
  solve_table<double, M, LF>(v, t, op, lds, t.C * M, false);
  if (t.active) {
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (t.j0 + k < n) st(out, t.at(t.j0 + k), v[k]);
  }
}

// Points per chunk the kernels are instantiated for.
static const int kMs[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};

int line_plan(int n, bool line_fast, LinePlan* P)
{
  if (n < 8) return X3D_ERR_SIZE;
  int C = 0, M = 0;
  // y/z: at most 32 chunks per line first (at 512 points, 16-point chunks in
  // 512-thread work-groups ran dery/derz 30% faster than 8-point chunks in
  // 1024-thread ones), 64 where no shorter split exists; x: up to 64
  auto pick = [&](int cmax) {
    for (int c = cmax; c >= 1 && !C; c >>= 1) {
      if (n % c) continue;
      const int m = n / c;
      for (int mm : kMs)
        if (mm == m) { C = c; M = m; break; }
    }
  };
  pick(line_fast ? 32 : 64);
  if (!C && line_fast) pick(64);
  if (!C) return X3D_ERR_UNSUPPORTED;
  // x-lines take their halo (at most 4 points) from the neighbouring chunks
  // of the same wavefront, at most ceil(4 / M) of them, and wrap it once
  if (!line_fast && (4 + M - 1) / M > C) return X3D_ERR_UNSUPPORTED;
  LinePlan p;
  p.n = n; p.M = M; p.C = C; p.line_fast = line_fast;
  if (line_fast) {
    p.G = C < 4 ? C : 4;
    p.Wc = C / p.G;
    p.Wl = p.Wc >= 4 ? 1 : 4 / p.Wc;
  } else {
    p.G = C;
    p.Wc = 1;
    p.Wl = 4;
  }
  p.LW = 64 / p.G;
  *P = p;
  return X3D_OK;
}

// Fused passes hold several line solves per thread, so their y/z tiling caps a
// work-group at 8 wavefronts (up to 256 VGPRs each) and takes longer chunks:
// the largest C <= 32 chunks per line with 4 chunk groups x 16 lines per
// wavefront, else 64 chunks with 8 groups x 8 lines.
int line_plan_fused(int n, bool line_fast, LinePlan* P, int variant)
{
  if (!line_fast) return line_plan(n, false, P);
  if (n < 8) return X3D_ERR_SIZE;
  auto okM = [](int m) { for (int mm : kMs) if (mm == m) return true; return false; };
  LinePlan p;
  p.n = n; p.line_fast = true;
  // explicit tiling "M,G,Wl" (chunk length, chunk groups per wavefront, line
  // groups per work-group) from X3D_FUSED_YZ, used when it fits this line
  if (const char* e = std::getenv("X3D_FUSED_YZ")) {
    int m = 0, g = 0, wl = 0;
    if (std::sscanf(e, "%d,%d,%d", &m, &g, &wl) == 3 && m > 0 && okM(m) && n % m == 0 && g > 0 && g <= 64 &&
        (64 % g) == 0 && wl > 0) {
      const int c = n / m;
      if (c <= 64 && c % g == 0 && (c / g) * wl <= 8) {
        p.C = c; p.M = m; p.G = g; p.Wc = c / g; p.Wl = wl; p.LW = 64 / g;
        *P = p;
        return X3D_OK;
      }
    }
  }
  // variant 1: 64 short chunks per line, 8 lines x 8 chunk groups per wavefront
  if (variant == 1 && n % 64 == 0 && okM(n / 64)) {
    p.C = 64; p.M = n / 64; p.G = 8; p.Wc = 8; p.Wl = 1; p.LW = 8;
    *P = p;
    return X3D_OK;
  }
  for (int c = 32; c >= 1; c >>= 1) {
    if (n % c || !okM(n / c)) continue;
    p.C = c; p.M = n / c; p.G = c < 4 ? c : 4; p.Wc = c / p.G;
    p.Wl = p.Wc >= 4 ? 1 : 4 / p.Wc; p.LW = 64 / p.G;
    *P = p;
    return X3D_OK;
  }
  if (n % 64 == 0 && okM(n / 64)) {
    p.C = 64; p.M = n / 64; p.G = 8; p.Wc = 8; p.Wl = 1; p.LW = 8;
    *P = p;
    return X3D_OK;
  }
  return X3D_ERR_UNSUPPORTED;
}

// FP32 fused y/z passes: the same scheme with chunks up to 32 points, tiled
// for 128-byte rows where the line allows (32 lines x 2 chunk groups per
// wavefront), else 16 x 4 or 8 x 8; at most 8 wavefronts per work-group.
static const int kMs32[] = {1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 32};

int line_plan_fused32(int n, bool line_fast, LinePlan* P)
{
  if (!line_fast) return line_plan(n, false, P);
  if (n < 8) return X3D_ERR_SIZE;
  auto okM = [](int m) { for (int mm : kMs32) if (mm == m) return true; return false; };
  for (int g : {2, 4, 8}) {
    for (int c = 8 * g; c >= g; c >>= 1) {
      if (n % c || !okM(n / c)) continue;
      LinePlan p;
      p.n = n; p.line_fast = true; p.C = c; p.M = n / c; p.G = g; p.Wc = c / g; p.LW = 64 / g;
      p.Wl = p.Wc >= 4 ? 1 : 4 / p.Wc;
      *P = p;
      return X3D_OK;
    }
  }
  return X3D_ERR_UNSUPPORTED;
}

void launch_dims(const x3d_ctx* c, int dir, Geom* geo, Tiling* tl, dim3* grid, dim3* block, size_t* smem,
                 size_t tbytes, bool fused)
{
  launch_dims_plan(c, dir, fused ? c->fplan[dir] : c->plan[dir], geo, tl, grid, block, smem, tbytes);
}

void launch_dims_plan(const x3d_ctx* c, int dir, const LinePlan& P, Geom* geo, Tiling* tl, dim3* grid,
                      dim3* block, size_t* smem, size_t tbytes)
{
  geo->dir = dir; geo->nx = c->nx; geo->nxy = (long long)c->nx * c->ny; geo->n = P.n;
  geo->nlines = dir == 0 ? c->ny * c->nz : dir == 1 ? c->nx * c->nz : c->nx * c->ny;
  tl->G = P.G; tl->Wc = P.Wc; tl->Wl = P.Wl; tl->LW = P.LW; tl->line_fast = P.line_fast ? 1 : 0;
  tl->fix = 1;
  // measured at 512^3: 4-5% faster y passes; the z passes are faster with
  // their blocks spread over the XCDs, the x passes are unchanged
  tl->xcd = dir == 1 ? c->xcd_remap : 0;
  const int lpw = P.Wl * P.LW;
  *grid = dim3((geo->nlines + lpw - 1) / lpw);
  *block = dim3(64 * P.Wc * P.Wl);
  *smem = tbytes * (size_t)line_lds_words(P.Wc, lpw);
  // x-lines: one staging tile of 64 rows x (M or M+1, the odd one) per wavefront
  if (!P.line_fast) *smem = tbytes * (size_t)(P.Wc * P.Wl) * 64u * (size_t)((P.M & 1) ? P.M : P.M + 1);
}

template <typename S, int KIND, bool CIRC, typename T = typename Store<S>::T>
static hipError_t launch_kind(int M, dim3 grid, dim3 block, size_t smem, hipStream_t st, const S* in, S* out,
                              OpDev<T> op, Coef<T> cf, Circ<T> cc, Geom geo, Tiling tl)
{
#define X3D_CASE(MM)                                                                                   \
  case MM:                                                                                             \
    if (tl.line_fast)                                                                                  \
      hipLaunchKernelGGL((k_line<S, KIND, MM, CIRC, true>), grid, block, smem, st, in, out, op, cf, cc, geo, tl); \
    else                                                                                               \
      hipLaunchKernelGGL((k_line<S, KIND, MM, CIRC, false>), grid, block, smem, st, in, out, op, cf, cc, geo, tl); \
    break;
  // FP32 y/z lines also take the FP32 plans' 24- and 32-point chunks
#define X3D_CASE_F32_YZ(MM)                                                                            \
  case MM:                                                                                             \
    if constexpr (std::is_same<S, float>::value) {                                                     \
      if (!tl.line_fast) return hipErrorInvalidValue;                                                  \
      hipLaunchKernelGGL((k_line<S, KIND, MM, CIRC, true>), grid, block, smem, st, in, out, op, cf, cc, geo, tl); \
      break;                                                                                           \
    } else {                                                                                           \
      return hipErrorInvalidValue;                                                                     \
    }
  switch (M) {
    X3D_CASE(1) X3D_CASE(2) X3D_CASE(3) X3D_CASE(4) X3D_CASE(5) X3D_CASE(6)
    X3D_CASE(8) X3D_CASE(12) X3D_CASE(16)
    X3D_CASE_F32_YZ(24) X3D_CASE_F32_YZ(32)
    default: return hipErrorInvalidValue;
  }
#undef X3D_CASE_F32_YZ
#undef X3D_CASE
  return hipGetLastError();
}

template <typename S, bool CIRC, typename T = typename Store<S>::T>
static hipError_t launch_any(int kind, int M, dim3 grid, dim3 block, size_t smem, hipStream_t st, const S* in,
                             S* out, OpDev<T> od, Coef<T> cf, Circ<T> cc, Geom geo, Tiling tl)
{
  switch (kind) {
    case K_D: return launch_kind<S, K_D, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    case K_DVP: return launch_kind<S, K_DVP, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    case K_DPV: return launch_kind<S, K_DPV, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    case K_IVP: return launch_kind<S, K_IVP, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    case K_IPV: return launch_kind<S, K_IPV, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    case K_DD: return launch_kind<S, K_DD, CIRC>(M, grid, block, smem, st, in, out, od, cf, cc, geo, tl);
    default: return hipErrorInvalidValue;
  }
}

static int line_raw(x3d_ctx* c, int dir, int kind, int op, const double* in, double* out, bool periodic_fix)
{
  if (dir < 0 || dir > 2 || op < 0 || op >= X3D_NOPS_ALL || !in || !out) return X3D_ERR_ARG;
  if (c->op[op].bc_unsupported) return X3D_ERR_UNSUPPORTED;
  if (c->op[op].np) {
    // non-periodic first derivatives: der?_ij; Thomas alone: the _12 sweeps
    if (kind == K_D) return launch_line_np(c, dir, op, in, out, true);
    if (kind == K_NONE) return launch_line_np(c, dir, op, in, out, false);
    return X3D_ERR_UNSUPPORTED;
  }
  if (!c->f64_lines) return X3D_ERR_UNSUPPORTED;
  const LinePlan& P = c->plan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (c->op[op].n != n || P.n != n) return X3D_ERR_SIZE;
  if (P.M == 0) return X3D_ERR_UNSUPPORTED;

  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  launch_dims(c, dir, &geo, &tl, &grid, &block, &smem, sizeof(double));
  tl.fix = periodic_fix ? 1 : 0;

  const double* tb = c->tab64 + c->off[op];
  OpDev<double> od{tb, tb + n, tb + 2 * n, tb + 3 * n, c->op[op].alfa};
  Coef<double> cf = stencil_coef(c->sc[dir], kind);
  Circ<double> cc{};
  hipError_t e;
  // the circulant solver applies when the operator carries the library's own
  // periodic tables and the full periodic solve is asked for
  const bool circ = kind != K_NONE && periodic_fix && c->op[op].circulant;
  if (circ) {
    cc = make_circ<double>(c->op[op].alfa, P);
    const double ig = 1.0 / circ_gamma(c->op[op].alfa);
    cf = {cf.a * ig, cf.b * ig, cf.c * ig, cf.d * ig};
    e = launch_any<double, true>(kind, P.M, grid, block, smem, c->stream, in, out, od, cf, cc, geo, tl);
  } else if (kind == K_NONE) {
    e = launch_kind<double, K_NONE, false>(P.M, grid, block, smem, c->stream, in, out, od, cf, cc, geo, tl);
  } else {
    e = launch_any<double, false>(kind, P.M, grid, block, smem, c->stream, in, out, od, cf, cc, geo, tl);
  }
  return e == hipSuccess ? X3D_OK : hip_status(e);
}

int launch_thomas_tables(x3d_ctx* c, int dir, const double* tab, double alfa, double* tt, bool periodic_fix)
{
  if (dir < 0 || dir > 2 || !tab || !tt) return X3D_ERR_ARG;
  if (!c->f64_lines) return X3D_ERR_UNSUPPORTED;
  const LinePlan& P = c->plan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (P.n != n) return X3D_ERR_SIZE;
  if (P.M == 0) return X3D_ERR_UNSUPPORTED;
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  launch_dims(c, dir, &geo, &tl, &grid, &block, &smem, sizeof(double));
  tl.fix = periodic_fix ? 1 : 0;
  const OpDev<double> od{tab, tab + n, tab + 2 * n, tab + 3 * n, alfa};
  const Coef<double> cf{};
  const Circ<double> cc{};
  const hipError_t e = launch_kind<double, K_NONE, false>(P.M, grid, block, smem, c->stream, tt, tt, od, cf, cc, geo, tl);
  return e == hipSuccess ? X3D_OK : hip_status(e);
}

// FP32 precision form of one operator: the circulant solver in FP32 (the
// library's periodic tables only; no Thomas-table path).
static int line_raw_f32(x3d_ctx* c, int dir, int kind, int op, const float* in, float* out)
{
  if (dir < 0 || dir > 2 || op < 0 || op >= X3D_NOPS_ALL || !in || !out || kind == K_NONE) return X3D_ERR_ARG;
  // y/z: the FP32 plans' tiling (chunks up to 32 points, 128-byte rows of 32
  // lines where the line allows) when there is one, as in the FP32 step
  const LinePlan& P = dir > 0 && c->fplan32[dir].M > 0 ? c->fplan32[dir] : c->plan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (c->op[op].n != n || P.n != n) return X3D_ERR_SIZE;
  if (!c->op[op].circulant) return X3D_ERR_UNSUPPORTED;
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  launch_dims_plan(c, dir, P, &geo, &tl, &grid, &block, &smem, sizeof(float));
  const Coef<double> k = stencil_coef(c->sc[dir], kind);
  const double ig = 1.0 / circ_gamma(c->op[op].alfa);
  const Coef<float> cf{(float)(k.a * ig), (float)(k.b * ig), (float)(k.c * ig), (float)(k.d * ig)};
  const Circ<float> cc = make_circ<float>(c->op[op].alfa, P);
  const OpDev<float> od{nullptr, nullptr, nullptr, nullptr, (float)c->op[op].alfa};
  const hipError_t e = launch_any<float, true>(kind, P.M, grid, block, smem, c->stream, in, out, od, cf, cc, geo, tl);
  return e == hipSuccess ? X3D_OK : hip_status(e);
}

NpRows np_rows(const x3d_ctx* c, int dir, int op)
{
  // first_deriv_exp_, ifirstder = 4 (src/x3d_operator_1d.f90:350-391)
  const double h = dir == 0 ? c->dx : dir == 1 ? c->dy : c->dz;
  NpRows r;
  r.n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  r.ncl1 = c->ncl1[dir]; r.ncln = c->ncln[dir]; r.npaire = c->op[op].npaire;
  r.afi = c->sc[dir].afi; r.bfi = c->sc[dir].bfi;
  r.af1 = -(5.0 / 2.0) / h; r.bf1 = 2.0 / h; r.cf1 = 0.5 / h;
  r.af2 = (3.0 / 4.0) / h;
  r.afn = -(5.0 / 2.0) / h; r.bfn = 2.0 / h; r.cfn = 0.5 / h;
  r.afm = (3.0 / 4.0) / h;
  return r;
}

// Chunking of a non-periodic line of any length n >= 8: C chunks (a power of
// two up to 64) of M <= 16 points with C M >= n and the least padding; y/z
// take up to 4 chunk groups per wavefront and C / 4 wavefronts per line set.
int np_plan(int n, bool line_fast, LinePlan* P)
{
  if (n < 8) return X3D_ERR_SIZE;
  int C = 0, M = 0, pad = 1 << 30;
  for (int c = 1; c <= 64; c <<= 1)
    for (int m = 1; m <= 16; ++m)
      if (c * m >= n && c * m - n < pad) { C = c; M = m; pad = c * m - n; }
  if (!C) return X3D_ERR_UNSUPPORTED;
  LinePlan p;
  p.n = n; p.M = M; p.C = C; p.line_fast = line_fast;
  if (line_fast) {
    p.G = C < 4 ? C : 4;
    p.Wc = C / p.G;
    p.Wl = p.Wc >= 4 ? 1 : 4 / p.Wc;
  } else {
    p.G = C;
    p.Wc = 1;
    p.Wl = 4;
  }
  p.LW = 64 / p.G;
  *P = p;
  return X3D_OK;
}

int launch_line_np(x3d_ctx* c, int dir, int op, const double* in, double* out, bool rhs)
{
  if (dir < 0 || dir > 2 || op < 0 || op >= X3D_NOPS_ALL || !in || !out) return X3D_ERR_ARG;
  if (!c->f64_lines) return X3D_ERR_UNSUPPORTED;
  const LinePlan& P = c->npplan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (c->op[op].n != n || P.n != n) return X3D_ERR_SIZE;
  if (!c->op[op].np || P.M == 0 || !c->nptab) return X3D_ERR_UNSUPPORTED;
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  launch_dims_plan(c, dir, P, &geo, &tl, &grid, &block, &smem, sizeof(double));
  tl.xcd = 0;
  if (!P.line_fast) smem = 0;  // x-lines: no staging tile, and the sweeps stay in the wavefront
  const size_t np = (size_t)P.C * P.M;
  const double* tb = c->nptab + c->op[op].npoff;
  const OpDev<double> od{tb, tb + np, tb + 2 * np, nullptr, c->op[op].alfa};
  NpRows r = np_rows(c, dir, op);
#define X3D_NP(MM)                                                                                            \
  case MM:                                                                                                    \
    if (P.line_fast) {                                                                                        \
      if (rhs) hipLaunchKernelGGL((k_der_np<MM, true, true>), grid, block, smem, c->stream, in, out, od, r, geo, tl); \
      else hipLaunchKernelGGL((k_der_np<MM, true, false>), grid, block, smem, c->stream, in, out, od, r, geo, tl); \
    } else {                                                                                                  \
      if (rhs) hipLaunchKernelGGL((k_der_np<MM, false, true>), grid, block, smem, c->stream, in, out, od, r, geo, tl); \
      else hipLaunchKernelGGL((k_der_np<MM, false, false>), grid, block, smem, c->stream, in, out, od, r, geo, tl); \
    }                                                                                                         \
    break;
  switch (P.M) {
    X3D_NP(1) X3D_NP(2) X3D_NP(3) X3D_NP(4) X3D_NP(5) X3D_NP(6) X3D_NP(7) X3D_NP(8)
    X3D_NP(9) X3D_NP(10) X3D_NP(11) X3D_NP(12) X3D_NP(13) X3D_NP(14) X3D_NP(15) X3D_NP(16)
    default: return X3D_ERR_UNSUPPORTED;
  }
#undef X3D_NP
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

// BFP32 form of one operator: BFP32 fields, the circulant solver in FP64 on
// the FP64 plans (y/z-line lane groups of 16 lines: one segment per row).
static int line_raw_bfp(x3d_ctx* c, int dir, int kind, int op, const bfp32* in, bfp32* out)
{
  if (dir < 0 || dir > 2 || op < 0 || op >= X3D_NOPS_ALL || !in || !out || kind == K_NONE) return X3D_ERR_ARG;
  if (c->nx % 16) return X3D_ERR_UNSUPPORTED;
  const LinePlan& P = c->plan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (c->op[op].n != n || P.n != n) return X3D_ERR_SIZE;
  if (!c->op[op].circulant || P.M == 0) return X3D_ERR_UNSUPPORTED;
  if (P.line_fast && P.LW % 16) return X3D_ERR_UNSUPPORTED;
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  launch_dims_plan(c, dir, P, &geo, &tl, &grid, &block, &smem, sizeof(double));
  if (dir == 2) tl.xcd = c->xcd_remap;  // 64 B rows: both halves of a 128 B line on one XCD
  const Coef<double> k = stencil_coef(c->sc[dir], kind);
  const double ig = 1.0 / circ_gamma(c->op[op].alfa);
  const Coef<double> cf{k.a * ig, k.b * ig, k.c * ig, k.d * ig};
  const Circ<double> cc = make_circ<double>(c->op[op].alfa, P);
  const OpDev<double> od{nullptr, nullptr, nullptr, nullptr, c->op[op].alfa};
  const hipError_t e = launch_any<bfp32, true>(kind, P.M, grid, block, smem, c->stream, in, out, od, cf, cc, geo, tl);
  return e == hipSuccess ? X3D_OK : hip_status(e);
}

// ---- y-stretching: the reference's separate loop after the line solve --------
// out(i, j, k) *= s(j) over the whole field; 256-thread blocks, a grid-stride
// loop.  BFP32: each 16-point x-segment lies in one y-row, so it is decoded,
// scaled by its row's factor and encoded again with a new exponent.
template <typename T>
__global__ __launch_bounds__(256) void k_scale_y(T* __restrict__ out, const T* __restrict__ s, int nx, int ny,
                                                 long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = out[i] * s[(i / nx) % ny];
}

__global__ __launch_bounds__(256) void k_scale_y_bfp(int* __restrict__ out, const double* __restrict__ s, int nx,
                                                     int ny, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = bfp_word(bfp_value(out[i]) * s[(i / nx) % ny]);
}

static dim3 scale_blocks(long long n)
{
  const long long b = (n + 255) / 256;
  return dim3((unsigned)(b < 65536 ? b : 65536));
}

int stretch_y(x3d_ctx* c, int which, double* out)
{
  const long long n = (long long)c->nx * c->ny * c->nz;
  hipLaunchKernelGGL(k_scale_y<double>, scale_blocks(n), dim3(256), 0, c->stream, out, stretch_vec<double>(c, which),
                     c->nx, c->ny, n);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

int stretch_y_f32(x3d_ctx* c, int which, float* out)
{
  const long long n = (long long)c->nx * c->ny * c->nz;
  hipLaunchKernelGGL(k_scale_y<float>, scale_blocks(n), dim3(256), 0, c->stream, out, stretch_vec<float>(c, which),
                     c->nx, c->ny, n);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

int stretch_y_bfp(x3d_ctx* c, int which, bfp32* out)
{
  // a 16-lane row must hold one whole segment: nx % 16 == 0 (checked by the
  // BFP32 entry points) and 256-thread blocks start on segment boundaries
  const long long n = (long long)c->nx * c->ny * c->nz;
  hipLaunchKernelGGL(k_scale_y_bfp, scale_blocks(n), dim3(256), 0, c->stream, reinterpret_cast<int*>(out),
                     stretch_vec<double>(c, which), c->nx, c->ny, n);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

#define X3D_TRY_(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

int launch_line(x3d_ctx* c, int dir, int kind, int op, const double* in, double* out, bool periodic_fix)
{
  const int w = stretch_which(c, dir, kind);
  if (w == -2) return X3D_ERR_UNSUPPORTED;
  X3D_TRY_(line_raw(c, dir, kind, op, in, out, periodic_fix));
  return w >= 0 ? stretch_y(c, w, out) : X3D_OK;
}

int launch_line_f32(x3d_ctx* c, int dir, int kind, int op, const float* in, float* out)
{
  const int w = stretch_which(c, dir, kind);
  if (w == -2) return X3D_ERR_UNSUPPORTED;
  X3D_TRY_(line_raw_f32(c, dir, kind, op, in, out));
  return w >= 0 ? stretch_y_f32(c, w, out) : X3D_OK;
}

int launch_line_bfp(x3d_ctx* c, int dir, int kind, int op, const bfp32* in, bfp32* out)
{
  const int w = stretch_which(c, dir, kind);
  if (w == -2) return X3D_ERR_UNSUPPORTED;
  X3D_TRY_(line_raw_bfp(c, dir, kind, op, in, out));
  return w >= 0 ? stretch_y_bfp(c, w, out) : X3D_OK;
}
#undef X3D_TRY_

}  // namespace x3d
