// Periodic compact operators on gfx950: stencil right-hand side, Thomas
// forward/backward sweeps and the Sherman-Morrison periodic correction in one
// pass over each line (one global read of the input window, one write).
//
// The reference runs each line's sweeps sequentially (src/thomas.f90:61-84
// and the y/z twins), so one line per thread needs the whole line's
// intermediate on chip.  Here a line of n points is cut into C chunks of M
// consecutive points, one chunk per thread, and both sweeps run as chunked
// first-order linear recurrences:
//   forward  t_j = r_j - s_j t_{j-1}        (carry-less local pass, then
//   backward x_j = (t_j - f_j x_{j+1}) w_j   a scan of the chunk maps
//                                           x -> A x + B and a fix-up).
// The scan composes the chunks of a line inside a wavefront with cross-lane
// shuffles and across wavefronts through LDS.  The coefficients are the
// operator's own f, s, w and periodic vectors (x3doperator1d), so the
// arithmetic is the reference's Thomas algorithm, reassociated.
//
// Lane layouts: x-lines are contiguous, so consecutive lanes take
// consecutive chunks of one line; y/z-lines are strided, so consecutive
// lanes take 16 consecutive lines (128 B rows) and the chunks of a line are
// spread over lane groups and wavefronts.
#include "x3d_internal.h"

namespace x3d {

template <int KIND>
struct Halo { static constexpr int L = 0, R = 0; };
template <> struct Halo<K_D> { static constexpr int L = 2, R = 2; };
template <> struct Halo<K_DVP> { static constexpr int L = 1, R = 2; };
template <> struct Halo<K_DPV> { static constexpr int L = 2, R = 1; };
template <> struct Halo<K_IVP> { static constexpr int L = 3, R = 4; };
template <> struct Halo<K_IPV> { static constexpr int L = 4, R = 3; };

template <typename T>
struct Coef { T a, b, c, d; };

// Right-hand side at window index q (window holds u_{j0-HL} ...).
// D    derx_00      src/x3d_derive.f90:134-137
// DVP  derxvp       src/x3d_staggered.f90:48-51
// DPV  derxpv       src/x3d_staggered.f90:246-249
// IVP  interxvp     src/x3d_staggered.f90:143-148
// IPV  interxpv     src/x3d_staggered.f90:326-331
// (the wrapped boundary rows are the same formulas with indices mod n)
template <int KIND, typename T>
__device__ __forceinline__ T rhs_at(const T* win, int q, const Coef<T>& k)
{
  if constexpr (KIND == K_D)
    return k.a * (win[q + 1] - win[q - 1]) + k.b * (win[q + 2] - win[q - 2]);
  else if constexpr (KIND == K_DVP)
    return k.a * (win[q + 1] - win[q]) + k.b * (win[q + 2] - win[q - 1]);
  else if constexpr (KIND == K_DPV)
    return k.a * (win[q] - win[q - 1]) + k.b * (win[q + 1] - win[q - 2]);
  else if constexpr (KIND == K_IVP)
    return k.a * (win[q + 1] + win[q]) + k.b * (win[q + 2] + win[q - 1]) +
           k.c * (win[q + 3] + win[q - 2]) + k.d * (win[q + 4] + win[q - 3]);
  else if constexpr (KIND == K_IPV)
    return k.a * (win[q] + win[q - 1]) + k.b * (win[q + 1] + win[q - 2]) +
           k.c * (win[q + 2] + win[q - 3]) + k.d * (win[q + 3] + win[q - 4]);
  else
    return win[q];
}

struct Geom {
  int dir;        // 0 x, 1 y, 2 z
  int nx;
  long long nxy;  // nx * ny
  int n;          // line length
  int nlines;
};

__device__ __forceinline__ long long line_base(const Geom& g, int line)
{
  if (g.dir == 0) return (long long)line * g.nx;
  if (g.dir == 1) {
    const int k = line / g.nx;
    return (long long)(line - k * g.nx) + (long long)k * g.nxy;
  }
  return line;
}

__device__ __forceinline__ long long point_stride(const Geom& g)
{
  return g.dir == 0 ? 1 : (g.dir == 1 ? (long long)g.nx : g.nxy);
}

struct Tiling {
  int G, Wc, Wl, LW, line_fast, fix;
};

template <typename T, int KIND, int M>
__global__ __launch_bounds__(1024) void k_line(const T* in, T* out, OpDev<T> op, Coef<T> cf,
                                               Geom geo, Tiling tl)
{
  constexpr int HL = Halo<KIND>::L, HR = Halo<KIND>::R, NW = M + HL + HR;
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);

  const int G = tl.G, Wc = tl.Wc, LW = tl.LW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wl = wave / Wc, wc = wave - wl * Wc;
  int lineLocal, g;
  if (tl.line_fast) { lineLocal = lane % LW; g = lane / LW; }
  else { lineLocal = lane / G; g = lane - lineLocal * G; }
  const int LPW = tl.Wl * LW;                 // lines per work-group
  const int ls = wl * LW + lineLocal;         // line slot in the work-group
  const int line = blockIdx.x * LPW + ls;
  const bool active = line < geo.nlines;
  const int n = geo.n;
  const int C = G * Wc;
  const int c = wc * G + g;                   // chunk index along the line
  const int j0 = c * M;
  // lane holding chunk gg of this thread's line inside the wavefront
  auto lane_of = [&](int gg) { return tl.line_fast ? lineLocal + gg * LW : lineLocal * G + gg; };

  T* sA = sm;                    // forward aggregates  [Wc][LPW]
  T* sB = sA + Wc * LPW;
  T* rA = sB + Wc * LPW;         // backward aggregates [Wc][LPW]
  T* rB = rA + Wc * LPW;
  T* s0 = rB + Wc * LPW;         // x_1 and x_n of each line [LPW]
  T* sN = s0 + LPW;

  long long base = 0, ps = 1;
  if (active) { base = line_base(geo, line); ps = point_stride(geo); }

  T win[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    int j = j0 - HL + q;
    j = j < 0 ? j + n : (j >= n ? j - n : j);
    win[q] = active ? in[base + (long long)j * ps] : T(0);
  }
  T v[M];
#pragma unroll
  for (int jj = 0; jj < M; ++jj) v[jj] = rhs_at<KIND>(win, jj + HL, cf);

  // ---- forward sweep t_j = r_j - s_j t_{j-1} (src/thomas.f90:73-75) ----
  T A = T(1), B = T(0);
  {
    T prev = T(0);
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      const T s = op.s[j0 + jj];
      v[jj] = v[jj] - prev * s;
      prev = v[jj];
      A = A * (-s);
    }
    B = prev;
  }
  for (int d = 1; d < G; d <<= 1) {
    const int src = lane_of(g >= d ? g - d : g);
    const T Ao = __shfl(A, src), Bo = __shfl(B, src);
    if (g >= d) { B = A * Bo + B; A = A * Ao; }
  }
  T carry;
  {
    const int src = lane_of(g > 0 ? g - 1 : 0);
    const T Ap = __shfl(A, src), Bp = __shfl(B, src);
    const T Ae = g > 0 ? Ap : T(1), Be = g > 0 ? Bp : T(0);
    if (Wc > 1) {
      if (g == G - 1) { sA[wc * LPW + ls] = A; sB[wc * LPW + ls] = B; }
      __syncthreads();
      T cw = T(0);
      for (int w2 = 0; w2 < wc; ++w2) cw = sA[w2 * LPW + ls] * cw + sB[w2 * LPW + ls];
      carry = Ae * cw + Be;
    } else {
      carry = Be;
    }
  }
  {
    T P = T(1);
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      P = P * (-op.s[j0 + jj]);
      v[jj] = v[jj] + P * carry;
    }
  }

  // ---- backward sweep x_j = (t_j - f_j x_{j+1}) w_j (src/thomas.f90:76-79) ----
  A = T(1);
  {
    T nxt = T(0);
#pragma unroll
    for (int jj = M - 1; jj >= 0; --jj) {
      const T f = op.f[j0 + jj], w = op.w[j0 + jj];
      v[jj] = (v[jj] - f * nxt) * w;
      nxt = v[jj];
      A = A * (-(f * w));
    }
    B = nxt;
  }
  for (int d = 1; d < G; d <<= 1) {
    const int src = lane_of(g + d < G ? g + d : g);
    const T Ao = __shfl(A, src), Bo = __shfl(B, src);
    if (g + d < G) { B = A * Bo + B; A = A * Ao; }
  }
  {
    const int src = lane_of(g < G - 1 ? g + 1 : g);
    const T Ap = __shfl(A, src), Bp = __shfl(B, src);
    const T Ae = g < G - 1 ? Ap : T(1), Be = g < G - 1 ? Bp : T(0);
    if (Wc > 1) {
      if (g == 0) { rA[wc * LPW + ls] = A; rB[wc * LPW + ls] = B; }
      __syncthreads();
      T cw = T(0);
      for (int w2 = Wc - 1; w2 > wc; --w2) cw = rA[w2 * LPW + ls] * cw + rB[w2 * LPW + ls];
      carry = Ae * cw + Be;
    } else {
      carry = Be;
    }
  }
  {
    T Q = T(1);
#pragma unroll
    for (int jj = M - 1; jj >= 0; --jj) {
      const T f = op.f[j0 + jj], w = op.w[j0 + jj];
      Q = Q * (-(f * w));
      v[jj] = v[jj] + Q * carry;
    }
  }

  // ---- periodic correction (src/thomas.f90:49-55) ----
  if (tl.fix) {
    T x0, xn;
    if (Wc > 1) {
      if (c == 0) s0[ls] = v[0];
      if (c == C - 1) sN[ls] = v[M - 1];
      __syncthreads();
      x0 = s0[ls];
      xn = sN[ls];
    } else {
      x0 = __shfl(v[0], lane_of(0));
      xn = __shfl(v[M - 1], lane_of(G - 1));
    }
    const T ss = (x0 - op.alfa * xn) / (T(1) + op.p[0] - op.alfa * op.p[n - 1]);
#pragma unroll
    for (int jj = 0; jj < M; ++jj) v[jj] = v[jj] - ss * op.p[j0 + jj];
  }

  if (active) {
#pragma unroll
    for (int jj = 0; jj < M; ++jj) out[base + (long long)(j0 + jj) * ps] = v[jj];
  }
}

// Points per chunk the kernels are instantiated for.
static const int kMs[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};

int line_plan(int n, bool line_fast, LinePlan* P)
{
  if (n < 8) return X3D_ERR_SIZE;
  int C = 0, M = 0;
  for (int c = 64; c >= 1 && !C; c >>= 1) {
    if (n % c) continue;
    const int m = n / c;
    for (int mm : kMs)
      if (mm == m) { C = c; M = m; break; }
  }
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

template <typename T, int KIND>
static hipError_t launch_kind(int M, dim3 grid, dim3 block, size_t smem, hipStream_t st,
                              const T* in, T* out, OpDev<T> op, Coef<T> cf, Geom geo, Tiling tl)
{
#define X3D_CASE(MM) \
  case MM: hipLaunchKernelGGL((k_line<T, KIND, MM>), grid, block, smem, st, in, out, op, cf, geo, tl); break;
  switch (M) {
    X3D_CASE(1) X3D_CASE(2) X3D_CASE(3) X3D_CASE(4) X3D_CASE(5) X3D_CASE(6)
    X3D_CASE(8) X3D_CASE(12) X3D_CASE(16)
    default: return hipErrorInvalidValue;
  }
#undef X3D_CASE
  return hipGetLastError();
}

int launch_line(x3d_ctx* c, int dir, int kind, int op, const double* in, double* out,
                bool periodic_fix)
{
  if (dir < 0 || dir > 2 || op < 0 || op >= X3D_NOPS || !in || !out) return X3D_ERR_ARG;
  const LinePlan& P = c->plan[dir];
  const int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  if (c->op[op].n != n || P.n != n) return X3D_ERR_SIZE;
  if (P.M == 0) return X3D_ERR_UNSUPPORTED;

  Geom geo;
  geo.dir = dir; geo.nx = c->nx; geo.nxy = (long long)c->nx * c->ny; geo.n = n;
  geo.nlines = dir == 0 ? c->ny * c->nz : dir == 1 ? c->nx * c->nz : c->nx * c->ny;
  Tiling tl;
  tl.G = P.G; tl.Wc = P.Wc; tl.Wl = P.Wl; tl.LW = P.LW; tl.line_fast = P.line_fast ? 1 : 0;
  tl.fix = periodic_fix ? 1 : 0;
  const int lpw = P.Wl * P.LW;
  dim3 grid((geo.nlines + lpw - 1) / lpw), block(64 * P.Wc * P.Wl);
  const size_t smem = P.Wc > 1 ? sizeof(double) * (size_t)(4 * P.Wc * lpw + 2 * lpw) : 0;

  const double* t = c->tab64 + c->off[op];
  OpDev<double> od{t, t + n, t + 2 * n, t + 3 * n, c->op[op].alfa};
  const Scheme& s = c->sc[dir];
  Coef<double> cf{0, 0, 0, 0};
  switch (kind) {
    case K_D: cf = {s.afi, s.bfi, 0, 0}; break;
    case K_DVP: case K_DPV: cf = {s.aci, s.bci, 0, 0}; break;
    case K_IVP: case K_IPV: cf = {s.aici, s.bici, s.cici, s.dici}; break;
    default: break;
  }
  hipError_t e;
  switch (kind) {
    case K_D: e = launch_kind<double, K_D>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    case K_DVP: e = launch_kind<double, K_DVP>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    case K_DPV: e = launch_kind<double, K_DPV>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    case K_IVP: e = launch_kind<double, K_IVP>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    case K_IPV: e = launch_kind<double, K_IPV>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    case K_NONE: e = launch_kind<double, K_NONE>(P.M, grid, block, smem, c->stream, in, out, od, cf, geo, tl); break;
    default: return X3D_ERR_ARG;
  }
  return e == hipSuccess ? X3D_OK : hip_status(e);
}

}  // namespace x3d
