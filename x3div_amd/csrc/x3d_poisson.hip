// Periodic spectral Poisson solve, poisson_000 (src/poisson.f90:276-419).
//
// The reference takes 2decomp's r2c with z halved, applies per mode the
// forward phase shifts e^{-i theta}, the sign flips for the upper halves of
// x and y, the division by -kxyz (zeroed when Re or Im of kxyz < 1e-16), the
// inverse shifts with conjugations, then c2r.  The shifts, flips and
// conjugations cancel exactly, leaving  c -> -c / (K ntot)  (0 where
// K < 1e-16), which is what the spectral kernel applies; a parity test
// against the statement-by-statement host restatement confirms it.  The
// multiplier depends on |wavenumber| per axis only, so the halved axis is
// free: here x (the contiguous axis) is halved.
//
// K = xk2 (Ty Tz)^2 + yk2 (Tx Tz)^2 + zk2 (Tx Ty)^2 is evaluated per mode
// from 1-D tables in the reference's operation order with contraction off,
// so it is bit-identical to kxyz (src/poisson.f90:673-695) and the null-mode
// test makes the same decisions.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "x3d_internal.h"

namespace x3d {

static const double PI = 3.14159265358979323846264338327950288;

// One grid of a grouped launch (the batching form: one launch per Poisson
// stage and line length over many grids): its half spectrum, its real field,
// its 1-D tables (xk2, tx, yk2, ty, zk2, tz), sizes and 1 / (nx ny nz), and
// t0, its first tile in the group's concatenated tile sequence.  A group's
// records end with one holding only t0 = the group's tile count.
struct PoisGrid {
  void* spec;
  void* rhs;
  const double* ktab;
  long long t0;
  int nx, ny, nz, pad;
  double inv_ntot;
};

// ktab layout: [xk2 nx][Tx nx][yk2 ny][Ty ny][zk2 nz][Tz nz]
__global__ void k_spectral(double2* __restrict__ spec, const double* __restrict__ ktab, int nx,
                           int ny, int nz, double inv_ntot)
{
#pragma clang fp contract(off)
  const int nxh = nx / 2 + 1;
  const long long N = (long long)nxh * ny * nz;
  const double* xk2 = ktab;
  const double* tx = xk2 + nx;
  const double* yk2 = tx + nx;
  const double* ty = yk2 + ny;
  const double* zk2 = ty + ny;
  const double* tz = zk2 + nz;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % nxh);
    const long long r = q / nxh;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const double tyz = ty[j] * tz[k], txz = tx[i] * tz[k], txy = tx[i] * ty[j];
    const double K = (xk2[i] * (tyz * tyz) + yk2[j] * (txz * txz)) + zk2[k] * (txy * txy);
    double2 v = spec[q];
    if (K < 1.0e-16) {
      v.x = 0.0;
      v.y = 0.0;
    } else {
      v.x = (v.x * inv_ntot) / (-K);
      v.y = (v.y * inv_ntot) / (-K);
    }
    spec[q] = v;
  }
}

// ---- line stages -------------------------------------------------------------
// For line lengths 2^a 3^b 5^c the 3-D transform is split into line stages.
// Forward: the x r2c (k_rfft_x) and the y DFT (k_spec_lines mode 1), or
// hipFFT's 2-D (x, y) plan where those lengths are not built.  Then one
// kernel does, per z-line of the half spectrum, the forward length-nz DFT,
// the division above and the inverse DFT, so the spectrum crosses HBM once
// for the three steps.  Inverse: the y DFT (mode 2), then the x c2r
// (k_irfft_x in FP32, hipFFT's batched 1-D c2r in FP64), or hipFFT's 2-D
// plan.  A work-group takes TI neighbouring x-columns (16 where they fit in
// LDS: 256-byte rows for FP64), holds the TI lines in LDS and runs Stockham
// radix-8 passes (then a radix-4 or radix-2 pass, then radix-3 and radix-5
// passes) with twiddles from a table of e^{-2 pi i t / n}.

template <typename C>
struct CplxT;
template <>
struct CplxT<double2> { using R = double; };
template <>
struct CplxT<float2> { using R = float; };

template <typename C>
__device__ __forceinline__ C cmul(C a, double2 w, bool conj)
{
  using R = typename CplxT<C>::R;
  const R wx = (R)w.x, wy = conj ? (R)-w.y : (R)w.y;
  C r;
  r.x = a.x * wx - a.y * wy;
  r.y = a.x * wy + a.y * wx;
  return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by -i (forward) or +i (inverse)
template <typename C>
__device__ __forceinline__ C cmi(C a, bool inv) { C r; if (inv) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; } return r; }

// In-place DFT of length R (2, 3, 4, 5, 8) with sign -1 (forward) or +1
// (inverse).
template <typename C, int R>
__device__ __forceinline__ void small_dft(C (&v)[R], bool inv)
{
  using T = typename CplxT<C>::R;
  if constexpr (R == 3) {
    // X1,2 = a - (b + c)/2 -/+ i (sqrt 3 / 2)(b - c) (forward)
    const T s3 = (T)0.86602540378443864676;
    const C a = v[0], t1 = cadd(v[1], v[2]);
    C m;
    m.x = a.x - (T)0.5 * t1.x;
    m.y = a.y - (T)0.5 * t1.y;
    C n = cmi(csub(v[1], v[2]), inv);
    n.x = s3 * n.x;
    n.y = s3 * n.y;
    v[0] = cadd(a, t1);
    v[1] = cadd(m, n);
    v[2] = csub(m, n);
  } else if constexpr (R == 5) {
    // t1 = b + e, t2 = c + d, t3 = b - e, t4 = c - d;
    // X1,4 = a + c1 t1 + c2 t2 -/+ i (s1 t3 + s2 t4),
    // X2,3 = a + c2 t1 + c1 t2 -/+ i (s2 t3 - s1 t4)   (forward)
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    const C a = v[0];
    const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    C m1, m2, n1, n2;
    m1.x = a.x + c1 * t1.x + c2 * t2.x;
    m1.y = a.y + c1 * t1.y + c2 * t2.y;
    m2.x = a.x + c2 * t1.x + c1 * t2.x;
    m2.y = a.y + c2 * t1.y + c1 * t2.y;
    n1.x = s1 * t3.x + s2 * t4.x;
    n1.y = s1 * t3.y + s2 * t4.y;
    n2.x = s2 * t3.x - s1 * t4.x;
    n2.y = s2 * t3.y - s1 * t4.y;
    n1 = cmi(n1, inv);
    n2 = cmi(n2, inv);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  } else if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const C t2 = cadd(v[1], v[3]), t3 = cmi(csub(v[1], v[3]), inv);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else {
    // X[2q] = DFT4(v_r + v_{r+4})[q], X[2q+1] = DFT4((v_r - v_{r+4}) W8^r)[q]
    const T h = (T)0.70710678118654752440;
    C a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = cadd(v[r], v[r + 4]);
      b[r] = csub(v[r], v[r + 4]);
    }
    // W8 = (h, -h) forward, (h, h) inverse; W8^2 = -i / +i; W8^3 = (-h, -h) / (-h, h)
    {
      const C x = b[1];
      b[1].x = h * (x.x + (inv ? -x.y : x.y));
      b[1].y = h * (x.y + (inv ? x.x : -x.x));
    }
    b[2] = cmi(b[2], inv);
    {
      const C x = b[3];
      b[3].x = h * (-x.x + (inv ? -x.y : x.y));
      b[3].y = h * (-x.y + (inv ? x.x : -x.x));
    }
    small_dft<C, 4>(a, inv);
    small_dft<C, 4>(b, inv);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
}

// Twiddles and DFT of butterfly jj (of N / R per line) of a radix-R Stockham
// pass, Ns = product of the radices already applied.  Its inputs are the
// points jj + r N / R; its outputs go to (jj / Ns) Ns R + jj % Ns + r Ns.
template <typename C, int N, int R>
__device__ __forceinline__ void stockham_bfly(C (&v)[R], const double2* __restrict__ tw, int jj, int Ns, bool inv)
{
  const int kin = jj % Ns;
  const int step = N / (Ns * R);
  // w^r from one table entry by repeated products (one load per butterfly)
  const double2 w1 = tw[kin * step];
  double2 wr = w1;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    v[r] = cmul(v[r], wr, inv);
    if (r + 1 < R) wr = double2{wr.x * w1.x - wr.y * w1.y, wr.x * w1.y + wr.y * w1.x};
  }
  small_dft<C, R>(v, inv);
}

// One Stockham pass of radix R over the TI lines held in buf[k][ST] (row
// stride ST >= TI elements; Ns = product of the radices already applied).
template <typename C, int N, int R, int TI, int NT, int ST = TI>
__device__ __forceinline__ void stockham_pass(C* buf, const double2* __restrict__ tw, int Ns, bool inv)
{
  constexpr int NB = TI * N / R;             // butterflies in the work-group
  constexpr int BPT = (NB + NT - 1) / NT;    // per thread
  C v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + NT * q;
    if (b < NB) {
      const int col = b % TI, jj = b / TI;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(jj + r * (N / R)) * ST + col];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + NT * q;
    if (b < NB) {
      const int col = b % TI, jj = b / TI;
      stockham_bfly<C, N, R>(v[q], tw, jj, Ns, inv);
      const int d = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(d + r * Ns) * ST + col] = v[q][r];
    }
  }
  __syncthreads();
}

// Lengths N = 2^a 3^b 5^c: the power-of-two part in radix-8 passes (then one
// radix-4 or radix-2 pass), then radix-3 and radix-5 passes; Stockham's
// autosort works for any order of the factors.
constexpr int pow_part(int n, int p) { return n % p ? 1 : p * pow_part(n / p, p); }
constexpr int count_factor(int n, int p) { return n % p ? 0 : 1 + count_factor(n / p, p); }
constexpr bool smooth235(int n) { return n == pow_part(n, 2) * pow_part(n, 3) * pow_part(n, 5); }

template <typename C, int N, int TI, int NT, int ST = TI>
__device__ __forceinline__ void line_dft(C* buf, const double2* __restrict__ tw, bool inv)
{
  static_assert(smooth235(N), "line lengths are 2^a 3^b 5^c");
  int Ns = 1;
  constexpr int L = __builtin_ctz(N), L3 = count_factor(N, 3), L5 = count_factor(N, 5);
#pragma unroll
  for (int p = 0; p < L / 3; ++p) {
    stockham_pass<C, N, 8, TI, NT, ST>(buf, tw, Ns, inv);
    Ns *= 8;
  }
  if constexpr (L % 3 == 2) {
    stockham_pass<C, N, 4, TI, NT, ST>(buf, tw, Ns, inv);
    Ns *= 4;
  }
  if constexpr (L % 3 == 1) {
    stockham_pass<C, N, 2, TI, NT, ST>(buf, tw, Ns, inv);
    Ns *= 2;
  }
#pragma unroll
  for (int p = 0; p < L3; ++p) {
    stockham_pass<C, N, 3, TI, NT, ST>(buf, tw, Ns, inv);
    Ns *= 3;
  }
#pragma unroll
  for (int p = 0; p < L5; ++p) {
    stockham_pass<C, N, 5, TI, NT, ST>(buf, tw, Ns, inv);
    Ns *= 5;
  }
}

// TI neighbouring x-columns per work-group of NT threads
template <typename C, int N, int TI, int NT>
__global__ __launch_bounds__(NT) void k_poisson_z(C* __restrict__ spec, const double2* __restrict__ tw,
                                                  const double* __restrict__ ktab, int nx, int ny,
                                                  double inv_ntot)
{
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  C* buf = reinterpret_cast<C*>(smem_raw);
  const int nxh = nx / 2 + 1;
  const int tiles = (nxh + TI - 1) / TI;
  const int j = blockIdx.x / tiles, i0 = (blockIdx.x - j * tiles) * TI;
  const long long S = (long long)nxh * ny;  // spectrum stride between z-planes
  const long long base = (long long)j * nxh + i0;
  const int tid = threadIdx.x;
  for (int e = tid; e < N * TI; e += NT) {
    const int k = e / TI, col = e % TI;
    C v;
    v.x = 0;
    v.y = 0;
    if (i0 + col < nxh) v = spec[k * S + base + col];
    buf[e] = v;
  }
  __syncthreads();
  line_dft<C, N, TI, NT>(buf, tw, false);
  // the division of the spectral kernel, same operation order
  const double* xk2 = ktab;
  const double* tx = xk2 + nx;
  const double* yk2 = tx + nx;
  const double* ty = yk2 + ny;
  const double* zk2 = ty + ny;
  const double* tz = zk2 + N;
  for (int e = tid; e < N * TI; e += NT) {
    const int k = e / TI, col = e % TI, i = i0 + col;
    if (i >= nxh) continue;
    const double tyz = ty[j] * tz[k], txz = tx[i] * tz[k], txy = tx[i] * ty[j];
    const double K = (xk2[i] * (tyz * tyz) + yk2[j] * (txz * txz)) + zk2[k] * (txy * txy);
    C v = buf[e];
    if (K < 1.0e-16) {
      v.x = 0;
      v.y = 0;
    } else {
      // one division per mode (the separate kernel divides each part)
      using R = typename CplxT<C>::R;
      const double q = inv_ntot / (-K);
      v.x = (R)((double)v.x * q);
      v.y = (R)((double)v.y * q);
    }
    buf[e] = v;
  }
  __syncthreads();
  line_dft<C, N, TI, NT>(buf, tw, true);
  for (int e = tid; e < N * TI; e += NT) {
    const int k = e / TI, col = e % TI;
    if (i0 + col < nxh) spec[k * S + base + col] = buf[e];
  }
}

// y stages for lines longer than the pipelined kernels take (768 and 1024
// points): one tile of TI x-columns of one z-plane per work-group, loaded,
// transformed forward (MODE 1) or inverse (MODE 2) in LDS, stored.
template <typename C, int N, int TI, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_y_plain(C* __restrict__ spec, const double2* __restrict__ tw, int nx, int ny)
{
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  C* buf = reinterpret_cast<C*>(smem_raw);
  const int nxh = nx / 2 + 1;
  const int tiles = (nxh + TI - 1) / TI;
  const int o = blockIdx.x / tiles, i0 = (blockIdx.x - o * tiles) * TI;
  const long long S = nxh;  // between y points
  const long long base = (long long)o * nxh * ny + i0;
  const int tid = threadIdx.x;
  for (int e = tid; e < N * TI; e += NT) {
    const int k = e / TI, col = e % TI;
    C v;
    v.x = 0;
    v.y = 0;
    if (i0 + col < nxh) v = spec[k * S + base + col];
    buf[e] = v;
  }
  __syncthreads();
  line_dft<C, N, TI, NT>(buf, tw, MODE == 2);
  for (int e = tid; e < N * TI; e += NT) {
    const int k = e / TI, col = e % TI;
    if (i0 + col < nxh) spec[k * S + base + col] = buf[e];
  }
}

// Persistent, software-pipelined line kernels over the half spectrum.  With
// one work-group per compute unit (a 128 KiB tile) the kernel above loads,
// then transforms, then stores, and the memory system idles during the
// transforms.  Here each work-group walks tiles blockIdx.x, +gridDim.x, ...,
// and the next tile's columns are loaded into registers while the current
// tile is transformed.  Vector-memory loads complete in issue order, so
// nothing the transforms wait on may be loaded from global memory after the
// prefetch: twiddles and the z tables are copied to LDS once, and the x and y
// factors of a tile are loaded before its successor's prefetch.
//   MODE 0 (z-lines): forward DFT, division, inverse DFT; the expressions of
//          k_poisson_z in the same order (products may fuse into
//          multiply-adds differently, so results agree to rounding).
//   MODE 1 / 2 (y-lines): the forward / inverse length-ny DFT alone, for
//          the y stages of the 3-D transform.  A tile is TI x-columns of
//          one z-plane, so its rows share the plane's pages.
// Tile t covers columns i0.. of outer line index o (y for MODE 0, z else).
template <typename C, int N, int TI, int MODE>
constexpr size_t spl_lds_bytes()
{
  return (size_t)N * TI * sizeof(C) + (size_t)(N / 2 > 0 ? N / 2 : 1) * sizeof(double2) +
         (MODE == 0 ? (size_t)2 * N * sizeof(double) : 0);
}

template <typename C, int N, int TI, int NT, int MODE, bool GRP = false>
__global__ __launch_bounds__(NT) void k_spec_lines(C* __restrict__ spec, const double2* __restrict__ tw,
                                                   const double* __restrict__ ktab, int nx, int ny, int nz,
                                                   double inv_ntot, const PoisGrid* __restrict__ grids,
                                                   const int* __restrict__ bgrid, long long gtiles)
{
#pragma clang fp contract(off)
  static_assert(NT % TI == 0, "a thread keeps one column");
  static_assert(MODE >= 0 && MODE <= 2, "z stage, forward or inverse y stage");
  constexpr int NE = N * TI;
  constexpr int E = (NE + NT - 1) / NT;  // elements per thread
  constexpr int NTW = N / 2 > 0 ? N / 2 : 1;
  // 512-point lines with one radix-8 butterfly per thread and pass: the
  // prefetched points of a thread (jj + 64 q of column col) are the inputs of
  // its first butterfly, and the outputs of its last forward butterfly are the
  // inputs of its first inverse one and, after the inverse, the points it
  // stores.  Those passes skip LDS, and the division is applied in registers.
  // (FP64 y-lines excepted: with only one transform after the prefetch,
  // which here starts one pass later, too little work hides it)
  constexpr bool REG = N == 512 && TI * (N / 8) == NT && !(MODE != 0 && sizeof(C) == 16);
  extern __shared__ unsigned char smem_raw[];
  C* buf = reinterpret_cast<C*>(smem_raw);
  double2* twl = reinterpret_cast<double2*>(smem_raw + sizeof(C) * NE);
  double* kz = reinterpret_cast<double*>(twl + NTW);  // MODE 0: zk2[0..N), then tz[0..N)
  const int tid = threadIdx.x, col = tid % TI;
  // the z tables of a grid into LDS (zk2, tz follow xk2, tx, yk2, ty in its
  // table block)
  auto load_kz = [&](const double* k, int gnx, int gny) {
    const double* zk2 = k + 2 * gnx + 2 * gny;
    for (int e = tid; e < N; e += NT) {
      kz[e] = zk2[e];
      kz[N + e] = zk2[N + e];
    }
  };
  for (int e = tid; e < NTW; e += NT) twl[e] = tw[e];
  if constexpr (MODE == 0 && !GRP) load_kz(ktab, nx, ny);
  // The tables are written by other wavefronts' threads: every wavefront
  // waits for them here, since the 512-point path's first butterfly reads
  // the twiddles before the tile loop's first barrier.
  __syncthreads();
  // A tile of one grid: TI columns from i0 of outer line index o (y for
  // MODE 0, z else), S elements between line points.  Grouped (GRP), the
  // tile is one of the grids' concatenated tile sequences and g its grid.
  // Element e = tid + NT q of a tile is point tid / TI + (NT / TI) q of
  // column col: a uniform row address plus this thread's 32-bit byte offset
  // voff (the host checks (NT / TI) S sizeof(C) < 4 GiB), so that no 64-bit
  // address stays live per element.
  struct Geo {
    C* p;  // column i0 of line point 0
    long long S;
    uint32_t voff;
    int i, o, nx, ny;
    bool in;
    const double* k;
    double inv;
  };
  // (one grid: the constants are computed once, the tile index is an int)
  const int nxh1 = nx / 2 + 1, tiles_x1 = (nxh1 + TI - 1) / TI;
  const long long S1 = MODE == 0 ? (long long)nxh1 * ny : (long long)nxh1;
  const long long OS1 = MODE == 0 ? (long long)nxh1 : (long long)nxh1 * ny;
  const uint32_t voff1 = ((uint32_t)(tid / TI) * (uint32_t)S1 + (uint32_t)col) * (uint32_t)sizeof(C);
  auto geo = [&](long long tile, int g) {
    Geo G;
    if constexpr (GRP) {
      const PoisGrid& R = grids[g];
      const long long lt = tile - R.t0;
      G.k = R.ktab;
      G.nx = R.nx;
      G.ny = R.ny;
      G.inv = R.inv_ntot;
      const int nxh = G.nx / 2 + 1, tiles_x = (nxh + TI - 1) / TI;
      G.o = (int)(lt / tiles_x);
      const int i0 = (int)(lt - (long long)G.o * tiles_x) * TI;
      G.S = MODE == 0 ? (long long)nxh * G.ny : (long long)nxh;
      const long long OS = MODE == 0 ? (long long)nxh : (long long)nxh * G.ny;
      G.p = reinterpret_cast<C*>(R.spec) + (long long)G.o * OS + i0;
      G.voff = ((uint32_t)(tid / TI) * (uint32_t)G.S + (uint32_t)col) * (uint32_t)sizeof(C);
      G.i = i0 + col;
      G.in = G.i < nxh;
    } else {
      G.k = ktab;
      G.nx = nx;
      G.ny = ny;
      G.inv = inv_ntot;
      const int t = (int)tile;
      G.o = t / tiles_x1;
      const int i0 = (t - G.o * tiles_x1) * TI;
      G.S = S1;
      G.p = spec + (long long)G.o * OS1 + i0;
      G.voff = voff1;
      G.i = i0 + col;
      G.in = G.i < nxh1;
    }
    return G;
  };
  auto at = [&](const Geo& G, int q) -> C& {
    C* row = G.p + (long long)(NT / TI) * q * G.S;
    return *reinterpret_cast<C*>(reinterpret_cast<char*>(row) + G.voff);
  };
  C P[E];
  auto fetch = [&](long long tile, int g) {
    const Geo G = geo(tile, g);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q;
      P[q].x = 0;
      P[q].y = 0;
      if (e < NE && G.in) P[q] = at(G, q);
    }
  };
  // One grid: work-group b walks tiles b, b + gridDim.x, ...  Grouped: a
  // contiguous run of tiles per work-group, starting in grid bgrid[b]; the
  // next tile is in the same grid or the next one (every grid has tiles).
  long long tile, tend, step;
  int g = 0;
  if constexpr (GRP) {
    const long long per = (gtiles + gridDim.x - 1) / gridDim.x;
    tile = (long long)blockIdx.x * per;
    tend = tile + per < gtiles ? tile + per : gtiles;
    step = 1;
    g = bgrid[blockIdx.x];
  } else {
    tile = blockIdx.x;
    tend = (long long)tiles_x1 * (MODE == 0 ? ny : nz);
    step = gridDim.x;
  }
  auto grid_of = [&](long long t) {
    if constexpr (GRP) return t >= grids[g + 1].t0 ? g + 1 : g;
    else return 0;
  };
  int gk = -1;  // grouped MODE 0: the grid whose z tables are in LDS
  if (tile < tend) fetch(tile, g);
  for (; tile < tend; tile += step) {
    g = grid_of(tile);
    const Geo G = geo(tile, g);
    const bool in = G.in;
    if constexpr (GRP && MODE == 0) {
      // (uniform; the last readers of the previous tables passed the
      // previous tile's closing barrier, the next readers wait for the
      // barrier after the tile's LDS fill)
      if (g != gk) {
        load_kz(G.k, G.nx, G.ny);
        gk = g;
      }
    }
    // this tile's x and y factors, issued before the next prefetch
    double kxx = 0, txx = 0, kyy = 0, tyy = 0;
    if constexpr (MODE == 0) {
      const double* xk2 = G.k;
      const double* tx = xk2 + G.nx;
      const double* yk2 = tx + G.nx;
      const double* ty = yk2 + G.ny;
      if (in) {
        kxx = xk2[G.i];
        txx = tx[G.i];
      }
      kyy = yk2[G.o];
      tyy = ty[G.o];
    }
    const bool more = tile + step < tend;
    if constexpr (REG) {
      // radix-8 passes with Ns = 1, 8, 64; thread = butterfly jj of column col
      const int jj = tid / TI;
      C v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = P[r];  // points jj + 64 r
      constexpr bool inv0 = MODE == 2;  // z: forward first; y: one direction only
      stockham_bfly<C, N, 8>(v, twl, jj, 1, inv0);
#pragma unroll
      for (int r = 0; r < 8; ++r) buf[(8 * jj + r) * TI + col] = v[r];
      // the prefetch once v is in LDS, so that the two are not live together
      if (more) fetch(tile + step, grid_of(tile + step));
      __syncthreads();
      stockham_pass<C, N, 8, TI, NT>(buf, twl, 8, inv0);
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = buf[(jj + 64 * r) * TI + col];
      stockham_bfly<C, N, 8>(v, twl, jj, 64, inv0);  // outputs: points jj + 64 r
      if constexpr (MODE == 0) {
        __syncthreads();  // every thread has read buf before it is rewritten
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          if (!in) continue;
          const int k = jj + 64 * r;
          const double tzk = kz[N + k];
          const double tyz = tyy * tzk, txz = txx * tzk, txy = txx * tyy;
          const double K = (kxx * (tyz * tyz) + kyy * (txz * txz)) + kz[k] * (txy * txy);
          if (K < 1.0e-16) {
            v[r].x = 0;
            v[r].y = 0;
          } else {
            using R = typename CplxT<C>::R;
            const double qd = G.inv / (-K);
            v[r].x = (R)((double)v[r].x * qd);
            v[r].y = (R)((double)v[r].y * qd);
          }
        }
        // the inverse's first butterfly takes the same points jj + 64 r
        stockham_bfly<C, N, 8>(v, twl, jj, 1, true);
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[(8 * jj + r) * TI + col] = v[r];
        __syncthreads();
        stockham_pass<C, N, 8, TI, NT>(buf, twl, 8, true);
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = buf[(jj + 64 * r) * TI + col];
        stockham_bfly<C, N, 8>(v, twl, jj, 64, true);
      }
      if (in) {
#pragma unroll
        for (int r = 0; r < 8; ++r) at(G, r) = v[r];  // point jj + 64 r
      }
      // buf is refilled by the next tile
      __syncthreads();
      continue;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q;
      if (e < NE) buf[e] = P[q];
    }
    __syncthreads();
    if (more) fetch(tile + step, grid_of(tile + step));
    if constexpr (MODE == 0) line_dft<C, N, TI, NT>(buf, twl, false);
    if constexpr (MODE == 0) {
      // (rolled loops below: their state is a few registers, the prefetch
      // is not, and unrolling them spills)
#pragma unroll 1
      for (int q = 0; q < E; ++q) {
        const int e = tid + NT * q;
        if (e >= NE || !in) continue;
        const int k = e / TI;
        const double tzk = kz[N + k];
        const double tyz = tyy * tzk, txz = txx * tzk, txy = txx * tyy;
        const double K = (kxx * (tyz * tyz) + kyy * (txz * txz)) + kz[k] * (txy * txy);
        C v = buf[e];
        if (K < 1.0e-16) {
          v.x = 0;
          v.y = 0;
        } else {
          using R = typename CplxT<C>::R;
          const double qd = G.inv / (-K);
          v.x = (R)((double)v.x * qd);
          v.y = (R)((double)v.y * qd);
        }
        buf[e] = v;
      }
      __syncthreads();
    }
    line_dft<C, N, TI, NT>(buf, twl, MODE != 1);
#pragma unroll 1
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q;
      if (e < NE && in) at(G, q) = buf[e];
    }
    // buf is refilled by the next tile
    __syncthreads();
  }
}

// BFP32 fields (the word format of the BFP32 entry points: a 16-point
// x-segment shares exponent e, word = (q << 1) | b with q = rint(v 2^(30-e))
// and b = bit k of e + 128 in the segment's points k < 8) read and written two
// neighbouring points per lane: a segment is then 8 lanes, and its exponent
// bits sit in the first four lanes' two words each.
__device__ __forceinline__ unsigned bfp_spread4(unsigned x)
{
  x = (x | (x << 2)) & 0x33u;
  return (x | (x << 1)) & 0x55u;
}
// Values of this lane's two words (every lane of the segment takes part).
__device__ __forceinline__ double2 bfp_pair_value(int2 w)
{
  const unsigned long long m0 = __ballot(w.x & 1), m1 = __ballot(w.y & 1);
  const int g = (int)(threadIdx.x & 63) & ~7;
  const unsigned a = (unsigned)(m0 >> g) & 0xFu, b = (unsigned)(m1 >> g) & 0xFu;  // points 0,2,4,6 / 1,3,5,7
  const int eb = (int)(bfp_spread4(a) | (bfp_spread4(b) << 1));
  return double2{ldexp((double)(w.x >> 1), eb - 158), ldexp((double)(w.y >> 1), eb - 158)};
}
// Words of this lane's two values: the segment's exponent from the largest
// FP64 exponent field over its 8 lanes, as the single-point encoder forms it.
__device__ __forceinline__ int2 bfp_pair_word(double2 v)
{
  unsigned h = max((unsigned)__double2hiint(v.x) & 0x7FFFFFFFu, (unsigned)__double2hiint(v.y) & 0x7FFFFFFFu);
  h = max(h, (unsigned)__shfl_xor((int)h, 1));
  h = max(h, (unsigned)__shfl_xor((int)h, 2));
  h = max(h, (unsigned)__shfl_xor((int)h, 4));
  int e = (int)(h >> 20) - 1022;
  e = min(max(e, -126), 127);
  int q0 = (int)rint(ldexp(v.x, 30 - e)), q1 = (int)rint(ldexp(v.y, 30 - e));
  q0 = min(max(q0, -1073741824), 1073741823);
  q1 = min(max(q1, -1073741824), 1073741823);
  const unsigned eb = (unsigned)(e + 128), k = 2u * (threadIdx.x & 7u);  // points k, k + 1
  return int2{(int)(((unsigned)q0 << 1) | ((eb >> k) & 1u)), (int)(((unsigned)q1 << 1) | ((eb >> (k + 1)) & 1u))};
}

// Forward x transform (r2c) of every x-line: a real line of nx = 2M points
// is read as M complex points z_j = x_{2j} + i x_{2j+1} (its own memory,
// one 2-word load per point), one length-M DFT gives Z, and the half
// spectrum is X_k = E_k - i W^k O_k, k = 0..M, with
// E_k = (Z_k + conj Z_{M-k}) / 2, O_k = (Z_k - conj Z_{M-k}) / 2,
// W = e^{-2 pi i / nx} and Z_M = Z_0.  A tile is TI consecutive lines; each
// work-group copies the twiddles to LDS once and walks tiles blockIdx.x,
// +gridDim.x, ...  Loads and stores run along the lines (coalesced), and the
// column layout in LDS, buf[point * ST + line], has the odd row stride
// ST = TI + 1 so that those accesses spread over the banks.
template <typename C, int M, int TI, int NT, bool BFP = false, bool GRP = false>
__global__ __launch_bounds__(NT) void k_rfft_x(const void* __restrict__ rhs, C* __restrict__ spec,
                                               const double2* __restrict__ tw, long long nlines,
                                               const PoisGrid* __restrict__ grids, const int* __restrict__ bgrid,
                                               long long gtiles)
{
#pragma clang fp contract(off)
  using R = typename CplxT<C>::R;
  constexpr int ST = TI + 1;
  constexpr int NE = M * TI;
  constexpr int NTW = M / 2 > 0 ? M / 2 : 1;
  constexpr int NH = M + 1;
  extern __shared__ unsigned char smem_raw[];
  C* buf = reinterpret_cast<C*>(smem_raw);
  double2* twl = reinterpret_cast<double2*>(smem_raw + sizeof(C) * M * ST);  // e^{-2 pi i t / M}, t < M / 2
  double2* wn = twl + NTW;                                                   // W^k, k < M
  const int tid = threadIdx.x;
  for (int e = tid; e < NTW; e += NT) twl[e] = tw[2 * e];
  for (int e = tid; e < M; e += NT) wn[e] = tw[e];
  // (BFP32: the words of two points per lane; decoded when stored to LDS,
  // where the loads are waited on anyway)
  using PT = typename std::conditional<BFP, int2, C>::type;
  // A tile: TI lines from l0 of one grid with nl lines (grouped: of grid g)
  struct Geo {
    const PT* in;
    C* spec;
    long long l0, nl;
  };
  auto geo = [&](long long tile, int g) {
    Geo G{reinterpret_cast<const PT*>(rhs), spec, tile * TI, nlines};
    if constexpr (GRP) {
      const PoisGrid& Rg = grids[g];
      G.in = reinterpret_cast<const PT*>(Rg.rhs);
      G.spec = reinterpret_cast<C*>(Rg.spec);
      G.l0 = (tile - Rg.t0) * TI;
      G.nl = (long long)Rg.ny * Rg.nz;
    }
    return G;
  };
  // the next tile's points are loaded into registers while this one is
  // transformed (the transforms read LDS only, so they do not wait on them)
  constexpr int E = (NE + NT - 1) / NT;
  PT P[E];
  auto fetch = [&](long long tile, int g) {
    const Geo G = geo(tile, g);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q, lt = e / M, j = e - lt * M;
      P[q].x = 0;
      P[q].y = 0;
      if (e < NE && G.l0 + lt < G.nl) P[q] = G.in[(G.l0 + lt) * M + j];
    }
  };
  // one grid: work-group b walks tiles b, b + gridDim.x, ...; grouped: a
  // contiguous run of tiles from grid bgrid[b] on
  long long tile, tend, step;
  int g = 0;
  if constexpr (GRP) {
    const long long per = (gtiles + gridDim.x - 1) / gridDim.x;
    tile = (long long)blockIdx.x * per;
    tend = tile + per < gtiles ? tile + per : gtiles;
    step = 1;
    g = bgrid[blockIdx.x];
  } else {
    tile = blockIdx.x;
    tend = (nlines + TI - 1) / TI;
    step = gridDim.x;
  }
  auto grid_of = [&](long long t) {
    if constexpr (GRP) return t >= grids[g + 1].t0 ? g + 1 : g;
    else return 0;
  };
  if (tile < tend) fetch(tile, g);
  for (; tile < tend; tile += step) {
    g = grid_of(tile);
    const Geo G = geo(tile, g);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q, lt = e / M, j = e - lt * M;
      C v;
      if constexpr (BFP) v = bfp_pair_value(P[q]);  // every lane: NE is a multiple of 64
      else v = P[q];
      if (e < NE) buf[j * ST + lt] = v;
    }
    // (the first barrier also orders the twiddle copy before its readers)
    __syncthreads();
    if (tile + step < tend) fetch(tile + step, grid_of(tile + step));
    line_dft<C, M, TI, NT, ST>(buf, twl, false);
    for (int e = tid; e < NH * TI; e += NT) {
      const int lt = e / NH, k = e - lt * NH;
      if (G.l0 + lt >= G.nl) continue;
      const C a = buf[(k == M ? 0 : k) * ST + lt];
      const C b = buf[(k == 0 ? 0 : M - k) * ST + lt];  // Z_{M-k}, unconjugated
      const R ex = (a.x + b.x) * (R)0.5, ey = (a.y - b.y) * (R)0.5;
      const R ox = (a.x - b.x) * (R)0.5, oy = (a.y + b.y) * (R)0.5;
      const R wx = k == M ? (R)-1 : (R)wn[k].x, wy = k == M ? (R)0 : (R)wn[k].y;
      const R txx = wx * ox - wy * oy, tyy = wx * oy + wy * ox;  // W^k O_k
      C x;
      x.x = ex + tyy;
      x.y = ey - txx;
      G.spec[(G.l0 + lt) * NH + k] = x;
    }
    // buf is refilled by the next tile
    __syncthreads();
  }
}

// Inverse x transform (c2r, unnormalised like hipFFT's) of every x-line: the
// half spectrum X_0..X_M is folded into M complex points
// Z_k = (X_k + conj X_{M-k}) + i W^{-k} (X_k - conj X_{M-k}), k < M (the
// imaginary parts of X_0 and X_M ignored, as a c2r does), one inverse
// length-M DFT gives z_j, and the line's points are x_{2j} = Re z_j,
// x_{2j+1} = Im z_j, stored as one 2-word value.  A thread loads the pair
// X_k, X_{M-k} (k <= M/2: lanes read one run upwards and one downwards, both
// coalesced) and folds it in registers into Z_k and Z_{M-k}.  Tiles, LDS
// layout, persistence and prefetch as k_rfft_x.
template <typename C, int M, int TI, int NT, bool BFP = false, bool GRP = false>
__global__ __launch_bounds__(NT) void k_irfft_x(const C* __restrict__ spec, void* __restrict__ rhs,
                                                const double2* __restrict__ tw, long long nlines,
                                                const PoisGrid* __restrict__ grids, const int* __restrict__ bgrid,
                                                long long gtiles)
{
#pragma clang fp contract(off)
  using R = typename CplxT<C>::R;
  constexpr int ST = TI + 1;
  constexpr int NTW = M / 2 > 0 ? M / 2 : 1;
  constexpr int NH = M + 1;
  constexpr int NP = M / 2 + 1;  // pairs (k, M - k) per line
  constexpr int NEP = NP * TI;
  extern __shared__ unsigned char smem_raw[];
  C* buf = reinterpret_cast<C*>(smem_raw);
  double2* twl = reinterpret_cast<double2*>(smem_raw + sizeof(C) * M * ST);  // e^{-2 pi i t / M}, t < M / 2
  double2* wn = twl + NTW;                                                   // W^k, k <= M / 2
  const int tid = threadIdx.x;
  for (int e = tid; e < NTW; e += NT) twl[e] = tw[2 * e];
  for (int e = tid; e < NP; e += NT) wn[e] = tw[e];
  // the fold reads W^k before the tile loop's first barrier: every
  // wavefront waits here for the threads that wrote the tables
  __syncthreads();
  using OT = typename std::conditional<BFP, int2, C>::type;  // BFP32: two words per lane
  // A tile: TI lines from l0 of one grid with nl lines (grouped: of grid g)
  struct Geo {
    const C* spec;
    OT* out;
    long long l0, nl;
  };
  auto geo = [&](long long tile, int g) {
    Geo G{spec, reinterpret_cast<OT*>(rhs), tile * TI, nlines};
    if constexpr (GRP) {
      const PoisGrid& Rg = grids[g];
      G.spec = reinterpret_cast<const C*>(Rg.spec);
      G.out = reinterpret_cast<OT*>(Rg.rhs);
      G.l0 = (tile - Rg.t0) * TI;
      G.nl = (long long)Rg.ny * Rg.nz;
    }
    return G;
  };
  constexpr int E = (NEP + NT - 1) / NT;
  C PA[E], PB[E];  // X_k, X_{M-k}
  auto fetch = [&](long long tile, int g) {
    const Geo G = geo(tile, g);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q, lt = e / NP, kp = e - lt * NP;
      PA[q].x = 0;
      PA[q].y = 0;
      PB[q] = PA[q];
      if (e < NEP && G.l0 + lt < G.nl) {
        const C* row = G.spec + (G.l0 + lt) * NH;
        PA[q] = row[kp];
        PB[q] = row[M - kp];
      }
    }
  };
  // one grid: work-group b walks tiles b, b + gridDim.x, ...; grouped: a
  // contiguous run of tiles from grid bgrid[b] on
  long long tile, tend, step;
  int g = 0;
  if constexpr (GRP) {
    const long long per = (gtiles + gridDim.x - 1) / gridDim.x;
    tile = (long long)blockIdx.x * per;
    tend = tile + per < gtiles ? tile + per : gtiles;
    step = 1;
    g = bgrid[blockIdx.x];
  } else {
    tile = blockIdx.x;
    tend = (nlines + TI - 1) / TI;
    step = gridDim.x;
  }
  auto grid_of = [&](long long t) {
    if constexpr (GRP) return t >= grids[g + 1].t0 ? g + 1 : g;
    else return 0;
  };
  if (tile < tend) fetch(tile, g);
  for (; tile < tend; tile += step) {
    g = grid_of(tile);
    const Geo G = geo(tile, g);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + NT * q, lt = e / NP, kp = e - lt * NP;
      if (e >= NEP) continue;
      C xa = PA[q], xb = PB[q];
      if (kp == 0) {
        xa.y = 0;
        xb.y = 0;
      }
      const R wx = (R)wn[kp].x, wy = (R)wn[kp].y;  // W^k; W^{-k} = conj, W^{-(M-k)} = -W^k
      {
        // Z_k: a = X_k, b = conj X_{M-k}
        const R sx = xa.x + xb.x, sy = xa.y - xb.y;              // a + b
        const R dx = xa.x - xb.x, dy = xa.y + xb.y;              // a - b
        const R px = wx * dx + wy * dy, py = wx * dy - wy * dx;  // W^{-k} (a - b)
        C z;
        z.x = sx - py;
        z.y = sy + px;
        buf[kp * ST + lt] = z;
      }
      if (kp > 0 && 2 * kp < M) {
        // Z_{M-k}: a = X_{M-k}, b = conj X_k
        const R sx = xb.x + xa.x, sy = xb.y - xa.y;
        const R dx = xb.x - xa.x, dy = xb.y + xa.y;
        const R px = -(wx * dx - wy * dy), py = -(wx * dy + wy * dx);  // -W^k (a - b)
        C z;
        z.x = sx - py;
        z.y = sy + px;
        buf[(M - kp) * ST + lt] = z;
      }
    }
    __syncthreads();
    if (tile + step < tend) fetch(tile + step, grid_of(tile + step));
    line_dft<C, M, TI, NT, ST>(buf, twl, true);
    // (BFP32: every lane of a segment encodes; M TI is a multiple of 64)
    for (int e = tid; e < M * TI; e += NT) {
      const int lt = e / M, j = e - lt * M;
      OT w;
      if constexpr (BFP) w = bfp_pair_word(buf[j * ST + lt]);
      else w = buf[j * ST + lt];
      if (G.l0 + lt < G.nl) G.out[(G.l0 + lt) * M + j] = w;
    }
    // buf is refilled by the next tile
    __syncthreads();
  }
}

// Tile (measured at 512^3): 16 columns where 16 lines fit in 128 KiB of LDS,
// else 8; 1024 threads for tiles of 64 KiB and more (the work-group is then
// the compute unit's only one or one of two, and its own waves hide the
// latency), 256 otherwise.
template <typename C, int N>
struct LineTile {
  static constexpr int TI = ((size_t)N * 16 * sizeof(C) <= 131072) ? 16 : 8;
  static constexpr int NT = (size_t)N * TI * sizeof(C) >= 65536 ? 1024 : 256;
};

// Point stride of the line kernels (elements) and whether their 32-bit
// per-thread offsets, (NT / TI) strides, fit.
template <typename C, int N, int MODE>
static bool spl_off32_ok(const x3d_ctx* c)
{
  const long long S = (long long)(c->nx / 2 + 1) * (MODE == 0 ? c->ny : 1);
  return (long long)(LineTile<C, N>::NT / LineTile<C, N>::TI) * S * (long long)sizeof(C) < (1LL << 32);
}

// The pipelined kernels cover lines of up to 768 points whose LDS fits a
// compute unit (at 1024 points the prefetch spills).
template <typename C, int N, int MODE>
constexpr bool spl_ok()
{
  return N >= 8 && N <= 768 && spl_lds_bytes<C, N, LineTile<C, N>::TI, MODE>() <= 160 * 1024;
}

// Slot of a line kernel's resident work-groups per compute unit in the
// context: [FP64 / FP32][z stage / forward y stage / inverse y stage].
template <typename C, int MODE>
static int& spl_slot(x3d_ctx* c)
{
  return c->spl_wg[sizeof(C) == 16 ? 0 : 1][MODE];
}

// At context set-up, on the context's device: raise the line kernel's LDS
// limit and record its resident work-groups per compute unit, so that a
// launch (possibly under graph capture) makes no such call.
template <typename C, int N, int MODE>
static int spl_prepare_n(x3d_ctx* c)
{
  constexpr int TI = LineTile<C, N>::TI, NT = LineTile<C, N>::NT;
  constexpr size_t smem = spl_lds_bytes<C, N, TI, MODE>();
  const void* fn = (const void*)k_spec_lines<C, N, TI, NT, MODE>;
  if (smem > 65536) X3D_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  int nb = 0;
  X3D_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, NT, smem));
  spl_slot<C, MODE>(c) = nb > 0 ? nb : 1;
  return X3D_OK;
}

template <typename C, int MODE>
static int spl_prepare(x3d_ctx* c, int n)
{
  switch (n) {
#define X3D_SPL_N(NN) \
    case NN: if constexpr (spl_ok<C, NN, MODE>()) return spl_prepare_n<C, NN, MODE>(c); break;
    X3D_SPL_N(8) X3D_SPL_N(16) X3D_SPL_N(32) X3D_SPL_N(64) X3D_SPL_N(128) X3D_SPL_N(256) X3D_SPL_N(512)
    X3D_SPL_N(24) X3D_SPL_N(40) X3D_SPL_N(48) X3D_SPL_N(80) X3D_SPL_N(96) X3D_SPL_N(160) X3D_SPL_N(192)
    X3D_SPL_N(320) X3D_SPL_N(384) X3D_SPL_N(640) X3D_SPL_N(768) X3D_SPL_N(1024)
#undef X3D_SPL_N
    default: break;
  }
  return X3D_OK;
}

template <typename C, int N, int MODE>
static int launch_spl(x3d_ctx* c, C* spec, const void* twdev)
{
  constexpr int TI = LineTile<C, N>::TI, NT = LineTile<C, N>::NT;
  constexpr size_t smem = spl_lds_bytes<C, N, TI, MODE>();
  if (!spl_off32_ok<C, N, MODE>(c)) return X3D_ERR_UNSUPPORTED;
  const int per_cu = spl_slot<C, MODE>(c);
  if (per_cu <= 0) return X3D_ERR_HIP;  // set-up did not prepare this kernel
  // one launch wave of work-groups: compute units x resident work-groups each
  const int nxh = c->nx / 2 + 1;
  const long long ntiles = (long long)((nxh + TI - 1) / TI) * (MODE == 0 ? c->ny : c->nz);
  const long long cap = (long long)(c->ncu > 0 ? c->ncu : 1) * per_cu;
  const unsigned blocks = (unsigned)(ntiles < cap ? ntiles : cap);
  const double inv_ntot = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
  hipLaunchKernelGGL((k_spec_lines<C, N, TI, NT, MODE>), dim3(blocks), dim3(NT), smem, c->stream, spec,
                     reinterpret_cast<const double2*>(twdev), c->ktab, c->nx, c->ny, c->nz, inv_ntot,
                     (const PoisGrid*)nullptr, (const int*)nullptr, 0LL);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename C, int N>
static int launch_pz_n(x3d_ctx* c, C* spec)
{
  constexpr int TI = LineTile<C, N>::TI, NT = LineTile<C, N>::NT;
  if constexpr (spl_ok<C, N, 0>()) {
    if (!std::getenv("X3D_PZ_PLAIN") && spl_off32_ok<C, N, 0>(c)) return launch_spl<C, N, 0>(c, spec, c->ztw);
  }
  const int nxh = c->nx / 2 + 1;
  const unsigned blocks = (unsigned)(((nxh + TI - 1) / TI) * c->ny);
  const size_t smem = (size_t)N * TI * sizeof(C);
  const double inv_ntot = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
  const double2* tw = reinterpret_cast<const double2*>(c->ztw);
  if (smem > 65536)
    (void)hipFuncSetAttribute((const void*)k_poisson_z<C, N, TI, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
  hipLaunchKernelGGL((k_poisson_z<C, N, TI, NT>), dim3(blocks), dim3(NT), smem, c->stream, spec, tw, c->ktab,
                     c->nx, c->ny, inv_ntot);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename C, int N, int MODE>
static int launch_y_plain(x3d_ctx* c, C* spec)
{
  constexpr int TI = LineTile<C, N>::TI, NT = LineTile<C, N>::NT;
  constexpr size_t smem = (size_t)N * TI * sizeof(C);
  static_assert(smem <= 160 * 1024, "y tile fits the LDS");
  const int nxh = c->nx / 2 + 1;
  const unsigned blocks = (unsigned)(((nxh + TI - 1) / TI) * c->nz);
  if (smem > 65536)
    (void)hipFuncSetAttribute((const void*)k_y_plain<C, N, TI, NT, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
  hipLaunchKernelGGL((k_y_plain<C, N, TI, NT, MODE>), dim3(blocks), dim3(NT), smem, c->stream, spec,
                     reinterpret_cast<const double2*>(c->ytw), c->nx, c->ny);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

// y stages of the 3-D transform (MODE 1 forward, 2 inverse), for the y
// lengths ysplit_ok admits (768 and 1024 points: the one-tile kernel)
template <typename C, int MODE>
static int launch_poisson_y(x3d_ctx* c, C* spec)
{
  switch (c->ny) {
    case 768:
      if constexpr (spl_ok<C, 768, MODE>()) return launch_spl<C, 768, MODE>(c, spec, c->ytw);
      return launch_y_plain<C, 768, MODE>(c, spec);
    case 1024:
      if constexpr (spl_ok<C, 1024, MODE>()) return launch_spl<C, 1024, MODE>(c, spec, c->ytw);
      return launch_y_plain<C, 1024, MODE>(c, spec);
    case 8: return launch_spl<C, 8, MODE>(c, spec, c->ytw);
    case 16: return launch_spl<C, 16, MODE>(c, spec, c->ytw);
    case 32: return launch_spl<C, 32, MODE>(c, spec, c->ytw);
    case 64: return launch_spl<C, 64, MODE>(c, spec, c->ytw);
    case 128: return launch_spl<C, 128, MODE>(c, spec, c->ytw);
    case 256: return launch_spl<C, 256, MODE>(c, spec, c->ytw);
    case 512: return launch_spl<C, 512, MODE>(c, spec, c->ytw);
    case 24: return launch_spl<C, 24, MODE>(c, spec, c->ytw);
    case 40: return launch_spl<C, 40, MODE>(c, spec, c->ytw);
    case 48: return launch_spl<C, 48, MODE>(c, spec, c->ytw);
    case 80: return launch_spl<C, 80, MODE>(c, spec, c->ytw);
    case 96: return launch_spl<C, 96, MODE>(c, spec, c->ytw);
    case 160: return launch_spl<C, 160, MODE>(c, spec, c->ytw);
    case 192: return launch_spl<C, 192, MODE>(c, spec, c->ytw);
    case 320: return launch_spl<C, 320, MODE>(c, spec, c->ytw);
    case 384: return launch_spl<C, 384, MODE>(c, spec, c->ytw);
    default: return X3D_ERR_UNSUPPORTED;
  }
}

// The x r2c kernel for nx = 2M: its tile (16 lines; 1024 threads from 128
// points on, which fills a compute unit with two FP64 or FP32 work-groups),
// the LDS of the tile and tables, and the set-up call that raises its LDS
// limit (outside any capture).
template <typename C, int M>
struct RfxTile {
  static constexpr int TI = LineTile<C, M>::TI;
  static constexpr int NT = M >= 128 ? 1024 : LineTile<C, M>::NT;
};
template <typename C, int M, bool INV>
constexpr size_t rfx_lds_bytes()
{
  return (size_t)M * (RfxTile<C, M>::TI + 1) * sizeof(C) + (size_t)((M / 2 > 0 ? M / 2 : 1) + M) * sizeof(double2);
}

template <typename C, int M, bool INV, bool BFP>
static int rfx_run(x3d_ctx* c, void* rhs, C* spec, bool prepare)
{
  constexpr int TI = RfxTile<C, M>::TI, NT = RfxTile<C, M>::NT;
  constexpr size_t smem = rfx_lds_bytes<C, M, INV>();
  static_assert(smem <= 160 * 1024, "x r2c / c2r tile fits the LDS");
  static_assert(!BFP || (sizeof(C) == 16 && M % 8 == 0), "BFP32: FP64 arithmetic, whole 16-point segments");
  const void* fn = INV ? (const void*)k_irfft_x<C, M, TI, NT, BFP> : (const void*)k_rfft_x<C, M, TI, NT, BFP>;
  int& per_cu = c->rfx_wg[BFP ? 2 : (sizeof(C) == 16 ? 0 : 1)][INV ? 1 : 0];
  if (prepare) {
    if (smem > 65536) X3D_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    int nb = 0;
    X3D_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, NT, smem));
    per_cu = nb > 0 ? nb : 1;
    return X3D_OK;
  }
  if (per_cu <= 0) return X3D_ERR_HIP;  // set-up did not prepare this kernel
  const long long nlines = (long long)c->ny * c->nz;
  const long long ntiles = (nlines + TI - 1) / TI;
  // one launch wave of work-groups: compute units x resident work-groups each
  const long long cap = (long long)(c->ncu > 0 ? c->ncu : 1) * per_cu;
  const long long blocks = ntiles < cap ? ntiles : cap;
  if constexpr (INV)
    hipLaunchKernelGGL((k_irfft_x<C, M, TI, NT, BFP>), dim3((unsigned)blocks), dim3(NT), smem, c->stream, spec, rhs,
                       reinterpret_cast<const double2*>(c->xtw), nlines, (const PoisGrid*)nullptr,
                       (const int*)nullptr, 0LL);
  else
    hipLaunchKernelGGL((k_rfft_x<C, M, TI, NT, BFP>), dim3((unsigned)blocks), dim3(NT), smem, c->stream,
                       (const void*)rhs, spec, reinterpret_cast<const double2*>(c->xtw), nlines,
                       (const PoisGrid*)nullptr, (const int*)nullptr, 0LL);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

// x r2c (INV false: rhs -> spec) or c2r (true: spec -> rhs), nx = 2M with M
// one of the line kernels' lengths up to 512
template <typename C, bool INV, bool BFP = false>
static int launch_rfx(x3d_ctx* c, void* rhs, C* spec, bool prepare = false)
{
  switch (c->nx / 2) {
#define X3D_RFX(MM)                                                                   \
  case MM:                                                                            \
    if constexpr (!BFP || MM % 8 == 0) return rfx_run<C, MM, INV, BFP>(c, rhs, spec, prepare); \
    break;
    X3D_RFX(8) X3D_RFX(16) X3D_RFX(32) X3D_RFX(64) X3D_RFX(128) X3D_RFX(256) X3D_RFX(512)
    X3D_RFX(24) X3D_RFX(40) X3D_RFX(48) X3D_RFX(80) X3D_RFX(96) X3D_RFX(160) X3D_RFX(192) X3D_RFX(320)
    X3D_RFX(384)
#undef X3D_RFX
    default: return X3D_ERR_UNSUPPORTED;
  }
  return X3D_ERR_UNSUPPORTED;
}

template <typename C>
static int launch_poisson_z(x3d_ctx* c, C* spec)
{
  switch (c->nz) {
    case 8: return launch_pz_n<C, 8>(c, spec);
    case 16: return launch_pz_n<C, 16>(c, spec);
    case 32: return launch_pz_n<C, 32>(c, spec);
    case 64: return launch_pz_n<C, 64>(c, spec);
    case 128: return launch_pz_n<C, 128>(c, spec);
    case 256: return launch_pz_n<C, 256>(c, spec);
    case 512: return launch_pz_n<C, 512>(c, spec);
    case 1024: return launch_pz_n<C, 1024>(c, spec);
    case 24: return launch_pz_n<C, 24>(c, spec);
    case 40: return launch_pz_n<C, 40>(c, spec);
    case 48: return launch_pz_n<C, 48>(c, spec);
    case 80: return launch_pz_n<C, 80>(c, spec);
    case 96: return launch_pz_n<C, 96>(c, spec);
    case 160: return launch_pz_n<C, 160>(c, spec);
    case 192: return launch_pz_n<C, 192>(c, spec);
    case 320: return launch_pz_n<C, 320>(c, spec);
    case 384: return launch_pz_n<C, 384>(c, spec);
    case 640: return launch_pz_n<C, 640>(c, spec);
    case 768: return launch_pz_n<C, 768>(c, spec);
    default: return X3D_ERR_UNSUPPORTED;
  }
}

#define X3D_TRY_RC(expr) do { const int _r = (expr); if (_r) return _r; } while (0)

// the lengths the line kernels are built for: powers of two and these
// 3- and 5-smooth ones
static bool smooth_len(int n)
{
  static const int extra[] = {24, 40, 48, 80, 96, 160, 192, 320, 384, 640, 768};
  if ((n & (n - 1)) == 0) return true;
  for (int e : extra)
    if (e == n) return true;
  return false;
}
static bool zsplit_ok(int nz) { return nz >= 8 && nz <= 1024 && smooth_len(nz); }
static bool ysplit_ok(int ny) { return ny >= 8 && ny <= 1024 && smooth_len(ny) && ny != 640; }
static bool xsplit_ok(int nx) { return nx % 2 == 0 && nx / 2 >= 8 && nx / 2 <= 512 && smooth_len(nx / 2); }

// Table of e^{-2 pi i t / n}, t < n, on the device.
static int twiddles(int n, void** dev)
{
  std::vector<double2> tw(n);
  for (int t = 0; t < n; ++t) {
    const double a = -2.0 * PI * (double)t / (double)n;
    tw[t] = double2{std::cos(a), std::sin(a)};
  }
  X3D_CHECK_HIP(hipMalloc(dev, tw.size() * sizeof(double2)));
  X3D_CHECK_HIP(hipMemcpy(*dev, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
  return X3D_OK;
}

// hipFFT plans of the stages it runs.  Forward: the 2-D (x, y) transform of
// every z-plane when the z stage is this file's kernel (hipFFT fuses the x
// r2c into one pass there, while its 1-D batched r2c takes two), else the
// whole 3-D transform.  Inverse: the x c2r of every row when the y stage is
// this file's kernel too, else the 2-D resp. 3-D transform.
static bool make_plans(const x3d_ctx* c, hipfftHandle* r2c, hipfftHandle* c2r, hipfftType tr, hipfftType tc)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = nx / 2 + 1;
  int n2[2] = {ny, nx}, re2[2] = {ny, nx}, ce2[2] = {ny, nxh};
  const int rd = nx * ny, cd = nxh * ny;
  if (c->zsplit) {
    if (hipfftPlanMany(r2c, 2, n2, re2, 1, rd, ce2, 1, cd, tr, nz) != HIPFFT_SUCCESS) return false;
  } else if (hipfftPlan3d(r2c, nz, ny, nx, tr) != HIPFFT_SUCCESS) {
    return false;
  }
  bool ok;
  if (c->ysplit) {
    int n1[1] = {nx}, re1[1] = {nx}, ce1[1] = {nxh};
    ok = hipfftPlanMany(c2r, 1, n1, ce1, 1, nxh, re1, 1, nx, tc, ny * nz) == HIPFFT_SUCCESS;
  } else if (c->zsplit) {
    ok = hipfftPlanMany(c2r, 2, n2, ce2, 1, cd, re2, 1, rd, tc, nz) == HIPFFT_SUCCESS;
  } else {
    ok = hipfftPlan3d(c2r, nz, ny, nx, tc) == HIPFFT_SUCCESS;
  }
  if (!ok) hipfftDestroy(*r2c);
  return ok;
}

// 1-D tables of abxyz (src/poisson.f90:436-476) and waves (:522-616):
// modified wavenumber squared of the staggered derivative and the
// interpolation transfer function, per axis, mirrored above n/2.
static void axis_tables(int n, double h, double L, const Scheme& s, std::vector<double>& a,
                        std::vector<double>& b, std::vector<double>& k2, std::vector<double>& e,
                        std::vector<double>& T)
{
  const double twopi = 2.0 * PI;
  a.resize(n); b.resize(n); k2.assign(n, 0.0); e.assign(n, 0.0); T.resize(n);
  for (int i = 0; i < n; ++i) {
    a[i] = std::sin((double)i * PI / (double)n);
    b[i] = std::cos((double)i * PI / (double)n);
  }
  for (int i = 0; i < n / 2 + 1; ++i) {
    const double w = twopi * (double)i / (double)n;
    double wp = s.aci * 2.0 * h * std::sin(w * 0.5) + s.bci * 2.0 * h * std::sin(3.0 * 0.5 * w);
    wp = wp / (1.0 + 2.0 * s.alcai * std::cos(w));
    const double v = (double)n * wp / L;
    k2[i] = v * v;
    e[i] = (double)n * w / L;
  }
  for (int i = n / 2 + 1; i < n; ++i) { k2[i] = k2[n - i]; e[i] = e[n - i]; }
  for (int i = 0; i < n; ++i) {
    const double rl = e[i] * h;
    const double tt = 2.0 * (s.bici * std::cos(rl * 1.5) + s.cici * std::cos(rl * 2.5) + s.dici * std::cos(rl * 3.5));
    const double tt1 = 2.0 * s.aici * std::cos(rl * 0.5);
    const double t1 = 1.0 + 2.0 * s.ailcai * std::cos(rl);
    T[i] = (tt1 + tt) / t1;
  }
}

int poisson_setup(x3d_ctx* c)
{
  std::vector<double> tx, ty, tz;
  axis_tables(c->nx, c->dx, c->xlx, c->sc[0], c->ax, c->bx, c->xk2, c->exs, tx);
  axis_tables(c->ny, c->dy, c->yly, c->sc[1], c->ay, c->by, c->yk2, c->eys, ty);
  axis_tables(c->nz, c->dz, c->zlz, c->sc[2], c->az, c->bz, c->zk2, c->ezs, tz);
  std::vector<double> h;
  h.insert(h.end(), c->xk2.begin(), c->xk2.end());
  h.insert(h.end(), tx.begin(), tx.end());
  h.insert(h.end(), c->yk2.begin(), c->yk2.end());
  h.insert(h.end(), ty.begin(), ty.end());
  h.insert(h.end(), c->zk2.begin(), c->zk2.end());
  h.insert(h.end(), tz.begin(), tz.end());
  X3D_CHECK_HIP(hipMalloc(&c->ktab, h.size() * sizeof(double)));
  X3D_CHECK_HIP(hipMemcpy(c->ktab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  const size_t ns = (size_t)(c->nx / 2 + 1) * c->ny * c->nz;
  X3D_CHECK_HIP(hipMalloc(&c->spec, ns * sizeof(double2)));
  c->zsplit = zsplit_ok(c->nz);
  // (the y kernel's 32-bit offsets span 64 rows of the half spectrum)
  c->ysplit = c->zsplit && ysplit_ok(c->ny) && (long long)(c->nx / 2 + 1) * 64 * 16 < (1LL << 32);
  // the forward x r2c as a line kernel as well (its 32-bit-free indexing
  // takes any size; X3D_XSPLIT=0 keeps hipFFT's 2-D forward plan)
  const char* xs = std::getenv("X3D_XSPLIT");
  c->xsplit = c->ysplit && xsplit_ok(c->nx) && !(xs && std::atoi(xs) == 0);
  if (c->zsplit) X3D_TRY_RC(twiddles(c->nz, &c->ztw));
  if (c->ysplit) X3D_TRY_RC(twiddles(c->ny, &c->ytw));
  if (c->xsplit) X3D_TRY_RC(twiddles(c->nx, &c->xtw));
  if (!make_plans(c, &c->fft_r2c, &c->fft_c2r, HIPFFT_D2Z, HIPFFT_Z2D)) return X3D_ERR_FFT;
  X3D_CHECK_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  if (c->zsplit) X3D_TRY_RC((spl_prepare<double2, 0>(c, c->nz)));
  if (c->ysplit) X3D_TRY_RC((spl_prepare<double2, 2>(c, c->ny)));
  if (c->xsplit) {
    X3D_TRY_RC((spl_prepare<double2, 1>(c, c->ny)));
    X3D_TRY_RC((launch_rfx<double2, false>(c, nullptr, nullptr, true)));
    if (c->nx % 16 == 0) {  // the BFP32 form's x kernels
      X3D_TRY_RC((launch_rfx<double2, false, true>(c, nullptr, nullptr, true)));
      X3D_TRY_RC((launch_rfx<double2, true, true>(c, nullptr, nullptr, true)));
    }
  }
  c->fft_ready = true;
  if (hipfftSetStream(c->fft_r2c, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  if (hipfftSetStream(c->fft_c2r, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  return X3D_OK;
}

void poisson_release(x3d_ctx* c)
{
  if (c->fft_ready) {
    hipfftDestroy(c->fft_r2c);
    hipfftDestroy(c->fft_c2r);
    c->fft_ready = false;
  }
  if (c->fft32_ready) {
    hipfftDestroy(c->fft_r2c32);
    hipfftDestroy(c->fft_c2r32);
    c->fft32_ready = false;
  }
  if (c->spec) (void)hipFree(c->spec);
  if (c->ktab) (void)hipFree(c->ktab);
  if (c->ztw) (void)hipFree(c->ztw);
  if (c->ytw) (void)hipFree(c->ytw);
  if (c->xtw) (void)hipFree(c->xtw);
  c->spec = nullptr;
  c->ktab = nullptr;
  c->ztw = nullptr;
  c->ytw = nullptr;
  c->xtw = nullptr;
}

// FP32 spectrum, FP64 division: the modified wavenumber and the quotient are
// formed in FP64 (the division near K = 0 is where FP32 would lose digits) and
// rounded once to FP32.
__global__ void k_spectral_f32(float2* __restrict__ spec, const double* __restrict__ ktab, int nx, int ny, int nz,
                               double inv_ntot)
{
#pragma clang fp contract(off)
  const int nxh = nx / 2 + 1;
  const long long N = (long long)nxh * ny * nz;
  const double* xk2 = ktab;
  const double* tx = xk2 + nx;
  const double* yk2 = tx + nx;
  const double* ty = yk2 + ny;
  const double* zk2 = ty + ny;
  const double* tz = zk2 + nz;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % nxh);
    const long long r = q / nxh;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const double tyz = ty[j] * tz[k], txz = tx[i] * tz[k], txy = tx[i] * ty[j];
    const double K = (xk2[i] * (tyz * tyz) + yk2[j] * (txz * txz)) + zk2[k] * (txy * txy);
    float2 v = spec[q];
    if (K < 1.0e-16) {
      v.x = 0.0f;
      v.y = 0.0f;
    } else {
      v.x = (float)(((double)v.x * inv_ntot) / (-K));
      v.y = (float)(((double)v.y * inv_ntot) / (-K));
    }
    spec[q] = v;
  }
}

int poisson_prepare_f32(x3d_ctx* c)
{
  if (!c->fft_ready) return X3D_ERR_FFT;
  if (!c->fft32_ready) {
    if (!make_plans(c, &c->fft_r2c32, &c->fft_c2r32, HIPFFT_R2C, HIPFFT_C2R)) return X3D_ERR_FFT;
    if (c->zsplit) X3D_TRY_RC((spl_prepare<float2, 0>(c, c->nz)));
    if (c->ysplit) X3D_TRY_RC((spl_prepare<float2, 2>(c, c->ny)));
    if (c->xsplit) {
      X3D_TRY_RC((spl_prepare<float2, 1>(c, c->ny)));
      X3D_TRY_RC((launch_rfx<float2, false>(c, nullptr, nullptr, true)));
      X3D_TRY_RC((launch_rfx<float2, true>(c, nullptr, nullptr, true)));
    }
    c->fft32_ready = true;
  }
  return X3D_OK;
}

int poisson_solve_f32(x3d_ctx* c, float* rhs)
{
  const int rc = poisson_prepare_f32(c);
  if (rc) return rc;
  if (hipfftSetStream(c->fft_r2c32, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  if (hipfftSetStream(c->fft_c2r32, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  // the FP64 spectrum buffer holds the FP32 one
  float2* spec = reinterpret_cast<float2*>(c->spec);
  if (c->xsplit) {
    X3D_TRY_RC((launch_rfx<float2, false>(c, rhs, spec)));
    X3D_TRY_RC((launch_poisson_y<float2, 1>(c, spec)));
  } else if (hipfftExecR2C(c->fft_r2c32, rhs, (hipfftComplex*)spec) != HIPFFT_SUCCESS) {
    return X3D_ERR_FFT;
  }
  if (c->zsplit) {
    X3D_TRY_RC(launch_poisson_z<float2>(c, spec));
    if (c->ysplit) X3D_TRY_RC((launch_poisson_y<float2, 2>(c, spec)));
    if (c->xsplit) return launch_rfx<float2, true>(c, rhs, spec);
    if (hipfftExecC2R(c->fft_c2r32, (hipfftComplex*)spec, rhs) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
    return X3D_OK;
  }
  const long long N = (long long)(c->nx / 2 + 1) * c->ny * c->nz;
  long long blocks = (N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const double inv_ntot = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
  hipLaunchKernelGGL(k_spectral_f32, dim3((unsigned)blocks), dim3(256), 0, c->stream, spec, c->ktab, c->nx, c->ny,
                     c->nz, inv_ntot);
  X3D_CHECK_HIP(hipGetLastError());
  if (hipfftExecC2R(c->fft_c2r32, (hipfftComplex*)spec, rhs) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  return X3D_OK;
}

int poisson_solve(x3d_ctx* c, double* rhs)
{
  if (!c->fft_ready) return X3D_ERR_FFT;
  if (c->xsplit) {
    X3D_TRY_RC((launch_rfx<double2, false>(c, rhs, c->spec)));
    X3D_TRY_RC((launch_poisson_y<double2, 1>(c, c->spec)));
  } else if (hipfftExecD2Z(c->fft_r2c, rhs, (hipfftDoubleComplex*)c->spec) != HIPFFT_SUCCESS) {
    return X3D_ERR_FFT;
  }
  if (c->zsplit) {
    X3D_TRY_RC(launch_poisson_z<double2>(c, c->spec));
    if (c->ysplit) X3D_TRY_RC((launch_poisson_y<double2, 2>(c, c->spec)));
    // (the x c2r stays hipFFT's in FP64: 0.403 ms at 512^3 against 0.455 ms
    // for k_irfft_x, which FP32 runs at 0.245 ms against 0.254)
    if (hipfftExecZ2D(c->fft_c2r, (hipfftDoubleComplex*)c->spec, rhs) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
    return X3D_OK;
  }
  const long long N = (long long)(c->nx / 2 + 1) * c->ny * c->nz;
  long long blocks = (N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const double inv_ntot = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
  hipLaunchKernelGGL(k_spectral, dim3((unsigned)blocks), dim3(256), 0, c->stream, c->spec, c->ktab,
                     c->nx, c->ny, c->nz, inv_ntot);
  X3D_CHECK_HIP(hipGetLastError());
  if (hipfftExecZ2D(c->fft_c2r, (hipfftDoubleComplex*)c->spec, rhs) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  return X3D_OK;
}

int poisson_solve_bfp(x3d_ctx* c, int32_t* rhs)
{
  if (!c->fft_ready || !c->xsplit || c->nx % 16) return X3D_ERR_UNSUPPORTED;
  X3D_TRY_RC((launch_rfx<double2, false, true>(c, rhs, c->spec)));
  X3D_TRY_RC((launch_poisson_y<double2, 1>(c, c->spec)));
  X3D_TRY_RC(launch_poisson_z<double2>(c, c->spec));
  X3D_TRY_RC((launch_poisson_y<double2, 2>(c, c->spec)));
  return launch_rfx<double2, true, true>(c, rhs, c->spec);
}

// ---- grouped Poisson of the batching form ----------------------------------
// Per grid the stages are those of poisson_solve / poisson_solve_f32 with
// every stage a line kernel: x r2c, forward y, z (forward, division,
// inverse), inverse y, x c2r (here the line kernel in FP64 too: at batch
// sizes one launch over many grids beats a hipFFT call per grid).  A grouped
// launch covers every grid with the stage's line length: its work-groups
// take contiguous runs of the grids' concatenated tiles.

// The grouped kernel of a stage for line length n (x stages: n = nx / 2),
// its threads, LDS bytes and tile width.  Lines of up to 384 points: a grid
// with longer ones fills the GPU on its own and is solved by itself (the
// grouped 512-point FP64 kernels would spill).
struct GrpKernel {
  const void* fn;
  int nt;
  size_t smem;
  int ti;
};

template <typename C, int N, int MODE>
static bool grp_spec_n(GrpKernel* k)
{
  if constexpr (spl_ok<C, N, MODE>()) {
    constexpr int TI = LineTile<C, N>::TI, NT = LineTile<C, N>::NT;
    *k = GrpKernel{(const void*)k_spec_lines<C, N, TI, NT, MODE, true>, NT, spl_lds_bytes<C, N, TI, MODE>(), TI};
    return true;
  }
  return false;
}

template <typename C, int M, bool INV>
static bool grp_rfx_m(GrpKernel* k)
{
  constexpr int TI = RfxTile<C, M>::TI, NT = RfxTile<C, M>::NT;
  const void* fn = INV ? (const void*)k_irfft_x<C, M, TI, NT, false, true>
                       : (const void*)k_rfft_x<C, M, TI, NT, false, true>;
  *k = GrpKernel{fn, NT, rfx_lds_bytes<C, M, INV>(), TI};
  return true;
}

template <typename C, int MODE>
static bool grp_spec(int n, GrpKernel* k)
{
  switch (n) {
#define X3D_GRP_N(NN) \
    case NN: return grp_spec_n<C, NN, MODE>(k);
    X3D_GRP_N(8) X3D_GRP_N(16) X3D_GRP_N(32) X3D_GRP_N(64) X3D_GRP_N(128) X3D_GRP_N(256)
    X3D_GRP_N(24) X3D_GRP_N(40) X3D_GRP_N(48) X3D_GRP_N(80) X3D_GRP_N(96) X3D_GRP_N(160) X3D_GRP_N(192)
    X3D_GRP_N(320) X3D_GRP_N(384)
#undef X3D_GRP_N
    default: return false;
  }
}

template <typename C, bool INV>
static bool grp_rfx(int m, GrpKernel* k)
{
  switch (m) {
#define X3D_GRP_M(MM) \
    case MM: return grp_rfx_m<C, MM, INV>(k);
    X3D_GRP_M(8) X3D_GRP_M(16) X3D_GRP_M(32) X3D_GRP_M(64) X3D_GRP_M(128) X3D_GRP_M(256) X3D_GRP_M(24) X3D_GRP_M(40) X3D_GRP_M(48) X3D_GRP_M(80) X3D_GRP_M(96) X3D_GRP_M(160) X3D_GRP_M(192)
    X3D_GRP_M(320) X3D_GRP_M(384)
#undef X3D_GRP_M
    default: return false;
  }
}

// Stage s of grid c: its line length, kernel and tile count.
template <typename C>
static bool grp_stage(const x3d_ctx* c, int s, int* n, GrpKernel* k, long long* tiles)
{
  const int nxh = c->nx / 2 + 1;
  bool ok = false;
  switch (s) {
    case 0: ok = grp_rfx<C, false>(*n = c->nx / 2, k); break;
    case 1: ok = grp_spec<C, 1>(*n = c->ny, k); break;
    case 2: ok = grp_spec<C, 0>(*n = c->nz, k); break;
    case 3: ok = grp_spec<C, 2>(*n = c->ny, k); break;
    case 4: ok = grp_rfx<C, true>(*n = c->nx / 2, k); break;
    default: return false;
  }
  if (!ok) return false;
  if (s == 0 || s == 4) {
    *tiles = ((long long)c->ny * c->nz + k->ti - 1) / k->ti;
  } else {
    // the y / z kernels' 32-bit per-thread offsets span (NT / TI) rows
    const long long S = s == 2 ? (long long)nxh * c->ny : nxh;
    if ((long long)(k->nt / k->ti) * S * (long long)sizeof(C) >= (1LL << 32)) return false;
    *tiles = (long long)((nxh + k->ti - 1) / k->ti) * (s == 2 ? c->ny : c->nz);
  }
  return true;
}

template <typename C>
static int grp_build(x3d_batch* b)
{
  std::vector<int> grouped;
  for (int i = 0; i < b->ng; ++i) {
    const x3d_ctx* c = b->g[i];
    bool ok = c->fft_ready && c->xsplit && c->ysplit && c->zsplit;
    for (int s = 0; ok && s < 5; ++s) {
      int n;
      GrpKernel k;
      long long t;
      ok = grp_stage<C>(c, s, &n, &k, &t);
    }
    (ok ? grouped : b->psolo).push_back(i);
  }
  std::vector<unsigned char> host;
  auto put = [&host](const void* p, size_t n) {
    const size_t at = (host.size() + 255) & ~size_t(255);
    host.resize(at + n);
    std::memcpy(host.data() + at, p, n);
    return at;
  };
  for (int s = 0; s < 5; ++s) {
    std::map<int, std::vector<int>> groups;  // line length -> grids, in batch order
    for (int i : grouped) {
      int n;
      GrpKernel k;
      long long t;
      grp_stage<C>(b->g[i], s, &n, &k, &t);
      groups[n].push_back(i);
    }
    for (auto& kv : groups) {
      const std::vector<int>& gi = kv.second;
      x3d_ctx* c0 = b->g[gi[0]];
      GrpKernel k;
      int n;
      long long t;
      grp_stage<C>(c0, s, &n, &k, &t);
      if (k.smem > 65536)
        X3D_CHECK_HIP(hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)k.smem));
      int per_cu = 0;
      X3D_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.fn, k.nt, k.smem));
      std::vector<PoisGrid> rec;
      long long total = 0;
      for (int i : gi) {
        x3d_ctx* c = b->g[i];
        grp_stage<C>(c, s, &n, &k, &t);
        PoisGrid r{};
        r.spec = c->spec;
        r.rhs = b->field(6, i);
        r.ktab = c->ktab;
        r.t0 = total;
        r.nx = c->nx;
        r.ny = c->ny;
        r.nz = c->nz;
        r.inv_ntot = 1.0 / ((double)c->nx * (double)c->ny * (double)c->nz);
        rec.push_back(r);
        total += t;
      }
      PoisGrid end{};
      end.t0 = total;
      rec.push_back(end);
      // one launch wave; each work-group a contiguous run of `per` tiles
      const long long cap = (long long)(b->ncu > 0 ? b->ncu : 1) * (per_cu > 0 ? per_cu : 1);
      const long long blocks0 = total < cap ? total : cap;
      const long long per = (total + blocks0 - 1) / blocks0;
      const long long blocks = (total + per - 1) / per;
      std::vector<int> bg((size_t)blocks);
      int g = 0;
      for (long long w = 0; w < blocks; ++w) {
        while (rec[(size_t)g + 1].t0 <= w * per) ++g;
        bg[(size_t)w] = g;
      }
      x3d_batch::PLaunch L{};
      L.fn = k.fn;
      L.blocks = (unsigned)blocks;
      L.threads = (unsigned)k.nt;
      L.smem = k.smem;
      L.stage = s;
      L.tw = (s == 0 || s == 4) ? c0->xtw : (s == 2 ? c0->ztw : c0->ytw);
      L.grids = put(rec.data(), rec.size() * sizeof(PoisGrid));
      L.bgrid = put(bg.data(), bg.size() * sizeof(int));
      L.tiles = total;
      b->plaunch.push_back(L);
    }
  }
  if (!host.empty()) {
    X3D_CHECK_HIP(hipMalloc(&b->pargs, host.size()));
    X3D_CHECK_HIP(hipMemcpy(b->pargs, host.data(), host.size(), hipMemcpyHostToDevice));
  }
  return X3D_OK;
}

int poisson_batch_build(x3d_batch* b)
{
  if (b->pargs) (void)hipFree(b->pargs);
  b->pargs = nullptr;
  b->plaunch.clear();
  b->psolo.clear();
  const char* e = std::getenv("X3D_BATCH_PGROUP");
  if (e && std::atoi(e) == 0) {
    for (int i = 0; i < b->ng; ++i) b->psolo.push_back(i);
    return X3D_OK;
  }
  return b->prec == 8 ? grp_build<double2>(b) : grp_build<float2>(b);
}

static int grp_launch(x3d_batch* b, const x3d_batch::PLaunch& L, hipStream_t st)
{
  {
    const PoisGrid* grids = reinterpret_cast<const PoisGrid*>(static_cast<char*>(b->pargs) + L.grids);
    const int* bgrid = reinterpret_cast<const int*>(static_cast<char*>(b->pargs) + L.bgrid);
    const double2* tw = static_cast<const double2*>(L.tw);
    long long tiles = L.tiles;
    // (the single-grid arguments are unused by the grouped kernels)
    void* none = nullptr;
    const double* ktab = nullptr;
    int zero = 0;
    long long nlines = 0;
    double inv = 0.0;
    if (L.stage == 0 || L.stage == 4) {
      void* args[] = {&none, &none, &tw, &nlines, &grids, &bgrid, &tiles};
      X3D_CHECK_HIP(hipLaunchKernel(L.fn, dim3(L.blocks), dim3(L.threads), args, L.smem, st));
    } else {
      void* args[] = {&none, &tw, &ktab, &zero, &zero, &zero, &inv, &grids, &bgrid, &tiles};
      X3D_CHECK_HIP(hipLaunchKernel(L.fn, dim3(L.blocks), dim3(L.threads), args, L.smem, st));
    }
  }
  return X3D_OK;
}

int poisson_batch_run(x3d_batch* b, int ns)
{
  for (int s = 0; s < 5; ++s) {
    std::vector<const x3d_batch::PLaunch*> ls;
    for (const auto& L : b->plaunch)
      if (L.stage == s) ls.push_back(&L);
    if (ns <= 0 || ls.size() < 2) {
      for (const auto* L : ls) X3D_TRY_RC(grp_launch(b, *L, b->stream));
      continue;
    }
    // the stage's groups (disjoint grids) side by side; the next stage waits
    // for all of them
    X3D_TRY_RC(batch_fork(b));
    int rc = X3D_OK;
    for (size_t k = 0; !rc && k < ls.size(); ++k) rc = grp_launch(b, *ls[k], b->pstreams[k % ns]);
    X3D_TRY_RC(batch_join(b, rc));
  }
  return X3D_OK;
}

}  // namespace x3d

extern "C" int x3d_poisson_tables(const x3d_ctx* c, double* ax, double* bx, double* ay, double* by,
                                  double* az, double* bz, double* kre)
{
  if (!c) return X3D_ERR_ARG;
  if (!x3d::all_periodic(c)) return X3D_ERR_UNSUPPORTED;  // no Poisson tables were made
  auto cp = [](double* dst, const std::vector<double>& src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  cp(ax, c->ax); cp(bx, c->bx); cp(ay, c->ay); cp(by, c->by); cp(az, c->az); cp(bz, c->bz);
  if (kre) {
    // kxyz on the reference's spectral grid (nx, ny, nz/2+1), x fastest
    std::vector<double> tx(c->nx), ty(c->ny), tz(c->nz);
    std::vector<double> a, b, k2, e;
    x3d::axis_tables(c->nx, c->dx, c->xlx, c->sc[0], a, b, k2, e, tx);
    x3d::axis_tables(c->ny, c->dy, c->yly, c->sc[1], a, b, k2, e, ty);
    x3d::axis_tables(c->nz, c->dz, c->zlz, c->sc[2], a, b, k2, e, tz);
    const int nzh = c->nz / 2 + 1;
    for (int k = 0; k < nzh; ++k)
      for (int j = 0; j < c->ny; ++j)
        for (int i = 0; i < c->nx; ++i) {
          const double tyz = ty[j] * tz[k], txz = tx[i] * tz[k], txy = tx[i] * ty[j];
          kre[(size_t)i + (size_t)c->nx * ((size_t)j + (size_t)c->ny * k)] =
              (c->xk2[i] * (tyz * tyz) + c->yk2[j] * (txz * txz)) + c->zk2[k] * (txy * txy);
        }
  }
  return X3D_OK;
}
