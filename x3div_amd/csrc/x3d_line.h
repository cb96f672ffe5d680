// Device building blocks of the periodic line kernels (stencil right-hand
// sides, thread tiling and the two line solvers).
//
// A line of n points is cut into C chunks of M consecutive points, one chunk
// per thread.  Lane layouts: x-lines are contiguous, so consecutive lanes take
// consecutive chunks of one line (tl.line_fast = 0); y/z-lines are strided,
// so consecutive lanes take consecutive lines (16 lines = 128 B rows) and the
// chunks of a line are spread over 4 lane groups and over wavefronts.
//
// Solvers (both solve the reference's periodic tridiagonal system: Thomas
// sweeps on the matrix prepared by src/x3d_operator_1d.f90:301-323, then the
// Sherman-Morrison correction of src/thomas.f90:49-55):
//  * solve_table: the reference's sweeps with the operator's own f, s, w and
//    periodic vectors, reassociated into chunked first-order recurrences
//    (used for the Thomas entry points and for caller-supplied tables);
//  * solve_circ: the same periodic system [alfa, 1, alfa] solved with the
//    limit factorisation of its Thomas sweeps, gamma (1 - beta L)(1 - beta U)
//    with gamma = (1 + sqrt(1 - 4 alfa^2)) / 2 and beta = -alfa / gamma, and
//    the cyclic wrap closed exactly by a geometric series instead of the
//    rank-one correction.  It needs no per-point coefficient loads.
#pragma once

#include <type_traits>

#include "x3d_internal.h"

namespace x3d {

template <int KIND>
struct Halo { static constexpr int L = 0, R = 0; };
template <> struct Halo<K_D> { static constexpr int L = 2, R = 2; };
template <> struct Halo<K_DVP> { static constexpr int L = 1, R = 2; };
template <> struct Halo<K_DPV> { static constexpr int L = 2, R = 1; };
template <> struct Halo<K_IVP> { static constexpr int L = 3, R = 4; };
template <> struct Halo<K_IPV> { static constexpr int L = 4, R = 3; };
template <> struct Halo<K_DD> { static constexpr int L = 2, R = 2; };

template <typename T>
struct Coef { T a, b, c, d; };

// Right-hand side at window index q.
// D    derx_00      src/x3d_derive.f90:134-137
// DVP  derxvp       src/x3d_staggered.f90:48-51
// DPV  derxpv       src/x3d_staggered.f90:246-249
// IVP  interxvp     src/x3d_staggered.f90:143-148
// IPV  interxpv     src/x3d_staggered.f90:326-331
// DD   derxx        (declared only, src/x3d_derive.f90:54-59): sixth-order
//                   compact second derivative, a (u+1 - 2u + u-1) + b (u+2 - 2u + u-2)
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
  else if constexpr (KIND == K_DD)
    return k.a * ((win[q + 1] + win[q - 1]) - (T)2 * win[q]) + k.b * ((win[q + 2] + win[q - 2]) - (T)2 * win[q]);
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

struct Tiling {
  int G, Wc, Wl, LW, line_fast, fix;
  int xcd;  // single-grid launches: line blocks regrouped per XCD (xcd_block)
};

// Work-group b of a launch of nb runs on XCD b % 8: the dispatcher deals
// work-groups to the 8 XCDs in turn.  This gives XCD x the contiguous range
// of line blocks [x nb / 8, (x + 1) nb / 8) instead, so that the blocks that
// share rows (neighbouring x ranges of the same y/z rows, each row in its own
// 2 MiB page at 512^3) run on one XCD and share its L2 and address
// translation caches.  A bijection of [0, nb) for any nb.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Constants of the circulant solver for one alfa and one tiling.
template <typename T>
struct Circ {
  T beta;
  T close;       // 1 / (1 - beta^n)
  T pwMG;        // beta^(M G)
  T pwM2[7];     // beta^(M 2^k)
  T pw[33];      // beta^k, k = 0..M (M <= 32)
  T pMc[64];     // beta^(M c), c = 0..63, formed as pow_Mc forms it
};

struct LineThread {
  int lineLocal, g, wc, ls, LPW, line, c, C, j0, G, Wc, LW, lf;
  bool active;
  uint32_t boff;  // byte offset of the line's first point
  uint32_t bps;   // byte stride between consecutive points of the line
  __device__ __forceinline__ int lane_of(int gg) const { return lf ? lineLocal + gg * LW : lineLocal * G + gg; }
  __device__ __forceinline__ uint32_t at(int j) const { return boff + (uint32_t)j * bps; }
};

// Element load/store at a 32-bit byte offset from a uniform field base (fields
// are below 4 GiB; the host checks).  They are buffer instructions: the base
// lives in an SGPR resource descriptor and each access costs one VGPR offset,
// with constant offsets folded into the instruction.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void* p)
{
  // raw buffer, stride 0, full 32-bit range; dword 3 as for gfx9 data buffers
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xFFFFFFFFu, 0x00020000);
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ T ld(const T* base, uint32_t boff)
{
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(field_rsrc(base), boff, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(field_rsrc(base), boff, 0, 0));
}
template <typename T>
__device__ __forceinline__ void st(T* base, uint32_t boff, T v)
{
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), field_rsrc(base), boff, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), field_rsrc(base), boff, 0, 0);
}

// ---- field storage ------------------------------------------------------------
// A field is stored as FP64, FP32 or BFP32 elements (S); the line kernels
// compute in Store<S>::T.
//
// BFP32 (block floating point, 4 bytes per point): x is cut into segments of
// 16 points (nx a multiple of 16); the points of a segment share the exponent
// e, the smallest with |v| < 2^e over the segment (clamped to [-126, 127]).
// Each point is the 32-bit word (q << 1) | b with the 31-bit mantissa
//   q = rint(v 2^(30-e)), clamped to [-2^30, 2^30 - 1],  v = q 2^(e-30),
// so the error is at most 2^(e-31), and 2^(e-30) for values within half a unit
// of 2^e: at most 2^-29 of the segment's largest value (FP32 rounds each value
// to 2^-24 of itself); b, the low bit, is bit k of
// e + 128 in the segment's point k for k < 8 (0 in points 8-15).  A field is
// thus one int32 array of the grid's shape and the exponent travels with the
// mantissas.  The kernels hold a segment in one aligned 16-lane row of a
// wavefront (the 16 consecutive y/z-lines of a lane group, or 16 consecutive
// staged x points), read the exponent with one ballot and form it on store
// with a row maximum; every lane of a row must take part in both.
template <typename S>
struct Store { using T = S; };
template <>
struct Store<bfp32> { using T = double; };
template <typename S>
inline constexpr bool is_bfp = std::is_same<S, bfp32>::value;

// Value of this lane's word; the exponent comes from the words of the lane's
// 16-lane row.
__device__ __forceinline__ double bfp_value(int w)
{
  const unsigned long long m = __ballot(w & 1);
  const int eb = (int)((m >> (threadIdx.x & 48)) & 0xFFull);
  return ldexp((double)(w >> 1), eb - 158);  // (w >> 1) 2^((eb - 128) - 30)
}

// Largest value over this lane's 16-lane row (DPP row rotations; every lane
// of the row is active, so no lane reads a disabled one).
__device__ __forceinline__ unsigned row16_max(unsigned x)
{
  unsigned y;
  y = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true);  // row_ror:8
  x = x > y ? x : y;
  y = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x124, 0xF, 0xF, true);  // row_ror:4
  x = x > y ? x : y;
  y = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x122, 0xF, 0xF, true);  // row_ror:2
  x = x > y ? x : y;
  y = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x121, 0xF, 0xF, true);  // row_ror:1
  return x > y ? x : y;
}

// Word of v in this lane's row: the row's common exponent, then v's mantissa
// and this lane's exponent bit.  The exponent comes from the largest FP64
// exponent field in the row (high word without the sign): e = biased - 1022
// is the smallest with |v| < 2^e for normal values; zero and FP64 subnormals
// give -1022 and clamp to -126.  |v| < 2^e bounds the rounded mantissa by
// 2^30, the one value the clamp changes.
__device__ __forceinline__ int bfp_word(double v)
{
  const unsigned hi = (unsigned)__double2hiint(v) & 0x7FFFFFFFu;
  int e = (int)(row16_max(hi) >> 20) - 1022;
  e = min(max(e, -126), 127);
  int q = (int)rint(ldexp(v, 30 - e));
  q = min(max(q, -1073741824), 1073741823);
  const unsigned bit = ((unsigned)(e + 128) >> (threadIdx.x & 15)) & 1u;  // 0 in points 8-15: e + 128 < 256
  return (int)(((unsigned)q << 1) | bit);
}

template <typename S>
__device__ __forceinline__ typename Store<S>::T sld(const S* base, uint32_t boff)
{
  if constexpr (is_bfp<S>)
    return bfp_value(ld(reinterpret_cast<const int*>(base), boff));
  else
    return ld(base, boff);
}
// With BFP32 every lane of the row calls this (on or not).
template <typename S>
__device__ __forceinline__ void sst(S* base, uint32_t boff, typename Store<S>::T v, bool on = true)
{
  if constexpr (is_bfp<S>) {
    const int w = bfp_word(v);
    if (on) st(reinterpret_cast<int*>(base), boff, w);
  } else {
    if (on) st(base, boff, v);
  }
}

// DIRC >= 0 fixes the direction at compile time; otherwise geo.dir selects it.
// blk: the work-group's block of lines (blockIdx.x, or a virtual block when
// one launch serves several grids).
template <typename T, int DIRC = -1>
__device__ __forceinline__ LineThread line_thread(const Geom& geo, const Tiling& tl, int M, int blk)
{
  LineThread t;
  t.G = tl.G; t.Wc = tl.Wc; t.LW = tl.LW; t.lf = tl.line_fast;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wl = wave / tl.Wc;
  t.wc = wave - wl * tl.Wc;
  if (tl.line_fast) { t.lineLocal = lane % tl.LW; t.g = lane / tl.LW; }
  else { t.lineLocal = lane / tl.G; t.g = lane - t.lineLocal * tl.G; }
  t.LPW = tl.Wl * tl.LW;
  t.ls = wl * tl.LW + t.lineLocal;
  t.line = blk * t.LPW + t.ls;
  t.active = t.line < geo.nlines;
  t.C = tl.G * tl.Wc;
  t.c = t.wc * tl.G + t.g;
  t.j0 = t.c * M;
  // lanes past the last line read the last line (results are not stored), so
  // every load is unconditional
  const int ln = t.active ? t.line : geo.nlines - 1;
  // Line base and point stride are formed without branches: a three-way
  // if/else on the (uniform) direction here was compiled by the ROCm 7.2
  // AMDGPU backend so that the z path kept the y stride.
  const int dir = DIRC >= 0 ? DIRC : geo.dir;
  const uint32_t nx = (uint32_t)geo.nx, nxy = (uint32_t)geo.nxy, l = (uint32_t)ln;
  const uint32_t k = l / nx;
  const uint32_t ex = l * nx, ey = (l - k * nx) + k * nxy, ez = l;
  const uint32_t e0 = dir == 0 ? ex : (dir == 1 ? ey : ez);
  const uint32_t eps = dir == 0 ? 1u : (dir == 1 ? nx : nxy);
  t.boff = e0 * (uint32_t)sizeof(T);
  t.bps = eps * (uint32_t)sizeof(T);
  return t;
}

// Strided y/z windows in two halves, so that a window can be loaded early
// and formed where it is used: window_issue loads the storage words of the
// chunk points and, in the lane groups at a wavefront's ends (whose
// neighbouring chunks live in other wavefronts), of the halo points; the
// masked loads touch a quarter of the rows.  window_finish decodes them
// (BFP32; every lane of a 16-lane row takes part, and the halo condition is
// uniform over a row, whose lanes share a chunk group) and takes the other
// halo points from the lane groups holding those chunks.
template <typename S>
using RawT = typename std::conditional<is_bfp<S>, int, typename Store<S>::T>::type;

template <typename S>
__device__ __forceinline__ RawT<S> ld_raw(const S* base, uint32_t boff)
{
  return ld(reinterpret_cast<const RawT<S>*>(base), boff);
}

template <typename S>
__device__ __forceinline__ typename Store<S>::T from_raw(RawT<S> w)
{
  if constexpr (is_bfp<S>) return bfp_value(w);
  else return w;
}

template <typename S, int M, int HL, int HR>
__device__ __forceinline__ void window_issue(const S* __restrict__ in, const LineThread& t, int n,
                                             RawT<S> (&raw)[M + HL + HR])
{
#pragma unroll
  for (int k = 0; k < M; ++k) raw[HL + k] = ld_raw(in, t.at(t.j0 + k));
#pragma unroll
  for (int h = 1; h <= HL; ++h) {
    const int d = (h + M - 1) / M;  // chunks to the left
    raw[HL - h] = 0;
    if (t.g < d) {
      int j = t.j0 - h;
      j = j < 0 ? j + n : j;
      raw[HL - h] = ld_raw(in, t.at(j));
    }
  }
#pragma unroll
  for (int h = 1; h <= HR; ++h) {
    const int d = (h + M - 1) / M;  // chunks to the right
    raw[HL + M - 1 + h] = 0;
    if (t.g + d >= t.G) {
      int j = t.j0 + M - 1 + h;
      j = j >= n ? j - n : j;
      raw[HL + M - 1 + h] = ld_raw(in, t.at(j));
    }
  }
}

template <typename S, int M, int HL, int HR>
__device__ __forceinline__ void window_finish(const LineThread& t, const RawT<S> (&raw)[M + HL + HR],
                                              typename Store<S>::T (&win)[M + HL + HR])
{
  using T = typename Store<S>::T;
#pragma unroll
  for (int k = 0; k < M; ++k) win[HL + k] = from_raw<S>(raw[HL + k]);
#pragma unroll
  for (int h = 1; h <= HL; ++h) {
    const int d = (h + M - 1) / M;  // chunks to the left
    const int idx = M * d - h;      // index inside that chunk
    T v = __shfl(win[HL + idx], t.lane_of(t.g >= d ? t.g - d : t.g));
    if (t.g < d) v = from_raw<S>(raw[HL - h]);
    win[HL - h] = v;
  }
#pragma unroll
  for (int h = 1; h <= HR; ++h) {
    const int d = (h + M - 1) / M;  // chunks to the right
    const int idx = h - 1 - M * (d - 1);
    T v = __shfl(win[HL + idx], t.lane_of(t.g + d < t.G ? t.g + d : t.g));
    if (t.g + d >= t.G) v = from_raw<S>(raw[HL + M - 1 + h]);
    win[HL + M - 1 + h] = v;
  }
}

// u at points j0-HL .. j0+M+HR-1 of this thread's line (wrapped mod n).
// x-lines (consecutive lanes hold consecutive chunks): the M chunk points are
// loaded from one address with immediate offsets and the halo comes from the
// neighbouring lanes; strided y/z-lines load every point.
template <typename S, int M, int HL, int HR, bool LF>
__device__ __forceinline__ void load_window(const S* __restrict__ in, const LineThread& t, int n,
                                            typename Store<S>::T (&win)[M + HL + HR])
{
  using T = typename Store<S>::T;
  static_assert(LF || !is_bfp<S>, "BFP32 x-lines go through the staged tile");
  if constexpr (!LF) {
    const uint32_t o = t.at(t.j0);
#pragma unroll
    for (int k = 0; k < M; ++k) win[HL + k] = ld(in, o + (uint32_t)(k * sizeof(T)));
#pragma unroll
    for (int h = 1; h <= HL; ++h) {
      const int d = (h + M - 1) / M;           // chunks to the left
      const int idx = M * d - h;               // index inside that chunk
      const int src = t.lane_of((t.g - d % t.G + t.G) % t.G);
      win[HL - h] = __shfl(win[HL + idx], src);
    }
#pragma unroll
    for (int h = 1; h <= HR; ++h) {
      const int d = (h + M - 1) / M;           // chunks to the right
      const int idx = h - 1 - M * (d - 1);
      const int src = t.lane_of((t.g + d) % t.G);
      win[HL + M - 1 + h] = __shfl(win[HL + idx], src);
    }
  } else {
    // strided lines: the chunk points from memory; each halo point from the
    // lane group holding that chunk in this wavefront, and from memory only
    // in the groups at the wavefront's ends (their neighbours live in other
    // wavefronts), so the masked loads touch a quarter of the rows (the
    // values of window_issue + window_finish, here with each halo formed as
    // soon as its chunk is in, which holds fewer registers)
#pragma unroll
    for (int k = 0; k < M; ++k) win[HL + k] = sld(in, t.at(t.j0 + k));
#pragma unroll
    for (int h = 1; h <= HL; ++h) {
      const int d = (h + M - 1) / M;  // chunks to the left
      const int idx = M * d - h;      // index inside that chunk
      T v = __shfl(win[HL + idx], t.lane_of(t.g >= d ? t.g - d : t.g));
      if (t.g < d) {
        int j = t.j0 - h;
        j = j < 0 ? j + n : j;
        v = sld(in, t.at(j));
      }
      win[HL - h] = v;
    }
#pragma unroll
    for (int h = 1; h <= HR; ++h) {
      const int d = (h + M - 1) / M;  // chunks to the right
      const int idx = h - 1 - M * (d - 1);
      T v = __shfl(win[HL + idx], t.lane_of(t.g + d < t.G ? t.g + d : t.g));
      if (t.g + d >= t.G) {
        int j = t.j0 + M - 1 + h;
        j = j >= n ? j - n : j;
        v = sld(in, t.at(j));
      }
      win[HL + M - 1 + h] = v;
    }
  }
}

template <typename S, int M>
__device__ __forceinline__ void load_chunk(const S* __restrict__ in, const LineThread& t,
                                           typename Store<S>::T (&v)[M])
{
#pragma unroll
  for (int k = 0; k < M; ++k) v[k] = sld(in, t.at(t.j0 + k));
}

// (y/z-lines: a 16-lane row holds 16 consecutive lines, all active or all not)
template <typename S, int M>
__device__ __forceinline__ void store_chunk(S* out, const LineThread& t, const typename Store<S>::T (&v)[M])
{
  if (t.active) {
#pragma unroll
    for (int k = 0; k < M; ++k) sst(out, t.at(t.j0 + k), v[k]);
  }
}

// ---- x-lines staged through LDS ---------------------------------------------
// On x-lines a lane's chunk is M consecutive points, so direct chunk loads put
// 64 lanes on addresses M*sizeof(T) apart.  Instead each wavefront moves its
// block through LDS: its LW lines are one contiguous run of 64*M points, read
// and written with lane-consecutive (fully coalesced) accesses and
// redistributed to chunks in a per-wavefront LDS tile of 64 rows of P words
// (P = M or M + 1, whichever is odd, so chunk reads do not conflict on banks).
// Only lanes of the same wavefront touch a tile, so the hand-offs are
// wavefront barriers, not work-group barriers.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int M>
struct XPitch { static constexpr int P = (M & 1) ? M : M + 1; };

// LDS words (of T) of the x tiles of one work-group of nwaves wavefronts.
template <int M>
__host__ __device__ constexpr int x_tile_words() { return 64 * XPitch<M>::P; }

template <typename S, int M>
struct XBlock {
  typename Store<S>::T* tile;  // this wavefront's LDS tile [64][P]
  uint32_t base;  // byte offset of the block's first point
  uint32_t lim;   // byte offset of the field's last point (clamp for partial blocks)
  int valid;      // points of the block that belong to existing lines
  int lane;
};

// tiles: the work-group's x tiles, one per wavefront, in wavefront order.
template <typename S, int M>
__device__ __forceinline__ XBlock<S, M> x_block(const Geom& geo, const LineThread& t, typename Store<S>::T* tiles)
{
  XBlock<S, M> b;
  const int wave = threadIdx.x >> 6;
  b.lane = threadIdx.x & 63;
  b.tile = tiles + wave * x_tile_words<M>();
  const int first = t.line - t.lineLocal;  // first line of this wavefront
  const long long left = (long long)(geo.nlines - first) * geo.n;
  b.valid = left <= 0 ? 0 : (left >= 64 * M ? 64 * M : (int)left);
  b.base = (uint32_t)first * (uint32_t)geo.n * (uint32_t)sizeof(S);
  b.lim = ((uint32_t)geo.nlines * (uint32_t)geo.n - 1u) * (uint32_t)sizeof(S);
  return b;
}

// Two consecutive points in one buffer access (16 B for FP64, 8 B for FP32).
// The loaded vector is reinterpreted as a whole: with ROCm 7.2, extracting
// the 32-bit elements of a raw_buffer_load_b64 result one by one and
// bit-casting each to float compiled to a single-dword load whose value was
// used for both points.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef double f64x2v __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ void ld2(const T* base, uint32_t boff, T& a, T& b)
{
  if constexpr (sizeof(T) == 8) {
    const f64x2v v = __builtin_bit_cast(f64x2v, __builtin_amdgcn_raw_buffer_load_b128(field_rsrc(base), boff, 0, 0));
    a = v.x;
    b = v.y;
  } else {
    const f32x2v v = __builtin_bit_cast(f32x2v, __builtin_amdgcn_raw_buffer_load_b64(field_rsrc(base), boff, 0, 0));
    a = v.x;
    b = v.y;
  }
}
template <typename T>
__device__ __forceinline__ void st2(T* base, uint32_t boff, T a, T b)
{
  if constexpr (sizeof(T) == 8) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, f64x2v{a, b}), field_rsrc(base), boff, 0, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, f32x2v{a, b}), field_rsrc(base), boff, 0, 0);
  }
}

// Coalesced load of the block into the tile (point e of the block goes to
// row e / M, column e % M).  With M even each lane moves two neighbouring
// points per access (both in one chunk); lines have even length then, so
// pairs never straddle the end of the field.
// BFP32 moves one point per lane and access, so that a 16-lane row holds one
// segment.
template <typename S, int M>
__device__ __forceinline__ void x_stage_in(const S* __restrict__ in, const XBlock<S, M>& b)
{
  using T = typename Store<S>::T;
  constexpr int P = XPitch<M>::P;
  wave_lds_sync();
  if constexpr (M % 2 == 0 && !is_bfp<S>) {
    const uint32_t last = b.lim - (uint32_t)sizeof(T);  // first point of the field's last pair
#pragma unroll
    for (int k = 0; k < M / 2; ++k) {
      const int e = 2 * b.lane + 128 * k;
      uint32_t off = b.base + (uint32_t)e * (uint32_t)sizeof(T);
      off = off < last ? off : last;
      T v0, v1;
      ld2(in, off, v0, v1);
      T* dst = b.tile + (e / M) * P + e % M;
      dst[0] = v0;
      dst[1] = v1;
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = b.lane + 64 * k;
      uint32_t off = b.base + (uint32_t)e * (uint32_t)sizeof(S);
      off = off < b.lim ? off : b.lim;
      b.tile[(e / M) * P + e % M] = sld(in, off);
    }
  }
  wave_lds_sync();
}

// Window j0-HL .. j0+M+HR-1 of this lane's chunk from the staged tile; the
// halo is in the neighbouring chunks of the same line (wrapped mod n).
template <typename S, int M, int HL, int HR>
__device__ __forceinline__ void x_window(const XBlock<S, M>& b, const LineThread& t,
                                         typename Store<S>::T (&win)[M + HL + HR])
{
  using T = typename Store<S>::T;
  constexpr int P = XPitch<M>::P;
  const T* own = b.tile + b.lane * P;
#pragma unroll
  for (int k = 0; k < M; ++k) win[HL + k] = own[k];
  // the chunk d to the left / right, wrapped: the x plans keep the halo
  // inside the neighbouring ceil(4 / M) <= G chunks, so one correction does
#pragma unroll
  for (int h = 1; h <= HL; ++h) {
    const int d = (h + M - 1) / M;
    const int gl = t.g - d;
    const int src = t.lane_of(gl < 0 ? gl + t.G : gl);
    win[HL - h] = b.tile[src * P + (M * d - h)];
  }
#pragma unroll
  for (int h = 1; h <= HR; ++h) {
    const int d = (h + M - 1) / M;
    const int gr = t.g + d;
    const int src = t.lane_of(gr >= t.G ? gr - t.G : gr);
    win[HL + M - 1 + h] = b.tile[src * P + (h - 1 - M * (d - 1))];
  }
}

template <typename S, int M, int HL, int HR>
__device__ __forceinline__ void x_load_window(const S* __restrict__ in, const XBlock<S, M>& b,
                                              const LineThread& t, typename Store<S>::T (&win)[M + HL + HR])
{
  x_stage_in<S, M>(in, b);
  x_window<S, M, HL, HR>(b, t, win);
}

// x_stage_in in two halves (M even, not BFP32): the coalesced loads into
// registers, issued early, and their copy into the tile once it is free.
template <typename S, int M>
__device__ __forceinline__ void x_fetch(const S* __restrict__ in, const XBlock<S, M>& b,
                                        typename Store<S>::T (&r)[M])
{
  static_assert(M % 2 == 0 && !is_bfp<S>, "pairs of plain values");
  using T = typename Store<S>::T;
  const uint32_t last = b.lim - (uint32_t)sizeof(T);
#pragma unroll
  for (int k = 0; k < M / 2; ++k) {
    const int e = 2 * b.lane + 128 * k;
    uint32_t off = b.base + (uint32_t)e * (uint32_t)sizeof(T);
    off = off < last ? off : last;
    ld2(in, off, r[2 * k], r[2 * k + 1]);
  }
}
template <typename S, int M>
__device__ __forceinline__ void x_put(const XBlock<S, M>& b, const typename Store<S>::T (&r)[M])
{
  constexpr int P = XPitch<M>::P;
  wave_lds_sync();
#pragma unroll
  for (int k = 0; k < M / 2; ++k) {
    const int e = 2 * b.lane + 128 * k;
    auto* dst = b.tile + (e / M) * P + e % M;
    dst[0] = r[2 * k];
    dst[1] = r[2 * k + 1];
  }
  wave_lds_sync();
}

// This lane's chunk of a field, through the tile.
template <typename S, int M>
__device__ __forceinline__ void x_load_chunk(const S* __restrict__ in, const XBlock<S, M>& b,
                                             typename Store<S>::T (&v)[M])
{
  using T = typename Store<S>::T;
  x_stage_in<S, M>(in, b);
  const T* own = b.tile + b.lane * XPitch<M>::P;
#pragma unroll
  for (int k = 0; k < M; ++k) v[k] = own[k];
}

// Chunk values to the tile, then coalesced stores of the block's valid points.
template <typename S, int M>
__device__ __forceinline__ void x_store(S* out, const XBlock<S, M>& b, const typename Store<S>::T (&v)[M])
{
  using T = typename Store<S>::T;
  constexpr int P = XPitch<M>::P;
  wave_lds_sync();
  T* own = b.tile + b.lane * P;
#pragma unroll
  for (int k = 0; k < M; ++k) own[k] = v[k];
  wave_lds_sync();
  if constexpr (is_bfp<S>) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = b.lane + 64 * k;
      sst(out, b.base + (uint32_t)e * (uint32_t)sizeof(S), b.tile[(e / M) * P + e % M], e < b.valid);
    }
  } else if constexpr (M % 2 == 0) {
#pragma unroll
    for (int k = 0; k < M / 2; ++k) {
      const int e = 2 * b.lane + 128 * k;
      const T* src = b.tile + (e / M) * P + e % M;
      if (e < b.valid) st2(out, b.base + (uint32_t)e * (uint32_t)sizeof(T), src[0], src[1]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = b.lane + 64 * k;
      if (e < b.valid) st(out, b.base + (uint32_t)e * (uint32_t)sizeof(T), b.tile[(e / M) * P + e % M]);
    }
  }
}

// Right-hand side of KIND for the M chunk points from a window whose halo is
// HLW points on the left (HLW >= Halo<KIND>::L).
template <int KIND, typename T, int M, int HLW, int NW>
__device__ __forceinline__ void rhs_chunk(const T (&win)[NW], const Coef<T>& cf, T (&v)[M])
{
#pragma unroll
  for (int k = 0; k < M; ++k) v[k] = rhs_at<KIND>(win, k + HLW, cf);
}

// LDS words (of T) the solvers use per work-group.
__host__ __device__ inline int line_lds_words(int Wc, int lpw) { return Wc > 1 ? 4 * Wc * lpw + 2 * lpw : 0; }

// beta^(M c) from the binary powers.
template <typename T>
__device__ __forceinline__ T pow_Mc(const Circ<T>& cc, int c)
{
  T r = T(1);
#pragma unroll
  for (int k = 0; k < 7; ++k) r = ((c >> k) & 1) ? r * cc.pwM2[k] : r;
  return r;
}

// ---- circulant solver ------------------------------------------------------
template <typename T, int M, bool LF>
__device__ __forceinline__ void solve_circ(T (&v)[M], const LineThread& t, const Circ<T>& cc, T* lds)
{
  const int G = t.G, g = t.g;
  T* agf = lds;                          // [Wc][LPW] forward wave aggregates
  T* agb = lds + t.Wc * t.LPW;           // [Wc][LPW] backward wave aggregates
  // forward y_j = q_j + beta y_{j-1}
  T prev = T(0);
#pragma unroll
  for (int k = 0; k < M; ++k) { prev = v[k] + cc.beta * prev; v[k] = prev; }
  T E = prev;
  {
    int s = 0;
    for (int d = 1; d < G; d <<= 1, ++s) {
      const T o = __shfl(E, t.lane_of(g >= d ? g - d : g));
      if (g >= d) E = E + cc.pwM2[s] * o;
    }
  }
  T Eex = __shfl(E, t.lane_of(g > 0 ? g - 1 : 0));
  if (g == 0) Eex = T(0);
  T cw = T(0), total;
  if (LF && t.Wc > 1) {
    if (g == G - 1) agf[t.wc * t.LPW + t.ls] = E;
    __syncthreads();
    T acc = T(0);
    for (int w = 0; w < t.Wc; ++w) {
      if (w == t.wc) cw = acc;
      acc = cc.pwMG * acc + agf[w * t.LPW + t.ls];
    }
    total = acc;
  } else {
    total = __shfl(E, t.lane_of(G - 1));
  }
  const T ylast = total * cc.close;
  // beta^(M c): x-lines take it from the table in the kernel arguments (one
  // load rather than seven selected multiplies); y/z-lines form it, as the
  // loads would hold registers their work-groups of up to 16 wavefronts lack
  const T carry = (LF ? pow_Mc(cc, t.c) : cc.pMc[t.c]) * ylast + (Eex + (LF ? pow_Mc(cc, g) : cc.pMc[g]) * cw);
#pragma unroll
  for (int k = 0; k < M; ++k) v[k] = v[k] + cc.pw[k + 1] * carry;

  // backward x_j = y_j + beta x_{j+1}
  T nxt = T(0);
#pragma unroll
  for (int k = M - 1; k >= 0; --k) { nxt = v[k] + cc.beta * nxt; v[k] = nxt; }
  T F = nxt;
  {
    int s = 0;
    for (int d = 1; d < G; d <<= 1, ++s) {
      const T o = __shfl(F, t.lane_of(g + d < G ? g + d : g));
      if (g + d < G) F = F + cc.pwM2[s] * o;
    }
  }
  T Fex = __shfl(F, t.lane_of(g < G - 1 ? g + 1 : g));
  if (g == G - 1) Fex = T(0);
  T cw2 = T(0), total2;
  if (LF && t.Wc > 1) {
    if (g == 0) agb[t.wc * t.LPW + t.ls] = F;
    __syncthreads();
    T acc = T(0);
    for (int w = t.Wc - 1; w >= 0; --w) {
      if (w == t.wc) cw2 = acc;
      acc = cc.pwMG * acc + agb[w * t.LPW + t.ls];
    }
    total2 = acc;
  } else {
    total2 = __shfl(F, t.lane_of(0));
  }
  const T x0 = total2 * cc.close;
  const T carry2 = (Fex + (LF ? pow_Mc(cc, G - 1 - g) : cc.pMc[G - 1 - g]) * cw2) +
                   (LF ? pow_Mc(cc, t.C - 1 - t.c) : cc.pMc[t.C - 1 - t.c]) * x0;
#pragma unroll
  for (int k = 0; k < M; ++k) v[k] = v[k] + cc.pw[M - k] * carry2;
}

// ---- table solver (the reference's sweeps with the operator's vectors) -------
template <typename T, int M, bool LF>
__device__ __forceinline__ void solve_table(T (&v)[M], const LineThread& t, const OpDev<T>& op, T* lds,
                                            int n, bool fix)
{
  const int G = t.G, g = t.g, Wc = LF ? t.Wc : 1, LPW = t.LPW, ls = t.ls, j0 = t.j0;
  T* sA = lds;
  T* sB = sA + Wc * LPW;
  T* rA = sB + Wc * LPW;
  T* rB = rA + Wc * LPW;
  T* s0 = rB + Wc * LPW;
  T* sN = s0 + LPW;

  // forward t_j = r_j - s_j t_{j-1} (src/thomas.f90:73-75)
  T A = T(1), B = T(0);
  {
    T prev = T(0);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const T s = op.s[j0 + k];
      v[k] = v[k] - prev * s;
      prev = v[k];
      A = A * (-s);
    }
    B = prev;
  }
  for (int d = 1; d < G; d <<= 1) {
    const int src = t.lane_of(g >= d ? g - d : g);
    const T Ao = __shfl(A, src), Bo = __shfl(B, src);
    if (g >= d) { B = A * Bo + B; A = A * Ao; }
  }
  T carry;
  {
    const int src = t.lane_of(g > 0 ? g - 1 : 0);
    const T Ap = __shfl(A, src), Bp = __shfl(B, src);
    const T Ae = g > 0 ? Ap : T(1), Be = g > 0 ? Bp : T(0);
    if (Wc > 1) {
      if (g == G - 1) { sA[t.wc * LPW + ls] = A; sB[t.wc * LPW + ls] = B; }
      __syncthreads();
      T cw = T(0);
      for (int w2 = 0; w2 < t.wc; ++w2) cw = sA[w2 * LPW + ls] * cw + sB[w2 * LPW + ls];
      carry = Ae * cw + Be;
    } else {
      carry = Be;
    }
  }
  {
    T P = T(1);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      P = P * (-op.s[j0 + k]);
      v[k] = v[k] + P * carry;
    }
  }

  // backward x_j = (t_j - f_j x_{j+1}) w_j (src/thomas.f90:76-79)
  A = T(1);
  {
    T nxt = T(0);
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
      const T f = op.f[j0 + k], w = op.w[j0 + k];
      v[k] = (v[k] - f * nxt) * w;
      nxt = v[k];
      A = A * (-(f * w));
    }
    B = nxt;
  }
  for (int d = 1; d < G; d <<= 1) {
    const int src = t.lane_of(g + d < G ? g + d : g);
    const T Ao = __shfl(A, src), Bo = __shfl(B, src);
    if (g + d < G) { B = A * Bo + B; A = A * Ao; }
  }
  {
    const int src = t.lane_of(g < G - 1 ? g + 1 : g);
    const T Ap = __shfl(A, src), Bp = __shfl(B, src);
    const T Ae = g < G - 1 ? Ap : T(1), Be = g < G - 1 ? Bp : T(0);
    if (Wc > 1) {
      if (g == 0) { rA[t.wc * LPW + ls] = A; rB[t.wc * LPW + ls] = B; }
      __syncthreads();
      T cw = T(0);
      for (int w2 = Wc - 1; w2 > t.wc; --w2) cw = rA[w2 * LPW + ls] * cw + rB[w2 * LPW + ls];
      carry = Ae * cw + Be;
    } else {
      carry = Be;
    }
  }
  {
    T Q = T(1);
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
      const T f = op.f[j0 + k], w = op.w[j0 + k];
      Q = Q * (-(f * w));
      v[k] = v[k] + Q * carry;
    }
  }

  // periodic correction (src/thomas.f90:49-55)
  if (fix) {
    T x0, xn;
    if (Wc > 1) {
      if (t.c == 0) s0[ls] = v[0];
      if (t.c == t.C - 1) sN[ls] = v[M - 1];
      __syncthreads();
      x0 = s0[ls];
      xn = sN[ls];
    } else {
      x0 = __shfl(v[0], t.lane_of(0));
      xn = __shfl(v[M - 1], t.lane_of(G - 1));
    }
    const T ss = (x0 - op.alfa * xn) / (T(1) + op.p[0] - op.alfa * op.p[n - 1]);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = v[k] - ss * op.p[j0 + k];
  }
}

// Host: circulant constants for alfa on a line plan.
template <typename T>
Circ<T> make_circ(double alfa, const LinePlan& P);
// Host: stencil coefficients of a kind in a direction; scaled by 1/gamma for
// the circulant solver when gamma_scale is true.
Coef<double> stencil_coef(const Scheme& s, int kind);
double circ_gamma(double alfa);

}  // namespace x3d
