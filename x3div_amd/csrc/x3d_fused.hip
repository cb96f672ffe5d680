// Fused timestep passes: every line solve a pencil direction needs, for all
// fields, in one sweep over the lines, with the reference's elementwise glue
// (products, combinations, Euler update, velocity correction) folded into the
// prologue/epilogue.  At one rank the pencils are one array, so the step is
//   transeq   Z -> Y -> X(+combine, Euler)        src/transeq.f90:44-355
//             (FP32 single grids: X -> Y -> Z(+combine, Euler))
//   divergence X -> Y -> Z                         src/navier.f90:98-236
//   Poisson                                        src/poisson.f90:276-419
//   gradp     Z -> Y -> X(+cor_vel)                src/navier.f90:249-359, 64-91
// Each pass reads its input fields once and writes its outputs once; all
// periodic operators use the circulant line solver.
#include <cstring>
#include <map>

#include "x3d_line.h"

namespace x3d {

// ---- transeq: r_c = D(u_d u_c) + u_d D(u_c), d = DIR ----------------------
// ZF (z first): Z: du_c = r_c; Y: tmp_c = du_c + r_c;
//               X: du_c = tmp_c + r_c; if euler: u_c = gdt du_c + u_c.
// else:         X: du_c = r_c; Y: tmp_c = du_c + r_c;
//               Z: du_c = tmp_c + r_c; if euler: u_c = gdt du_c + u_c.
// (src/transeq.f90:100-115, 120-134, 139-153, 158-177 per component; the
// reference sums the x, y, z parts in that order, ZF the z, y, x parts: the
// same terms, rounded in another order.)  With ZF the z pass, whose strided
// rows cost an address translation each, reads u and writes one part (6
// fields), and the x pass, whose rows are contiguous, carries the sums and
// the Euler update (12 fields).  Measured at 512^3, that is faster for FP64
// (6.81 -> 6.75 ms for the three passes) and BFP32 (6.81 -> 5.86 ms), slower
// for FP32 (3.11 -> 3.21 ms), whose single grids keep x first.
// S: the fields' storage (FP64, FP32, or BFP32 computed in FP64).
template <typename S>
struct TranseqArgs {
  using T = typename Store<S>::T;
  S *ux, *uy, *uz, *dux, *duy, *duz, *tx, *ty, *tz;
  Coef<T> cf;
  Circ<T> cc;
  Geom geo;
  Tiling tl;
  T gdt;
  int euler;
  const T* sy;  // y-stretching ppy(j), or nullptr (y pass only)
};

template <typename S, int M, int DIR, bool ZF>
__device__ __forceinline__ void transeq_body(const TranseqArgs<S>& __restrict__ A, int blk)
{
  // this pass's place in the order: the first writes du_c, the last adds
  // tmp_c and applies the Euler update
  constexpr bool FIRST = DIR == (ZF ? 2 : 0), LAST = DIR == (ZF ? 0 : 2);
  using T = typename Store<S>::T;
  S* const ux = A.ux; S* const uy = A.uy; S* const uz = A.uz;
  S* const dux = A.dux; S* const duy = A.duy; S* const duz = A.duz;
  S* const tx = A.tx; S* const ty = A.ty; S* const tz = A.tz;
  const Coef<T> cf = A.cf;
  const Circ<T>& cc = A.cc;
  const Geom geo = A.geo;
  const T gdt = A.gdt;
  const int euler = A.euler;
  constexpr int H = 2, NW = M + 2 * H;
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<S, DIR>(geo, A.tl, M, blk);
  const int n = geo.n;
  S* const Ud = DIR == 0 ? ux : (DIR == 1 ? uy : uz);
  // x: a second per-wavefront tile holds u_x (= u_d) for the whole kernel,
  // staged once; the first tile stages each u_c
  XBlock<S, M> xb, xd;
  if constexpr (DIR == 0) {
    xb = x_block<S, M>(geo, t, lds);
    xd = x_block<S, M>(geo, t, lds + (blockDim.x >> 6) * x_tile_words<M>());
    x_stage_in<S, M>(Ud, xd);
  }
  // y/z: the u_d window (the same for the three components) and, for the
  // Euler update of a last z pass, each u_c chunk are kept in per-thread LDS
  // slots from their loads to their later uses instead of being loaded again:
  // strided loads cost address translation per row, LDS slots do not.  Slot
  // q of thread i is keep[q * blockDim.x + i], after the solver's words.
  T* const keep = lds + line_lds_words(t.Wc, t.LPW);
  const int nth = blockDim.x, tid = threadIdx.x;
  // a first z pass: the u_y window of the second component is loaded during
  // the first component's line solves (its strided rows wait on address
  // translation) and formed when it is used.  FP64 arithmetic (FP64 and
  // BFP32 fields) with 16-point chunks only, whose 512-thread work-groups
  // hold 2 waves per SIMD with or without the window's registers (shorter
  // chunks would lose occupancy).
  constexpr bool PREW = DIR == 2 && FIRST && sizeof(T) == 8 && M >= 16;
  RawT<S> pw[PREW ? NW : 1];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // components one after another: no loads hoisted across them
    __builtin_amdgcn_sched_barrier(0);
    // pointers straight from the kernel arguments (c is a constant here)
    S* const Uc = c == 0 ? ux : (c == 1 ? uy : uz);
    S* const DUc = c == 0 ? dux : (c == 1 ? duy : duz);
    S* const TMc = c == 0 ? tx : (c == 1 ? ty : tz);
    // right-hand sides of D(u_c) and D(u_d u_c) first, so that only the chunk
    // values stay live through the two line solves
    T v1[M], v2[M];
    {
      T wc[NW], wd[NW];
      if constexpr (DIR == 0) {
        x_window<S, M, H, H>(xd, t, wd);
        if (c == 0) {
#pragma unroll
          for (int q = 0; q < NW; ++q) wc[q] = wd[q];
        } else {
          x_load_window<S, M, H, H>(Uc, xb, t, wc);
        }
      } else {
        if (c == 0) {
          load_window<S, M, H, H, true>(Ud, t, n, wd);
#pragma unroll
          for (int q = 0; q < NW; ++q) keep[q * nth + tid] = wd[q];
        } else {
#pragma unroll
          for (int q = 0; q < NW; ++q) wd[q] = keep[q * nth + tid];
        }
        if (c == DIR) {
#pragma unroll
          for (int q = 0; q < NW; ++q) wc[q] = wd[q];
        } else {
          bool formed = false;
          if constexpr (PREW) {
            if (c == 1) {
              window_finish<S, M, H, H>(t, pw, wc);
              formed = true;
            }
          }
          if (!formed) load_window<S, M, H, H, true>(Uc, t, n, wc);
        }
        if constexpr (DIR == 2 && LAST) {
#pragma unroll
          for (int k = 0; k < M; ++k) keep[(NW + k) * nth + tid] = wc[H + k];
        }
      }
      rhs_chunk<K_D, T, M, H>(wc, cf, v2);
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int q = k + H;
        v1[k] = cf.a * (wd[q + 1] * wc[q + 1] - wd[q - 1] * wc[q - 1]) +
                cf.b * (wd[q + 2] * wc[q + 2] - wd[q - 2] * wc[q - 2]);
      }
    }
    // y, and a last z pass: the epilogue's strided loads (du_c of the first
    // pass, resp. tmp_c) are issued here and complete during the two line
    // solves, whose barriers wait on LDS only.  Neither field is written by
    // this pass.  (not for FP32 z-lines of 32-point chunks, which would spill)
    constexpr bool PF = DIR == 1 || (DIR == 2 && LAST && !(sizeof(T) == 4 && M >= 32));
    if constexpr (PREW) {
      if (c == 0) window_issue<S, M, H, H>(uy, t, n, pw);
    }
    // a last x pass in FP64: the coalesced loads of tmp_c likewise (measured:
    // FP64 512^3 2.45 -> 2.42 ms; the FP32 batch lost a wave and ran slower)
    constexpr bool PFX = DIR == 0 && LAST && M % 2 == 0 && !is_bfp<S> && sizeof(T) == 8;
    // (storage words, decoded at use: a BFP32 decode would wait for them here)
    T a[M];
    RawT<S> ar[M];
    if constexpr (PFX) x_fetch<S, M>(TMc, xb, a);
    if constexpr (PF) {
      if (t.active) {
        const S* src = DIR == 1 ? DUc : TMc;  // (DIR 2: tmp_c)
#pragma unroll
        for (int k = 0; k < M; ++k) ar[k] = ld_raw(src, t.at(t.j0 + k));
      }
    }
    // keep the two line solves apart: interleaving them doubles the live state
    __builtin_amdgcn_sched_barrier(0);
    solve_circ<T, M, DIR != 0>(v2, t, cc, lds);
    __builtin_amdgcn_sched_barrier(0);
    solve_circ<T, M, DIR != 0>(v1, t, cc, lds);
    __builtin_amdgcn_sched_barrier(0);
    // u_d at the chunk points again (a cache hit) rather than kept live
    if constexpr (DIR == 0) {
      const T* ud = xd.tile + xd.lane * XPitch<M>::P;
#pragma unroll
      for (int k = 0; k < M; ++k) v1[k] = v1[k] + ud[k] * v2[k];
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) v1[k] = v1[k] + keep[(H + k) * nth + tid] * v2[k];
    }
    if constexpr (DIR == 0 && FIRST) {
      x_store<S, M>(DUc, xb, v1);
    } else if constexpr (DIR == 0) {
      // du_c = tmp_c + r_c and the Euler update, through the staging tile:
      // u_c at the chunk points first (the u_x tile, resp. the tile still
      // holding this component's window load), then tmp_c restaged (its
      // loads were issued before the solves where PFX)
      T uc[M];
      {
        const T* own = (c == 0 ? xd.tile + xd.lane * XPitch<M>::P : xb.tile + xb.lane * XPitch<M>::P);
#pragma unroll
        for (int k = 0; k < M; ++k) uc[k] = own[k];
      }
      if constexpr (PFX) {
        x_put<S, M>(xb, a);
        const T* own = xb.tile + xb.lane * XPitch<M>::P;
#pragma unroll
        for (int k = 0; k < M; ++k) a[k] = own[k];
      } else {
        x_load_chunk<S, M>(TMc, xb, a);
      }
#pragma unroll
      for (int k = 0; k < M; ++k) v1[k] = a[k] + v1[k];
      x_store<S, M>(DUc, xb, v1);
      if (euler) {
#pragma unroll
        for (int k = 0; k < M; ++k) v1[k] = gdt * v1[k] + uc[k];
        x_store<S, M>(Uc, xb, v1);
      }
    } else if constexpr (DIR == 1) {
      // tmp_c = du_c (first pass's part) + r_c: the sum is formed here, where
      // a strided row stays inside one page, rather than in a last z pass
      if (A.sy) {
#pragma unroll
        for (int k = 0; k < M; ++k) v1[k] = v1[k] * A.sy[t.j0 + k];
      }
      if (t.active) {
#pragma unroll
        for (int k = 0; k < M; ++k) sst(TMc, t.at(t.j0 + k), from_raw<S>(ar[k]) + v1[k]);
      }
    } else if constexpr (FIRST) {
      if (t.active) {
#pragma unroll
        for (int k = 0; k < M; ++k) sst(DUc, t.at(t.j0 + k), v1[k]);
      }
    } else {
      if (t.active) {
        if constexpr (PF) {
#pragma unroll
          for (int k = 0; k < M; ++k) a[k] = from_raw<S>(ar[k]);
        } else {
#pragma unroll
          for (int k = 0; k < M; ++k) a[k] = sld(TMc, t.at(t.j0 + k));
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const uint32_t at = t.at(t.j0 + k);
          const T d = a[k] + v1[k];
          sst(DUc, at, d);
          if (euler) sst(Uc, at, gdt * d + keep[(NW + k) * nth + tid]);
        }
      }
    }
  }
}

// ---- divergence passes (src/navier.f90:166-202) ----------------------------
// X: o1 = Dvp(a), o2 = Ivp(b), o3 = Ivp(c)        a, b, c = ux, uy, uz
// Y: o1 = Dvp(b) + Ivp(a), o2 = Ivp(c)             a, b, c = pp1, pgy1, pgz1
// Z: o1 = Ivp(a) + Dvp(b)                          a, b = duydypi2, upi2
// Arguments of the divergence and gradient passes.
template <typename S>
struct PassArgs {
  using T = typename Store<S>::T;
  const S *a, *b, *c;
  S *o1, *o2, *o3;
  S *ux, *uy, *uz;
  Coef<T> cfd;
  Circ<T> ccd;
  Coef<T> cfi;
  Circ<T> cci;
  Geom geo;
  Tiling tl;
  int corr;
  const T* sy;  // y-stretching: ppyi(j) for Dvp (divergence), ppy(j) for Dpv (gradp), or nullptr
};

template <typename S, int M, int DIR>
__device__ __forceinline__ void div_body(const PassArgs<S>& __restrict__ A, int blk)
{
  using T = typename Store<S>::T;
  const S* const a = A.a; const S* const b = A.b; const S* const c = A.c;
  S* const o1 = A.o1; S* const o2 = A.o2; S* const o3 = A.o3;
  const Coef<T> cfd = A.cfd, cfi = A.cfi;
  const Circ<T>& ccd = A.ccd;
  const Circ<T>& cci = A.cci;
  const Geom geo = A.geo;
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<S, DIR>(geo, A.tl, M, blk);
  const int n = geo.n;
  if constexpr (DIR == 0) {
    const XBlock<S, M> xb = x_block<S, M>(geo, t, lds);
    T v[M];
    {
      T w[M + 3];
      x_load_window<S, M, 1, 2>(a, xb, t, w);
      rhs_chunk<K_DVP, T, M, 1>(w, cfd, v);
    }
    solve_circ<T, M, false>(v, t, ccd, lds);
    x_store<S, M>(o1, xb, v);
    {
      T w[M + 7];
      x_load_window<S, M, 3, 4>(b, xb, t, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, false>(v, t, cci, lds);
    x_store<S, M>(o2, xb, v);
    {
      T w[M + 7];
      x_load_window<S, M, 3, 4>(c, xb, t, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, false>(v, t, cci, lds);
    x_store<S, M>(o3, xb, v);
  } else if constexpr (DIR == 1) {
    T v[M], s[M];
    {
      T w[M + 3];
      load_window<S, M, 1, 2, DIR != 0>(b, t, n, w);
      rhs_chunk<K_DVP, T, M, 1>(w, cfd, v);
    }
    solve_circ<T, M, DIR != 0>(v, t, ccd, lds);
    if (A.sy) {
#pragma unroll
      for (int k = 0; k < M; ++k) v[k] = v[k] * A.sy[t.j0 + k];
    }
    {
      T w[M + 7];
      load_window<S, M, 3, 4, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, s);
    }
    solve_circ<T, M, DIR != 0>(s, t, cci, lds);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = v[k] + s[k];
    store_chunk<S, M>(o1, t, v);
    {
      T w[M + 7];
      load_window<S, M, 3, 4, DIR != 0>(c, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, DIR != 0>(v, t, cci, lds);
    store_chunk<S, M>(o2, t, v);
  } else {
    T v[M], s[M];
    {
      T w[M + 7];
      load_window<S, M, 3, 4, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    // the second window's loads in flight during the first solve; in FP64
    // arithmetic it is formed (halo shuffles, BFP32 decode) after the solve,
    // in FP32 before it (FP32's 32-point chunks would lose a wave per SIMD)
    T wb[M + 3];
    RawT<S> wbr[sizeof(T) == 8 ? M + 3 : 1];
    if constexpr (sizeof(T) == 8) window_issue<S, M, 1, 2>(b, t, n, wbr);
    else load_window<S, M, 1, 2, DIR != 0>(b, t, n, wb);
    __builtin_amdgcn_sched_barrier(0);
    solve_circ<T, M, DIR != 0>(v, t, cci, lds);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (sizeof(T) == 8) window_finish<S, M, 1, 2>(t, wbr, wb);
    rhs_chunk<K_DVP, T, M, 1>(wb, cfd, s);
    solve_circ<T, M, DIR != 0>(s, t, ccd, lds);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = v[k] + s[k];
    store_chunk<S, M>(o1, t, v);
  }
}

// ---- pressure-gradient passes (src/navier.f90:273-304) + cor_vel -----------
// Z: o1 = Ipv(a), o2 = Dpv(a)                       a = pp3
// Y: o1 = Ipv(a), o2 = Dpv(a), o3 = Ipv(b)          a, b = ppi3, pgz3
// X: o1 = Dpv(a), o2 = Ipv(b), o3 = Ipv(c)          a, b, c = ppi2, pgy2, pgzi2
//    if corr: u_k = u_k - o_k (src/navier.f90:84-86)
template <typename S, int M, int DIR>
__device__ __forceinline__ void grad_body(const PassArgs<S>& __restrict__ A, int blk)
{
  using T = typename Store<S>::T;
  const S* const a = A.a; const S* const b = A.b; const S* const c = A.c;
  S* const o1 = A.o1; S* const o2 = A.o2; S* const o3 = A.o3;
  S* const ux = A.ux; S* const uy = A.uy; S* const uz = A.uz;
  const Coef<T> cfd = A.cfd, cfi = A.cfi;
  const Circ<T>& ccd = A.ccd;
  const Circ<T>& cci = A.cci;
  const Geom geo = A.geo;
  const int corr = A.corr;
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<S, DIR>(geo, A.tl, M, blk);
  const int n = geo.n;
  if constexpr (DIR == 2 || DIR == 1) {
    T v[M];
    {
      T w[M + 7];
      load_window<S, M, 4, 3, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      T s[M];
      rhs_chunk<K_DPV, T, M, 4>(w, cfd, s);
      solve_circ<T, M, DIR != 0>(v, t, cci, lds);
      store_chunk<S, M>(o1, t, v);
      solve_circ<T, M, DIR != 0>(s, t, ccd, lds);
      if (DIR == 1 && A.sy) {
#pragma unroll
        for (int k = 0; k < M; ++k) s[k] = s[k] * A.sy[t.j0 + k];
      }
      store_chunk<S, M>(o2, t, s);
    }
    if constexpr (DIR == 1) {
      T w[M + 7];
      load_window<S, M, 4, 3, DIR != 0>(b, t, n, w);
      rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      solve_circ<T, M, DIR != 0>(v, t, cci, lds);
      store_chunk<S, M>(o3, t, v);
    }
  } else {
    const XBlock<S, M> xb = x_block<S, M>(geo, t, lds);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const S* const inq = q == 0 ? a : (q == 1 ? b : c);
      S* const outq = q == 0 ? o1 : (q == 1 ? o2 : o3);
      S* const Uq = q == 0 ? ux : (q == 1 ? uy : uz);
      T v[M];
      {
        T w[M + 7];
        x_load_window<S, M, 4, 3>(inq, xb, t, w);
        if (q == 0) rhs_chunk<K_DPV, T, M, 4>(w, cfd, v);
        else rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      }
      if (q == 0) solve_circ<T, M, false>(v, t, ccd, lds);
      else solve_circ<T, M, false>(v, t, cci, lds);
      x_store<S, M>(outq, xb, v);
      if (corr) {
        T u[M];
        x_load_chunk<S, M>(Uq, xb, u);
#pragma unroll
        for (int k = 0; k < M; ++k) u[k] = u[k] - v[k];
        x_store<S, M>(Uq, xb, u);
      }
    }
  }
}

// ---- kernels: one grid (blockIdx.x is the line block) or many grids in one
// persistent launch (each work-group walks virtual blocks; block0[g] is the
// first virtual block of grid g, block0[ngrids] the total) -----------------
__device__ __forceinline__ int grid_of(const int* __restrict__ block0, int ngrids, int vb)
{
  int lo = 0, hi = ngrids - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block0[mid] <= vb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename S, int M, int DIR, bool ZF>
__global__ __launch_bounds__(512) void k_transeq(TranseqArgs<S> A) { transeq_body<S, M, DIR, ZF>(A, A.tl.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x); }
// y passes: at most 128 VGPRs, so that two work-groups share a compute unit
// (measured faster at 512^3 in both precisions despite a few spills)
template <typename S, int M, int DIR>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(DIR == 1 ? 4 : 1))) void k_div(PassArgs<S> A) { div_body<S, M, DIR>(A, A.tl.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x); }
template <typename S, int M, int DIR>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(DIR == 1 ? 4 : 1))) void k_grad(PassArgs<S> A) { grad_body<S, M, DIR>(A, A.tl.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x); }

template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_transeq_batch(const TranseqArgs<T>* __restrict__ list,
                                                       const int* __restrict__ block0, int ngrids)
{
  const int total = block0[ngrids];
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int g = grid_of(block0, ngrids, vb);
    transeq_body<T, M, DIR, true>(list[g], vb - block0[g]);
  }
}
template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_div_batch(const PassArgs<T>* __restrict__ list,
                                                   const int* __restrict__ block0, int ngrids)
{
  const int total = block0[ngrids];
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int g = grid_of(block0, ngrids, vb);
    div_body<T, M, DIR>(list[g], vb - block0[g]);
  }
}
template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_grad_batch(const PassArgs<T>* __restrict__ list,
                                                    const int* __restrict__ block0, int ngrids)
{
  const int total = block0[ngrids];
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int g = grid_of(block0, ngrids, vb);
    grad_body<T, M, DIR>(list[g], vb - block0[g]);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Work-groups above 64 KiB of dynamic LDS (up to the CU's 160 KiB) need the
// kernel's limit raised once.
static void allow_lds(const void* fn, size_t bytes)
{
  if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename S>
struct DirSet {
  using T = typename Store<S>::T;
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  size_t smem_tq;  // transeq: plus the y/z keep slots
  int M;
  Coef<T> cf_d, cf_dvp, cf_dpv, cf_ivp, cf_ipv;  // scaled by 1/gamma
  Circ<T> cc_d, cc_s, cc_i;
};

// Plans: FP64 and BFP32 use the fused plans, FP32 its own (longer chunks,
// 128-byte rows).
template <typename S>
static const LinePlan& fused_plan(const x3d_ctx* c, int dir)
{
  return std::is_same<S, float>::value ? c->fplan32[dir] : c->fplan[dir];
}

template <typename S>
static DirSet<S> dirset(const x3d_ctx* c, int dir)
{
  using T = typename Store<S>::T;
  DirSet<S> s;
  const LinePlan& P = fused_plan<S>(c, dir);
  launch_dims_plan(c, dir, P, &s.geo, &s.tl, &s.grid, &s.block, &s.smem, sizeof(T));
  // BFP32 z-rows of 16 lines are 64 B, half a 128 B L2 line whose other half
  // belongs to the next line block: blocks regrouped per XCD fetch it into
  // one L2 once (the dispatcher's round robin puts the two on different XCDs)
  if constexpr (is_bfp<S>)
    if (dir == 2) s.tl.xcd = c->xcd_remap;
  s.M = P.M;
  // keep slots of transeq: the u_d window (M + 4)
  s.smem_tq = s.smem;
  if (P.line_fast)
    s.smem_tq += sizeof(T) * (size_t)s.block.x * (size_t)(P.M + 4);
  else
    s.smem_tq = 2 * s.smem;  // x: the staging tile and the u_x tile
  const Scheme& sc = c->sc[dir];
  // coefficients in FP64, rounded once to T
  auto scaled = [](Coef<double> k, double alfa) {
    const double ig = 1.0 / circ_gamma(alfa);
    return Coef<T>{(T)(k.a * ig), (T)(k.b * ig), (T)(k.c * ig), (T)(k.d * ig)};
  };
  s.cf_d = scaled(stencil_coef(sc, K_D), sc.alfai);
  s.cf_dvp = scaled(stencil_coef(sc, K_DVP), sc.alcai);
  s.cf_dpv = scaled(stencil_coef(sc, K_DPV), sc.alcai);
  s.cf_ivp = scaled(stencil_coef(sc, K_IVP), sc.ailcai);
  s.cf_ipv = scaled(stencil_coef(sc, K_IPV), sc.ailcai);
  s.cc_d = make_circ<T>(sc.alfai, P);
  s.cc_s = make_circ<T>(sc.alcai, P);
  s.cc_i = make_circ<T>(sc.ailcai, P);
  return s;
}

// Chunk lengths instantiated per storage: FP64 and BFP32 up to 16 points, FP32
// up to 32 (the same registers per thread).
#define X3D_MSWITCH_T(S, M, CALL_)                                                          \
  do {                                                                                      \
    if constexpr (!std::is_same<S, float>::value) {                                         \
      switch (M) {                                                                          \
        case 1: { constexpr int MM = 1; CALL_; } break;                          \
        case 2: { constexpr int MM = 2; CALL_; } break;                          \
        case 3: { constexpr int MM = 3; CALL_; } break;                          \
        case 4: { constexpr int MM = 4; CALL_; } break;                          \
        case 5: { constexpr int MM = 5; CALL_; } break;                          \
        case 6: { constexpr int MM = 6; CALL_; } break;                          \
        case 8: { constexpr int MM = 8; CALL_; } break;                          \
        case 12: { constexpr int MM = 12; CALL_; } break;                          \
        case 16: { constexpr int MM = 16; CALL_; } break;                          \
        default: return X3D_ERR_UNSUPPORTED;                                                \
      }                                                                                     \
    } else {                                                                                \
      switch (M) {                                                                          \
        case 1: { constexpr int MM = 1; CALL_; } break;                          \
        case 2: { constexpr int MM = 2; CALL_; } break;                          \
        case 3: { constexpr int MM = 3; CALL_; } break;                          \
        case 4: { constexpr int MM = 4; CALL_; } break;                          \
        case 5: { constexpr int MM = 5; CALL_; } break;                          \
        case 6: { constexpr int MM = 6; CALL_; } break;                          \
        case 8: { constexpr int MM = 8; CALL_; } break;                          \
        case 12: { constexpr int MM = 12; CALL_; } break;                          \
        case 16: { constexpr int MM = 16; CALL_; } break;                          \
        case 24: { constexpr int MM = 24; CALL_; } break;                          \
        case 32: { constexpr int MM = 32; CALL_; } break;                          \
        default: return X3D_ERR_UNSUPPORTED;                                                \
      }                                                                                     \
    }                                                                                       \
  } while (0)

// BFP32 fields: a y/z-line lane group must be 16 lines (one x-segment) or a
// multiple, and x rows whole segments.
template <typename S>
static bool storage_ok(const x3d_ctx* c, int dir)
{
  if constexpr (!is_bfp<S>) return true;
  const LinePlan& P = fused_plan<S>(c, dir);
  return c->nx % 16 == 0 && (!P.line_fast || P.LW % 16 == 0);
}

template <int DIR, typename S>
static int transeq_pass(x3d_ctx* c, S* ux, S* uy, S* uz, S* dux, S* duy, S* duz, S* tx, S* ty, S* tz, double gdt,
                        int euler)
{
  using T = typename Store<S>::T;
  if (!storage_ok<S>(c, DIR)) return X3D_ERR_UNSUPPORTED;
  const DirSet<S> s = dirset<S>(c, DIR);
  const TranseqArgs<S> A{ux,   uy,   uz,   dux,  duy,      duz,   tx,
                         ty,   tz,   s.cf_d, s.cc_d, s.geo, s.tl, (T)gdt,
                         euler, DIR == 1 ? stretch_vec<T>(c, 0) : nullptr};
  // (a last z pass keeps each u_c chunk for the Euler update as well)
  constexpr bool ZF = transeq_zfirst<S>();
  const size_t smem = s.smem_tq + (DIR == 2 && !ZF ? sizeof(T) * (size_t)s.block.x * (size_t)s.M : 0);
  X3D_MSWITCH_T(S, s.M, {
    allow_lds((const void*)k_transeq<S, MM, DIR, ZF>, smem);
    hipLaunchKernelGGL((k_transeq<S, MM, DIR, ZF>), s.grid, s.block, smem, c->stream, A);
  });
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <int DIR, typename S>
static int div_pass(x3d_ctx* c, const S* a, const S* b, const S* cc, S* o1, S* o2, S* o3)
{
  using T = typename Store<S>::T;
  if (!storage_ok<S>(c, DIR)) return X3D_ERR_UNSUPPORTED;
  const DirSet<S> s = dirset<S>(c, DIR);
  const PassArgs<S> A{a, b, cc, o1, o2, o3, nullptr, nullptr, nullptr, s.cf_dvp, s.cc_s, s.cf_ivp, s.cc_i,
                      s.geo, s.tl, 0, DIR == 1 ? stretch_vec<T>(c, 1) : nullptr};
  X3D_MSWITCH_T(S, s.M, hipLaunchKernelGGL((k_div<S, MM, DIR>), s.grid, s.block, s.smem, c->stream, A));
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <int DIR, typename S>
static int grad_pass(x3d_ctx* c, const S* a, const S* b, const S* cc, S* o1, S* o2, S* o3, S* ux, S* uy, S* uz,
                     int corr)
{
  using T = typename Store<S>::T;
  if (!storage_ok<S>(c, DIR)) return X3D_ERR_UNSUPPORTED;
  const DirSet<S> s = dirset<S>(c, DIR);
  const PassArgs<S> A{a,    b,    cc,       o1,     o2,       o3,     ux,    uy,   uz,
                      s.cf_dpv, s.cc_s, s.cf_ipv, s.cc_i, s.geo, s.tl, corr, DIR == 1 ? stretch_vec<T>(c, 0) : nullptr};
  X3D_MSWITCH_T(S, s.M, hipLaunchKernelGGL((k_grad<S, MM, DIR>), s.grid, s.block, s.smem, c->stream, A));
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

bool fused_ok(const x3d_ctx* c)
{
  for (int o = 0; o < X3D_NOPS; ++o)
    if (!c->op[o].circulant) return false;
  return true;
}

#define X3D_TRY(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

// Scratch fields: the FP64 work arrays, viewed as S (FP32 and BFP32 use half
// of each).
template <typename S>
static S* wk(x3d_ctx* c, int i) { return reinterpret_cast<S*>(c->work[i]); }

template <typename T>
int fused_momentum(x3d_ctx* c, double gdt, int euler, T* ux, T* uy, T* uz, T* dux, T* duy, T* duz)
{
  if (sizeof(T) == 8 && !c->f64_lines) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(ensure_work(c));
  T *tx = wk<T>(c, 5), *ty = wk<T>(c, 6), *tz = wk<T>(c, 7);
  constexpr int D0 = transeq_zfirst<T>() ? 2 : 0;  // first pass; the last is 2 - D0
  X3D_TRY((transeq_pass<D0, T>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, 0)));
  X3D_TRY(mark_pass(c, 1));
  X3D_TRY((transeq_pass<1, T>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, 0)));
  X3D_TRY(mark_pass(c, 2));
  X3D_TRY((transeq_pass<2 - D0, T>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, euler)));
  X3D_TRY(mark_pass(c, 3));
  return X3D_OK;
}

template <typename T>
int fused_divergence(x3d_ctx* c, T* pp3, const T* ux, const T* uy, const T* uz)
{
  if (sizeof(T) == 8 && !c->f64_lines) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(ensure_work(c));
  T *pp1 = wk<T>(c, 0), *pgy1 = wk<T>(c, 1), *pgz1 = wk<T>(c, 2), *dd2 = wk<T>(c, 3), *upi2 = wk<T>(c, 4);
  X3D_TRY((div_pass<0, T>(c, ux, uy, uz, pp1, pgy1, pgz1)));
  X3D_TRY(mark_pass(c, 4));
  X3D_TRY((div_pass<1, T>(c, pp1, pgy1, pgz1, dd2, upi2, nullptr)));
  X3D_TRY(mark_pass(c, 5));
  X3D_TRY((div_pass<2, T>(c, dd2, upi2, nullptr, pp3, nullptr, nullptr)));
  X3D_TRY(mark_pass(c, 6));
  return X3D_OK;
}

template <typename T>
int fused_gradp(x3d_ctx* c, T* px, T* py, T* pz, const T* pp3, T* ux, T* uy, T* uz, int corr)
{
  if (sizeof(T) == 8 && !c->f64_lines) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(ensure_work(c));
  T *ppi3 = wk<T>(c, 0), *pgz3 = wk<T>(c, 1), *ppi2 = wk<T>(c, 2), *pgy2 = wk<T>(c, 3), *pgzi2 = wk<T>(c, 4);
  X3D_TRY((grad_pass<2, T>(c, pp3, nullptr, nullptr, ppi3, pgz3, nullptr, nullptr, nullptr, nullptr, 0)));
  X3D_TRY(mark_pass(c, 8));
  X3D_TRY((grad_pass<1, T>(c, ppi3, pgz3, nullptr, ppi2, pgy2, pgzi2, nullptr, nullptr, nullptr, 0)));
  X3D_TRY(mark_pass(c, 9));
  X3D_TRY((grad_pass<0, T>(c, ppi2, pgy2, pgzi2, px, py, pz, ux, uy, uz, corr)));
  X3D_TRY(mark_pass(c, 10));
  return X3D_OK;
}

// ---- batched passes ----------------------------------------------------------
namespace {
struct PlanKey {
  int M, G, Wc, Wl, LW;
  bool operator<(const PlanKey& o) const
  {
    return M != o.M ? M < o.M : G != o.G ? G < o.G : Wc != o.Wc ? Wc < o.Wc : Wl != o.Wl ? Wl < o.Wl : LW < o.LW;
  }
};
constexpr int kPassDir[9] = {2, 1, 0, 0, 1, 2, 2, 1, 0};
}  // namespace

template <typename T>
int batch_build(x3d_batch* b, double gdt)
{
  // the batched passes carry no stretching vectors and run the circulant
  // solvers of the library's own tables
  for (const x3d_ctx* g : b->g)
    if (g->istret || !fused_ok(g)) return X3D_ERR_UNSUPPORTED;
  std::vector<unsigned char> host;
  auto put = [&host](const void* p, size_t n) {
    size_t at = (host.size() + 255) & ~size_t(255);
    host.resize(at + n);
    std::memcpy(host.data() + at, p, n);
    return at;
  };
  b->launches.clear();
  for (int i = 0; i < b->ng; ++i) X3D_TRY(ensure_work(b->g[i]));
  for (int pass = 0; pass < 9; ++pass) {
    const int dir = kPassDir[pass];
    std::map<PlanKey, std::vector<int>> groups;
    for (int i = 0; i < b->ng; ++i) {
      const LinePlan& P = fused_plan<T>(b->g[i], dir);
      if (!P.M) return X3D_ERR_UNSUPPORTED;
      groups[PlanKey{P.M, P.G, P.Wc, P.Wl, P.LW}].push_back(i);
    }
    for (auto& kv : groups) {
      std::vector<int> block0{0};
      x3d_batch::Launch L{};
      L.pass = pass;
      L.M = kv.first.M;
      std::vector<TranseqArgs<T>> ta;
      std::vector<PassArgs<T>> pa;
      for (int i : kv.second) {
        x3d_ctx* c = b->g[i];
        const DirSet<T> s = dirset<T>(c, dir);
        T* f[10];
        for (int k = 0; k < 10; ++k) f[k] = static_cast<T*>(b->field(k, i));
        T* w[8];
        for (int k = 0; k < 8; ++k) w[k] = wk<T>(c, k);
        if (pass < 3) {
          ta.push_back(TranseqArgs<T>{f[0], f[1], f[2], f[3], f[4], f[5], w[5], w[6], w[7], s.cf_d, s.cc_d, s.geo,
                                      s.tl, (T)gdt, pass == 2 ? 1 : 0});
        } else if (pass < 6) {
          const T *a = nullptr, *bb = nullptr, *cc = nullptr;
          T *o1 = nullptr, *o2 = nullptr, *o3 = nullptr;
          if (pass == 3) { a = f[0]; bb = f[1]; cc = f[2]; o1 = w[0]; o2 = w[1]; o3 = w[2]; }
          if (pass == 4) { a = w[0]; bb = w[1]; cc = w[2]; o1 = w[3]; o2 = w[4]; }
          if (pass == 5) { a = w[3]; bb = w[4]; o1 = f[6]; }
          pa.push_back(PassArgs<T>{a, bb, cc, o1, o2, o3, nullptr, nullptr, nullptr, s.cf_dvp, s.cc_s, s.cf_ivp,
                                   s.cc_i, s.geo, s.tl, 0});
        } else {
          const T *a = nullptr, *bb = nullptr, *cc = nullptr;
          T *o1 = nullptr, *o2 = nullptr, *o3 = nullptr, *u1 = nullptr, *u2 = nullptr, *u3 = nullptr;
          if (pass == 6) { a = f[6]; o1 = w[0]; o2 = w[1]; }
          if (pass == 7) { a = w[0]; bb = w[1]; o1 = w[2]; o2 = w[3]; o3 = w[4]; }
          if (pass == 8) { a = w[2]; bb = w[3]; cc = w[4]; o1 = f[7]; o2 = f[8]; o3 = f[9]; u1 = f[0]; u2 = f[1]; u3 = f[2]; }
          pa.push_back(PassArgs<T>{a, bb, cc, o1, o2, o3, u1, u2, u3, s.cf_dpv, s.cc_s, s.cf_ipv, s.cc_i, s.geo,
                                   s.tl, pass == 8 ? 1 : 0});
        }
        block0.push_back(block0.back() + (int)s.grid.x);
        L.block = s.block;
        L.smem = pass < 3 ? s.smem_tq : s.smem;
      }
      L.ngroup = (int)kv.second.size();
      L.args = pass < 3 ? put(ta.data(), ta.size() * sizeof(TranseqArgs<T>))
                        : put(pa.data(), pa.size() * sizeof(PassArgs<T>));
      L.block0 = put(block0.data(), block0.size() * sizeof(int));
      // persistent: at most a few work-groups per compute unit walk the blocks
      const int cap = 4 * b->ncu;
      L.grid = dim3((unsigned)(block0.back() < cap ? block0.back() : cap));
      b->launches.push_back(L);
    }
  }
  if (host.size() > b->dargs_bytes) {
    if (b->dargs) (void)hipFree(b->dargs);
    b->dargs = nullptr;
    b->dargs_bytes = 0;
    X3D_CHECK_HIP(hipMalloc(&b->dargs, host.size()));
    b->dargs_bytes = host.size();
  }
  X3D_CHECK_HIP(hipMemcpy(b->dargs, host.data(), host.size(), hipMemcpyHostToDevice));
  b->built = true;
  b->built_gdt = gdt;
  return X3D_OK;
}

template <typename T>
static int batch_launch(x3d_batch* b, const x3d_batch::Launch& L, hipStream_t st)
{
  const char* base = static_cast<const char*>(b->dargs);
  const int* block0 = reinterpret_cast<const int*>(base + L.block0);
  const void* args = base + L.args;
  const TranseqArgs<T>* ta = static_cast<const TranseqArgs<T>*>(args);
  const PassArgs<T>* pa = static_cast<const PassArgs<T>*>(args);
  switch (L.pass) {
#define X3D_BT(P, K, DIR, LIST)                                                                             \
  case P:                                                                                                 \
    X3D_MSWITCH_T(T, L.M, {                                                                               \
      allow_lds((const void*)K<T, MM, DIR>, L.smem);                                                      \
      hipLaunchKernelGGL((K<T, MM, DIR>), L.grid, L.block, L.smem, st, LIST, block0, L.ngroup);           \
    });                                                                                                   \
    break;
    X3D_BT(0, k_transeq_batch, 2, ta) X3D_BT(1, k_transeq_batch, 1, ta) X3D_BT(2, k_transeq_batch, 0, ta)
    X3D_BT(3, k_div_batch, 0, pa) X3D_BT(4, k_div_batch, 1, pa) X3D_BT(5, k_div_batch, 2, pa)
    X3D_BT(6, k_grad_batch, 2, pa) X3D_BT(7, k_grad_batch, 1, pa) X3D_BT(8, k_grad_batch, 0, pa)
#undef X3D_BT
    default: return X3D_ERR_ARG;
  }
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename T>
int batch_passes(x3d_batch* b, int which)
{
  if (!b->built) return X3D_ERR_ARG;
  const int ns = b->overlap ? (int)b->pstreams.size() : 0;
  for (int p = 3 * which; p < 3 * which + 3; ++p) {
    std::vector<const x3d_batch::Launch*> ls;
    for (const auto& L : b->launches)
      if (L.pass == p) ls.push_back(&L);
    if (ns == 0 || ls.size() < 2) {
      for (const auto* L : ls) X3D_TRY(batch_launch<T>(b, *L, b->stream));
      continue;
    }
    // the pass's groups (disjoint grids) side by side; the next pass waits
    // for all of them
    X3D_TRY(batch_fork(b));
    int rc = X3D_OK;
    for (size_t k = 0; !rc && k < ls.size(); ++k) rc = batch_launch<T>(b, *ls[k], b->pstreams[k % ns]);
    X3D_TRY(batch_join(b, rc));
  }
  return X3D_OK;
}

template int batch_build<double>(x3d_batch*, double);
template int batch_build<float>(x3d_batch*, double);
template int batch_passes<double>(x3d_batch*, int);
template int batch_passes<float>(x3d_batch*, int);

template int fused_momentum<double>(x3d_ctx*, double, int, double*, double*, double*, double*, double*, double*);
template int fused_momentum<float>(x3d_ctx*, double, int, float*, float*, float*, float*, float*, float*);
template int fused_momentum<bfp32>(x3d_ctx*, double, int, bfp32*, bfp32*, bfp32*, bfp32*, bfp32*, bfp32*);
template int fused_divergence<double>(x3d_ctx*, double*, const double*, const double*, const double*);
template int fused_divergence<float>(x3d_ctx*, float*, const float*, const float*, const float*);
template int fused_divergence<bfp32>(x3d_ctx*, bfp32*, const bfp32*, const bfp32*, const bfp32*);
template int fused_gradp<double>(x3d_ctx*, double*, double*, double*, const double*, double*, double*, double*,
                                 int);
template int fused_gradp<float>(x3d_ctx*, float*, float*, float*, const float*, float*, float*, float*, int);
template int fused_gradp<bfp32>(x3d_ctx*, bfp32*, bfp32*, bfp32*, const bfp32*, bfp32*, bfp32*, bfp32*, int);

#undef X3D_MSWITCH_T

}  // namespace x3d
