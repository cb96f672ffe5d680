// Fused timestep passes: every line solve a pencil direction needs, for all
// fields, in one sweep over the lines, with the reference's elementwise glue
// (products, combinations, Euler update, velocity correction) folded into the
// prologue/epilogue.  At one rank the pencils are one array, so the step is
//   transeq   X -> Y -> Z(+combine, Euler)        src/transeq.f90:44-355
//   divergence X -> Y -> Z                         src/navier.f90:98-236
//   Poisson                                        src/poisson.f90:276-419
//   gradp     Z -> Y -> X(+cor_vel)                src/navier.f90:249-359, 64-91
// Each pass reads its input fields once and writes its outputs once; all
// periodic operators use the circulant line solver.
#include "x3d_line.h"

namespace x3d {

// ---- transeq: r_c = D(u_d u_c) + u_d D(u_c), d = DIR ----------------------
// X: du_c = r_c; Y: tmp_c = r_c;
// Z: du_c = du_c + (tmp_c + r_c); if euler: u_c = gdt du_c + u_c.
// (src/transeq.f90:100-115, 120-134, 139-153, 158-177 per component)
template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_transeq(T* ux, T* uy, T* uz, T* dux, T* duy, T* duz, T* tx,
                                                  T* ty, T* tz, Coef<T> cf, Circ<T> cc, Geom geo, Tiling tl,
                                                  T gdt, int euler)
{
  constexpr int H = 2, NW = M + 2 * H;
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<T, DIR>(geo, tl, M);
  const int n = geo.n;
  T* const Ud = DIR == 0 ? ux : (DIR == 1 ? uy : uz);
  XBlock<T, M> xb;
  if constexpr (DIR == 0) xb = x_block<T, M>(geo, t, lds);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // components one after another: no loads hoisted across them
    __builtin_amdgcn_sched_barrier(0);
    // pointers straight from the kernel arguments (c is a constant here)
    T* const Uc = c == 0 ? ux : (c == 1 ? uy : uz);
    T* const DUc = c == 0 ? dux : (c == 1 ? duy : duz);
    T* const TMc = c == 0 ? tx : (c == 1 ? ty : tz);
    // right-hand sides of D(u_c) and D(u_d u_c) first, so that only the chunk
    // values stay live through the two line solves
    T v1[M], v2[M];
    {
      T wc[NW], wd[NW];
      if constexpr (DIR == 0) {
        x_load_window<T, M, H, H>(Uc, xb, t, wc);
        x_load_window<T, M, H, H>(Ud, xb, t, wd);
      } else {
        load_window<T, M, H, H, DIR != 0>(Uc, t, n, wc);
        load_window<T, M, H, H, DIR != 0>(Ud, t, n, wd);
      }
      rhs_chunk<K_D, T, M, H>(wc, cf, v2);
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int q = k + H;
        v1[k] = cf.a * (wd[q + 1] * wc[q + 1] - wd[q - 1] * wc[q - 1]) +
                cf.b * (wd[q + 2] * wc[q + 2] - wd[q - 2] * wc[q - 2]);
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
      T ud[M];
      x_load_chunk<T, M>(Ud, xb, ud);
#pragma unroll
      for (int k = 0; k < M; ++k) v1[k] = v1[k] + ud[k] * v2[k];
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) v1[k] = v1[k] + ld(Ud, t.at(t.j0 + k)) * v2[k];
    }
    if constexpr (DIR == 0) {
      x_store<T, M>(DUc, xb, v1);
    } else if constexpr (DIR == 1) {
      store_chunk<T, M>(TMc, t, v1);
    } else {
      if (t.active) {
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const uint32_t at = t.at(t.j0 + k);
          const T d = ld(DUc, at) + (ld(TMc, at) + v1[k]);
          st(DUc, at, d);
          if (euler) st(Uc, at, gdt * d + ld(Uc, at));
        }
      }
    }
  }
}

// ---- divergence passes (src/navier.f90:166-202) ----------------------------
// X: o1 = Dvp(a), o2 = Ivp(b), o3 = Ivp(c)        a, b, c = ux, uy, uz
// Y: o1 = Dvp(b) + Ivp(a), o2 = Ivp(c)             a, b, c = pp1, pgy1, pgz1
// Z: o1 = Ivp(a) + Dvp(b)                          a, b = duydypi2, upi2
template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_div(const T* a, const T* b, const T* c, T* o1, T* o2, T* o3,
                                              Coef<T> cfd, Circ<T> ccd, Coef<T> cfi, Circ<T> cci, Geom geo,
                                              Tiling tl)
{
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<T, DIR>(geo, tl, M);
  const int n = geo.n;
  if constexpr (DIR == 0) {
    const XBlock<T, M> xb = x_block<T, M>(geo, t, lds);
    T v[M];
    {
      T w[M + 3];
      x_load_window<T, M, 1, 2>(a, xb, t, w);
      rhs_chunk<K_DVP, T, M, 1>(w, cfd, v);
    }
    solve_circ<T, M, false>(v, t, ccd, lds);
    x_store<T, M>(o1, xb, v);
    {
      T w[M + 7];
      x_load_window<T, M, 3, 4>(b, xb, t, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, false>(v, t, cci, lds);
    x_store<T, M>(o2, xb, v);
    {
      T w[M + 7];
      x_load_window<T, M, 3, 4>(c, xb, t, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, false>(v, t, cci, lds);
    x_store<T, M>(o3, xb, v);
  } else if constexpr (DIR == 1) {
    T v[M], s[M];
    {
      T w[M + 3];
      load_window<T, M, 1, 2, DIR != 0>(b, t, n, w);
      rhs_chunk<K_DVP, T, M, 1>(w, cfd, v);
    }
    solve_circ<T, M, DIR != 0>(v, t, ccd, lds);
    {
      T w[M + 7];
      load_window<T, M, 3, 4, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, s);
    }
    solve_circ<T, M, DIR != 0>(s, t, cci, lds);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = v[k] + s[k];
    store_chunk<T, M>(o1, t, v);
    {
      T w[M + 7];
      load_window<T, M, 3, 4, DIR != 0>(c, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, DIR != 0>(v, t, cci, lds);
    store_chunk<T, M>(o2, t, v);
  } else {
    T v[M], s[M];
    {
      T w[M + 7];
      load_window<T, M, 3, 4, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IVP, T, M, 3>(w, cfi, v);
    }
    solve_circ<T, M, DIR != 0>(v, t, cci, lds);
    {
      T w[M + 3];
      load_window<T, M, 1, 2, DIR != 0>(b, t, n, w);
      rhs_chunk<K_DVP, T, M, 1>(w, cfd, s);
    }
    solve_circ<T, M, DIR != 0>(s, t, ccd, lds);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = v[k] + s[k];
    store_chunk<T, M>(o1, t, v);
  }
}

// ---- pressure-gradient passes (src/navier.f90:273-304) + cor_vel -----------
// Z: o1 = Ipv(a), o2 = Dpv(a)                       a = pp3
// Y: o1 = Ipv(a), o2 = Dpv(a), o3 = Ipv(b)          a, b = ppi3, pgz3
// X: o1 = Dpv(a), o2 = Ipv(b), o3 = Ipv(c)          a, b, c = ppi2, pgy2, pgzi2
//    if corr: u_k = u_k - o_k (src/navier.f90:84-86)
template <typename T, int M, int DIR>
__global__ __launch_bounds__(512) void k_grad(const T* a, const T* b, const T* c, T* o1, T* o2, T* o3,
                                               T* ux, T* uy, T* uz, Coef<T> cfd, Circ<T> ccd, Coef<T> cfi,
                                               Circ<T> cci, Geom geo, Tiling tl, int corr)
{
  extern __shared__ unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const LineThread t = line_thread<T, DIR>(geo, tl, M);
  const int n = geo.n;
  if constexpr (DIR == 2 || DIR == 1) {
    T v[M];
    {
      T w[M + 7];
      load_window<T, M, 4, 3, DIR != 0>(a, t, n, w);
      rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      T s[M];
      rhs_chunk<K_DPV, T, M, 4>(w, cfd, s);
      solve_circ<T, M, DIR != 0>(v, t, cci, lds);
      store_chunk<T, M>(o1, t, v);
      solve_circ<T, M, DIR != 0>(s, t, ccd, lds);
      store_chunk<T, M>(o2, t, s);
    }
    if constexpr (DIR == 1) {
      T w[M + 7];
      load_window<T, M, 4, 3, DIR != 0>(b, t, n, w);
      rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      solve_circ<T, M, DIR != 0>(v, t, cci, lds);
      store_chunk<T, M>(o3, t, v);
    }
  } else {
    const XBlock<T, M> xb = x_block<T, M>(geo, t, lds);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T* const inq = q == 0 ? a : (q == 1 ? b : c);
      T* const outq = q == 0 ? o1 : (q == 1 ? o2 : o3);
      T* const Uq = q == 0 ? ux : (q == 1 ? uy : uz);
      T v[M];
      {
        T w[M + 7];
        x_load_window<T, M, 4, 3>(inq, xb, t, w);
        if (q == 0) rhs_chunk<K_DPV, T, M, 4>(w, cfd, v);
        else rhs_chunk<K_IPV, T, M, 4>(w, cfi, v);
      }
      if (q == 0) solve_circ<T, M, false>(v, t, ccd, lds);
      else solve_circ<T, M, false>(v, t, cci, lds);
      x_store<T, M>(outq, xb, v);
      if (corr) {
        T u[M];
        x_load_chunk<T, M>(Uq, xb, u);
#pragma unroll
        for (int k = 0; k < M; ++k) u[k] = u[k] - v[k];
        x_store<T, M>(Uq, xb, u);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct DirSet {
  Geom geo;
  Tiling tl;
  dim3 grid, block;
  size_t smem;
  int M;
  Coef<double> cf_d, cf_dvp, cf_dpv, cf_ivp, cf_ipv;  // scaled by 1/gamma
  Circ<double> cc_d, cc_s, cc_i;
};

static DirSet dirset(const x3d_ctx* c, int dir)
{
  DirSet s;
  launch_dims(c, dir, &s.geo, &s.tl, &s.grid, &s.block, &s.smem, sizeof(double), true);
  const LinePlan& P = c->fplan[dir];
  s.M = P.M;
  const Scheme& sc = c->sc[dir];
  auto scaled = [](Coef<double> k, double alfa) {
    const double ig = 1.0 / circ_gamma(alfa);
    return Coef<double>{k.a * ig, k.b * ig, k.c * ig, k.d * ig};
  };
  s.cf_d = scaled(stencil_coef(sc, K_D), sc.alfai);
  s.cf_dvp = scaled(stencil_coef(sc, K_DVP), sc.alcai);
  s.cf_dpv = scaled(stencil_coef(sc, K_DPV), sc.alcai);
  s.cf_ivp = scaled(stencil_coef(sc, K_IVP), sc.ailcai);
  s.cf_ipv = scaled(stencil_coef(sc, K_IPV), sc.ailcai);
  s.cc_d = make_circ<double>(sc.alfai, P);
  s.cc_s = make_circ<double>(sc.alcai, P);
  s.cc_i = make_circ<double>(sc.ailcai, P);
  return s;
}

#define X3D_MSWITCH(M, CALL)       \
  switch (M) {                     \
    case 1: { constexpr int MM = 1; CALL; } break;   \
    case 2: { constexpr int MM = 2; CALL; } break;   \
    case 3: { constexpr int MM = 3; CALL; } break;   \
    case 4: { constexpr int MM = 4; CALL; } break;   \
    case 5: { constexpr int MM = 5; CALL; } break;   \
    case 6: { constexpr int MM = 6; CALL; } break;   \
    case 8: { constexpr int MM = 8; CALL; } break;   \
    case 12: { constexpr int MM = 12; CALL; } break; \
    case 16: { constexpr int MM = 16; CALL; } break; \
    default: return X3D_ERR_UNSUPPORTED;             \
  }

template <int DIR>
static int transeq_pass(x3d_ctx* c, double* ux, double* uy, double* uz, double* dux, double* duy, double* duz,
                        double* tx, double* ty, double* tz, double gdt, int euler)
{
  const DirSet s = dirset(c, DIR);
  X3D_MSWITCH(s.M, hipLaunchKernelGGL((k_transeq<double, MM, DIR>), s.grid, s.block, s.smem, c->stream, ux, uy,
                                      uz, dux, duy, duz, tx, ty, tz, s.cf_d, s.cc_d, s.geo, s.tl, gdt, euler));
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <int DIR>
static int div_pass(x3d_ctx* c, const double* a, const double* b, const double* cc, double* o1, double* o2,
                    double* o3)
{
  const DirSet s = dirset(c, DIR);
  X3D_MSWITCH(s.M, hipLaunchKernelGGL((k_div<double, MM, DIR>), s.grid, s.block, s.smem, c->stream, a, b, cc,
                                      o1, o2, o3, s.cf_dvp, s.cc_s, s.cf_ivp, s.cc_i, s.geo, s.tl));
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

template <int DIR>
static int grad_pass(x3d_ctx* c, const double* a, const double* b, const double* cc, double* o1, double* o2,
                     double* o3, double* ux, double* uy, double* uz, int corr)
{
  const DirSet s = dirset(c, DIR);
  X3D_MSWITCH(s.M, hipLaunchKernelGGL((k_grad<double, MM, DIR>), s.grid, s.block, s.smem, c->stream, a, b, cc,
                                      o1, o2, o3, ux, uy, uz, s.cf_dpv, s.cc_s, s.cf_ipv, s.cc_i, s.geo, s.tl,
                                      corr));
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}
#undef X3D_MSWITCH

bool fused_ok(const x3d_ctx* c)
{
  for (int o = 0; o < X3D_NOPS; ++o)
    if (!c->op[o].circulant) return false;
  return true;
}

#define X3D_TRY(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

int fused_momentum(x3d_ctx* c, double gdt, int euler, double* ux, double* uy, double* uz, double* dux,
                   double* duy, double* duz)
{
  X3D_TRY(ensure_work(c));
  double *tx = c->work[5], *ty = c->work[6], *tz = c->work[7];
  X3D_TRY(transeq_pass<0>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, 0));
  X3D_TRY(mark_pass(c, 1));
  X3D_TRY(transeq_pass<1>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, 0));
  X3D_TRY(mark_pass(c, 2));
  X3D_TRY(transeq_pass<2>(c, ux, uy, uz, dux, duy, duz, tx, ty, tz, gdt, euler));
  X3D_TRY(mark_pass(c, 3));
  return X3D_OK;
}

int fused_divergence(x3d_ctx* c, double* pp3, const double* ux, const double* uy, const double* uz)
{
  X3D_TRY(ensure_work(c));
  double *pp1 = c->work[0], *pgy1 = c->work[1], *pgz1 = c->work[2], *dd2 = c->work[3], *upi2 = c->work[4];
  X3D_TRY(div_pass<0>(c, ux, uy, uz, pp1, pgy1, pgz1));
  X3D_TRY(mark_pass(c, 4));
  X3D_TRY(div_pass<1>(c, pp1, pgy1, pgz1, dd2, upi2, nullptr));
  X3D_TRY(mark_pass(c, 5));
  X3D_TRY(div_pass<2>(c, dd2, upi2, nullptr, pp3, nullptr, nullptr));
  X3D_TRY(mark_pass(c, 6));
  return X3D_OK;
}

int fused_gradp(x3d_ctx* c, double* px, double* py, double* pz, const double* pp3, double* ux, double* uy,
                double* uz, int corr)
{
  X3D_TRY(ensure_work(c));
  double *ppi3 = c->work[0], *pgz3 = c->work[1], *ppi2 = c->work[2], *pgy2 = c->work[3], *pgzi2 = c->work[4];
  X3D_TRY(grad_pass<2>(c, pp3, nullptr, nullptr, ppi3, pgz3, nullptr, nullptr, nullptr, nullptr, 0));
  X3D_TRY(mark_pass(c, 8));
  X3D_TRY(grad_pass<1>(c, ppi3, pgz3, nullptr, ppi2, pgy2, pgzi2, nullptr, nullptr, nullptr, 0));
  X3D_TRY(mark_pass(c, 9));
  X3D_TRY(grad_pass<0>(c, ppi2, pgy2, pgzi2, px, py, pz, ux, uy, uz, corr));
  X3D_TRY(mark_pass(c, 10));
  return X3D_OK;
}

}  // namespace x3d
