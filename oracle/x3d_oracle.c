/*
 * x3d_oracle.c -- CPU oracle for the periodic x3div Euler-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (checker + CPU-baseline leg of bench.py); the HIP
 * product never links it.  Plain C, FP64, compiled without FP contraction so
 * every expression is evaluated in the reference's order.  Each function
 * names the reference routine (file:line under src/) it restates.  Periodic
 * boundary conditions only (ncl?1 = ncl?n = 0), which is the only case the
 * reference's Poisson solver accepts (src/poisson.f90:104-108).
 */
#include "x3d_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k, nx, ny) ((size_t)(i) + (size_t)(nx) * ((size_t)(j) + (size_t)(ny) * (size_t)(k)))

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------------ */
/* Thomas preparation: src/x3d_operator_1d.f90:301-323 (prepare)            */
/* ------------------------------------------------------------------------ */
void orc_prepare(const double *b, const double *cdiag, const double *f,
                 double *s, double *w, int n)
{
  for (int i = 0; i < n; ++i) w[i] = cdiag[i];
  for (int i = 1; i < n; ++i) {
    s[i] = b[i - 1] / w[i - 1];
    w[i] = w[i] - f[i - 1] * s[i];
  }
  for (int i = 0; i < n; ++i) w[i] = 1.0 / w[i];
}

/* src/thomas.f90:200-218 (thomas1d) */
void orc_thomas1d(double *tt, const double *ff, const double *fs, const double *fw, int n)
{
  for (int k = 1; k < n; ++k) tt[k] = tt[k] - tt[k - 1] * fs[k];
  tt[n - 1] = tt[n - 1] * fw[n - 1];
  for (int k = n - 2; k >= 0; --k) tt[k] = (tt[k] - ff[k] * tt[k + 1]) * fw[k];
}

/* Periodic tridiagonal rows shared by first_deriv_imp_ and interpol_imp_:
 * upper f, diagonal c, lower b with the periodic corner modification
 * (src/x3d_operator_1d.f90:414-471 and :576-599).  Assignment order follows
 * the reference so that small n behaves identically. */
static void periodic_rows(double a, double *f, double *c, double *b, int n)
{
  for (int i = 0; i < n; ++i) { f[i] = 0.0; c[i] = 0.0; b[i] = 0.0; }
  f[0] = a; f[1] = a; c[0] = 2.0; c[1] = 1.0; b[0] = a; b[1] = a;
  f[n - 3] = a; f[n - 2] = a; f[n - 1] = 0.0;
  c[n - 3] = 1.0; c[n - 2] = 1.0; c[n - 1] = 1.0 + a * a;
  b[n - 3] = a; b[n - 2] = a; b[n - 1] = 0.0;
  for (int i = 2; i <= n - 4; ++i) { f[i] = a; c[i] = 1.0; b[i] = a; }
}

/* src/x3d_operator_1d.f90:234-266 (init): pointers + periodic vector
 * periodic = thomas1d((-1, 0, ..., 0, alfa)). */
static int op_init(orc_op *op, const double *f, const double *s, const double *w,
                   int n, double alfa, int paire)
{
  op->n = n; op->npaire = paire; op->alfa = alfa;
  op->f = malloc(sizeof(double) * n); op->s = malloc(sizeof(double) * n);
  op->w = malloc(sizeof(double) * n); op->periodic = malloc(sizeof(double) * n);
  if (!op->f || !op->s || !op->w || !op->periodic) return -1;
  memcpy(op->f, f, sizeof(double) * n);
  memcpy(op->s, s, sizeof(double) * n);
  memcpy(op->w, w, sizeof(double) * n);
  op->periodic[0] = -1.0;
  for (int i = 1; i < n - 1; ++i) op->periodic[i] = 0.0;
  op->periodic[n - 1] = alfa;
  orc_thomas1d(op->periodic, f, s, w, n);
  return 0;
}

/* Scheme constants: first_deriv_exp_ (ifirstder=4, :350-353) and
 * interpol_exp_ (ifirstder=4, ipinter=3, :509-545). */
static void scheme_consts(orc_ctx *c, int d, double h)
{
  c->alfai[d] = 1.0 / 3.0;
  c->afi[d] = (7.0 / 9.0) / h;
  c->bfi[d] = (1.0 / 36.0) / h;
  c->alcai[d] = 9.0 / 62.0;
  c->aci[d] = (63.0 / 62.0) / h;
  c->bci[d] = (17.0 / 62.0) / 3.0 / h;
  c->ailcai[d] = 0.49;
  c->aici[d] = 1.0 / 128.0 * (75.0 + 70.0 * c->ailcai[d]);
  c->bici[d] = 1.0 / 256.0 * (126.0 * c->ailcai[d] - 25.0);
  c->cici[d] = 1.0 / 256.0 * (-10.0 * c->ailcai[d] + 3.0);
  c->dici[d] = 0.0;
  /* second derivative (isecondder = 4, src/parameters.f90:107; the reference
   * implements none): Lele's (1992) sixth-order tridiagonal scheme,
   * alfa = 2/11, a = 12/11 on (u+1 - 2u + u-1)/h^2, b = 3/11 on (u+2 - 2u + u-2)/(4h^2) */
  c->alsai[d] = 2.0 / 11.0;
  c->asi[d] = (12.0 / 11.0) / (h * h);
  c->bsi[d] = (3.0 / 44.0) / (h * h);
}

/* One direction of x3d_operator_1d_init (src/x3d_operator_1d.f90:76-189):
 * builds the D (derX / derXp), Ds (vp / pv) and I (vp / pv) operators. */
static int dir_ops(orc_ctx *c, int d, int n)
{
  double *f = calloc(n, sizeof(double)), *cc = calloc(n, sizeof(double));
  double *b = calloc(n, sizeof(double)), *s = calloc(n, sizeof(double));
  double *w = calloc(n, sizeof(double)), *fp = calloc(n, sizeof(double));
  double *sp = calloc(n, sizeof(double)), *wp = calloc(n, sizeof(double));
  int rc = 0;
  if (!f || !cc || !b || !s || !w || !fp || !sp || !wp) { rc = -1; goto out; }

  /* D: first_deriv_imp_ periodic branch; ffp = ff (ncl1 /= 1), so derX and
   * derXp carry identical coefficients. */
  periodic_rows(c->alfai[d], f, cc, b, n);
  orc_prepare(b, cc, f, s, w, n);
  rc |= op_init(&c->ops[ORC_OP_DERX + 2 * d], f, s, w, n, c->alfai[d], 0);
  rc |= op_init(&c->ops[ORC_OP_DERXP + 2 * d], f, s, w, n, c->alfai[d], 1);

  /* Ds: cfx6/csx6/cwx6, used for both vp (npaire 1) and pv (npaire 0). */
  periodic_rows(c->alcai[d], f, cc, b, n);
  memset(s, 0, sizeof(double) * n);
  orc_prepare(b, cc, f, s, w, n);
  rc |= op_init(&c->ops[ORC_OP_DERXVP + d], f, s, w, n, c->alcai[d], 1);
  rc |= op_init(&c->ops[ORC_OP_DERXPV + d], f, s, w, n, c->alcai[d], 0);

  /* I: intXpv uses cifx6/cisx6/ciwx6 and intXvp the "p" copies
   * cifxp6/cisxp6/ciwxp6 (src/x3d_operator_1d.f90:170,176).  cifxp6 is a
   * plain copy of cifx6 (:667-670; only cfxp6(1) and cfip6(1) are zeroed,
   * :675-676, and neither reaches the periodic path), so both prepare()
   * calls (:679-680) see identical rows. */
  periodic_rows(c->ailcai[d], f, cc, b, n);
  memset(s, 0, sizeof(double) * n);
  orc_prepare(b, cc, f, s, w, n);
  rc |= op_init(&c->ops[ORC_OP_INTXPV + d], f, s, w, n, c->ailcai[d], 1);
  memcpy(fp, f, sizeof(double) * n);
  memset(sp, 0, sizeof(double) * n);
  orc_prepare(b, cc, fp, sp, wp, n);
  rc |= op_init(&c->ops[ORC_OP_INTXVP + d], fp, sp, wp, n, c->ailcai[d], 1);

  /* DD: the same periodic rows with alfa = 2/11 */
  periodic_rows(c->alsai[d], f, cc, b, n);
  memset(s, 0, sizeof(double) * n);
  orc_prepare(b, cc, f, s, w, n);
  rc |= op_init(&c->ops[ORC_OP_DERXX + d], f, s, w, n, c->alsai[d], 0);
out:
  free(f); free(cc); free(b); free(s); free(w); free(fp); free(sp); free(wp);
  return rc;
}

/* Poisson tables: abxyz (src/poisson.f90:421-479), waves periodic branch
 * (:515-616, :660-699). */
static void poisson_tables(orc_ctx *c)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double twopi = 2.0 * ORC_PI;
  for (int i = 0; i < nx; ++i) {
    c->ax[i] = sin((double)i * ORC_PI / (double)nx);
    c->bx[i] = cos((double)i * ORC_PI / (double)nx);
  }
  for (int j = 0; j < ny; ++j) {
    c->ay[j] = sin((double)j * ORC_PI / (double)ny);
    c->by[j] = cos((double)j * ORC_PI / (double)ny);
  }
  for (int k = 0; k < nz; ++k) {
    c->az[k] = sin((double)k * ORC_PI / (double)nz);
    c->bz[k] = cos((double)k * ORC_PI / (double)nz);
  }
  /* modified wavenumbers, stored as the (equal) real part of (1+i)*value */
  for (int i = 0; i < nx / 2 + 1; ++i) {
    double w = twopi * (double)i / (double)nx;
    double wp = c->aci[0] * 2.0 * c->dx * sin(w * 0.5) + c->bci[0] * 2.0 * c->dx * sin(3.0 * 0.5 * w);
    wp = wp / (1.0 + 2.0 * c->alcai[0] * cos(w));
    double v = (double)nx * wp / c->xlx;
    c->xk2[i] = v * v;
    c->exs[i] = (double)nx * w / c->xlx;
  }
  for (int i = nx / 2 + 1; i < nx; ++i) { c->xk2[i] = c->xk2[nx - i]; c->exs[i] = c->exs[nx - i]; }
  for (int j = 0; j < ny / 2 + 1; ++j) {
    double w = twopi * (double)j / (double)ny;
    double wp = c->aci[1] * 2.0 * c->dy * sin(w * 0.5) + c->bci[1] * 2.0 * c->dy * sin(3.0 * 0.5 * w);
    wp = wp / (1.0 + 2.0 * c->alcai[1] * cos(w));
    double v = (double)ny * wp / c->yly;
    c->yk2[j] = v * v;
    c->eys[j] = (double)ny * w / c->yly;
  }
  for (int j = ny / 2 + 1; j < ny; ++j) { c->yk2[j] = c->yk2[ny - j]; c->eys[j] = c->eys[ny - j]; }
  for (int k = 0; k < nz / 2 + 1; ++k) {
    double w = twopi * (double)k / (double)nz;
    double wp = c->aci[2] * 2.0 * c->dz * sin(w * 0.5) + (c->bci[2] * 2.0 * c->dz) * sin(3.0 * 0.5 * w);
    wp = wp / (1.0 + 2.0 * c->alcai[2] * cos(w));
    double v = (double)nz * wp / c->zlz;
    c->zk2[k] = v * v;
    c->ezs[k] = (double)nz * w / c->zlz;
  }
  const int nzh = nz / 2 + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nzh; ++k) {
    double rlezs = c->ezs[k] * c->dz;
    double ztt = 2.0 * (c->bici[2] * cos(rlezs * 1.5) + c->cici[2] * cos(rlezs * 2.5) + c->dici[2] * cos(rlezs * 3.5));
    double ztt1 = 2.0 * c->aici[2] * cos(rlezs * 0.5);
    double zt1 = 1.0 + 2.0 * c->ailcai[2] * cos(rlezs);
    for (int j = 0; j < ny; ++j) {
      double rleys = c->eys[j] * c->dy;
      double ytt = 2.0 * (c->bici[1] * cos(rleys * 1.5) + c->cici[1] * cos(rleys * 2.5) + c->dici[1] * cos(rleys * 3.5));
      double ytt1 = 2.0 * c->aici[1] * cos(rleys * 0.5);
      double yt1 = 1.0 + 2.0 * c->ailcai[1] * cos(rleys);
      for (int i = 0; i < nx; ++i) {
        double rlexs = c->exs[i] * c->dx;
        double xtt = 2.0 * (c->bici[0] * cos(rlexs * 1.5) + c->cici[0] * cos(rlexs * 2.5) + c->dici[0] * cos(rlexs * 3.5));
        double xtt1 = 2.0 * c->aici[0] * cos(rlexs * 0.5);
        double xt1 = 1.0 + 2.0 * c->ailcai[0] * cos(rlexs);
        double tyz = ((ytt1 + ytt) / yt1) * ((ztt1 + ztt) / zt1);
        double txz = ((xtt1 + xtt) / xt1) * ((ztt1 + ztt) / zt1);
        double txy = ((xtt1 + xtt) / xt1) * ((ytt1 + ytt) / yt1);
        double xt2 = c->xk2[i] * (tyz * tyz);
        double yt2 = c->yk2[j] * (txz * txz);
        double zt2 = c->zk2[k] * (txy * txy);
        double K = xt2 + yt2 + zt2;
        size_t q = IDX(i, j, k, nx, ny);
        c->kxyz[2 * q] = K;
        c->kxyz[2 * q + 1] = K;
      }
    }
  }
}

/* init_xcompact3d's numerics part: parameter() (src/parameters.f90:24-90)
 * with periodic defaults, x3d_operator_1d_init and decomp_2d_poisson_init. */
int orc_init(orc_ctx *c, int nx, int ny, int nz, double xlx, double yly, double zlz)
{
  memset(c, 0, sizeof(*c));
  if (nx < 8 || ny < 8 || nz < 8) return -1;
  c->nx = nx; c->ny = ny; c->nz = nz;
  c->nxm = nx; c->nym = ny; c->nzm = nz;   /* periodic: nXm = nX */
  c->xlx = xlx; c->yly = yly; c->zlz = zlz;
  c->dx = xlx / (double)c->nxm; c->dy = yly / (double)c->nym; c->dz = zlz / (double)c->nzm;
  c->eps = 1.0e-16;
  scheme_consts(c, 0, c->dx);
  scheme_consts(c, 1, c->dy);
  scheme_consts(c, 2, c->dz);
  if (dir_ops(c, 0, nx) || dir_ops(c, 1, ny) || dir_ops(c, 2, nz)) return -2;
  c->ax = calloc(nx, sizeof(double)); c->bx = calloc(nx, sizeof(double));
  c->ay = calloc(ny, sizeof(double)); c->by = calloc(ny, sizeof(double));
  c->az = calloc(nz, sizeof(double)); c->bz = calloc(nz, sizeof(double));
  c->xk2 = calloc(nx, sizeof(double)); c->exs = calloc(nx, sizeof(double));
  c->yk2 = calloc(ny, sizeof(double)); c->eys = calloc(ny, sizeof(double));
  c->zk2 = calloc(nz / 2 + 1, sizeof(double)); c->ezs = calloc(nz / 2 + 1, sizeof(double));
  c->kxyz = calloc((size_t)2 * nx * ny * (nz / 2 + 1), sizeof(double));
  if (!c->ax || !c->kxyz) return -3;
  poisson_tables(c);
  return 0;
}

void orc_free(orc_ctx *c)
{
  for (int o = 0; o < ORC_NOPS_ALL; ++o) {
    free(c->ops[o].f); free(c->ops[o].s); free(c->ops[o].w); free(c->ops[o].periodic);
  }
  free(c->ax); free(c->bx); free(c->ay); free(c->by); free(c->az); free(c->bz);
  free(c->xk2); free(c->exs); free(c->yk2); free(c->eys); free(c->zk2); free(c->ezs);
  free(c->kxyz);
  free(c->ppy); free(c->ppyi);
  memset(c, 0, sizeof(*c));
}

orc_ctx *orc_new(int nx, int ny, int nz, double xlx, double yly, double zlz)
{
  orc_ctx *c = malloc(sizeof(orc_ctx));
  if (!c) return NULL;
  if (orc_init(c, nx, ny, nz, xlx, yly, zlz)) { orc_free(c); free(c); return NULL; }
  return c;
}

void orc_delete(orc_ctx *c)
{
  if (!c) return;
  orc_free(c);
  free(c);
}

int orc_get_op(const orc_ctx *c, int op, double *f, double *s, double *w,
               double *periodic, double *alfa, int *n, int *npaire)
{
  if (op < 0 || op >= ORC_NOPS_ALL) return -1;
  const orc_op *o = &c->ops[op];
  memcpy(f, o->f, sizeof(double) * o->n);
  memcpy(s, o->s, sizeof(double) * o->n);
  memcpy(w, o->w, sizeof(double) * o->n);
  memcpy(periodic, o->periodic, sizeof(double) * o->n);
  *alfa = o->alfa; *n = o->n; *npaire = o->npaire;
  return 0;
}

void orc_get_consts(const orc_ctx *c, double *out)
{
  for (int d = 0; d < 3; ++d) {
    double *o = out + 11 * d;
    o[0] = c->alfai[d]; o[1] = c->afi[d]; o[2] = c->bfi[d];
    o[3] = c->alcai[d]; o[4] = c->aci[d]; o[5] = c->bci[d];
    o[6] = c->ailcai[d]; o[7] = c->aici[d]; o[8] = c->bici[d]; o[9] = c->cici[d]; o[10] = c->dici[d];
  }
}

void orc_get_kxyz(const orc_ctx *c, double *re, double *im)
{
  size_t ns = (size_t)c->nx * c->ny * (c->nz / 2 + 1);
  for (size_t q = 0; q < ns; ++q) { re[q] = c->kxyz[2 * q]; im[q] = c->kxyz[2 * q + 1]; }
}

void orc_get_ab(const orc_ctx *c, double *out)
{
  memcpy(out, c->ax, sizeof(double) * c->nx); out += c->nx;
  memcpy(out, c->bx, sizeof(double) * c->nx); out += c->nx;
  memcpy(out, c->ay, sizeof(double) * c->ny); out += c->ny;
  memcpy(out, c->by, sizeof(double) * c->ny); out += c->ny;
  memcpy(out, c->az, sizeof(double) * c->nz); out += c->nz;
  memcpy(out, c->bz, sizeof(double) * c->nz);
}

/* ------------------------------------------------------------------------ */
/* Batched Thomas: src/thomas.f90:34-188                                     */
/* ------------------------------------------------------------------------ */
void orc_xthomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      double *t = tt + IDX(0, j, k, nx, ny);
      for (int i = 1; i < nx; ++i) t[i] = t[i] - t[i - 1] * fs[i];
      t[nx - 1] = t[nx - 1] * fw[nx - 1];
      for (int i = nx - 2; i >= 0; --i) t[i] = (t[i] - ff[i] * t[i + 1]) * fw[i];
    }
}

void orc_xthomas_0(double *tt, const orc_op *op, int nx, int ny, int nz)
{
  orc_xthomas_12(tt, op->f, op->s, op->w, nx, ny, nz);
  const double *p = op->periodic; const double a = op->alfa;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      double *t = tt + IDX(0, j, k, nx, ny);
      double ss = (t[0] - a * t[nx - 1]) / (1.0 + p[0] - a * p[nx - 1]);
      for (int i = 0; i < nx; ++i) t[i] = t[i] - ss * p[i];
    }
}

/* The y and z sweeps run the lines of a plane side by side (i innermost);
 * each line sees the same operations in the same order as the reference's
 * line-by-line sweeps (src/thomas.f90:114-136, 166-188). */
void orc_ythomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k) {
    for (int j = 1; j < ny; ++j) {
      double *t = tt + IDX(0, j, k, nx, ny); const double *tm = t - nx;
      for (int i = 0; i < nx; ++i) t[i] = t[i] - tm[i] * fs[j];
    }
    double *tl = tt + IDX(0, ny - 1, k, nx, ny);
    for (int i = 0; i < nx; ++i) tl[i] = tl[i] * fw[ny - 1];
    for (int j = ny - 2; j >= 0; --j) {
      double *t = tt + IDX(0, j, k, nx, ny); const double *tp = t + nx;
      for (int i = 0; i < nx; ++i) t[i] = (t[i] - ff[j] * tp[i]) * fw[j];
    }
  }
}

void orc_ythomas_0(double *tt, const orc_op *op, int nx, int ny, int nz)
{
  orc_ythomas_12(tt, op->f, op->s, op->w, nx, ny, nz);
  const double *p = op->periodic; const double a = op->alfa;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      double ss = (tt[IDX(i, 0, k, nx, ny)] - a * tt[IDX(i, ny - 1, k, nx, ny)]) / (1.0 + p[0] - a * p[ny - 1]);
      for (int j = 0; j < ny; ++j) tt[IDX(i, j, k, nx, ny)] = tt[IDX(i, j, k, nx, ny)] - ss * p[j];
    }
}

void orc_zthomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz)
{
  const size_t pl = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; ++j) {
    double *t0 = tt + IDX(0, j, 0, nx, ny);
    for (int k = 1; k < nz; ++k) {
      double *t = t0 + k * pl; const double *tm = t - pl;
      for (int i = 0; i < nx; ++i) t[i] = t[i] - tm[i] * fs[k];
    }
    double *tl = t0 + (nz - 1) * pl;
    for (int i = 0; i < nx; ++i) tl[i] = tl[i] * fw[nz - 1];
    for (int k = nz - 2; k >= 0; --k) {
      double *t = t0 + k * pl; const double *tp = t + pl;
      for (int i = 0; i < nx; ++i) t[i] = (t[i] - ff[k] * tp[i]) * fw[k];
    }
  }
}

void orc_zthomas_0(double *tt, const orc_op *op, int nx, int ny, int nz)
{
  orc_zthomas_12(tt, op->f, op->s, op->w, nx, ny, nz);
  const double *p = op->periodic; const double a = op->alfa;
  const size_t pl = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; ++j) {
    double *t0 = tt + IDX(0, j, 0, nx, ny);
    const double *tl = t0 + (nz - 1) * pl;
    const double den = 1.0 + p[0] - a * p[nz - 1];
    for (int i = 0; i < nx; ++i) {
      double ss = (t0[i] - a * tl[i]) / den;
      for (int k = 0; k < nz; ++k) t0[i + k * pl] = t0[i + k * pl] - ss * p[k];
    }
  }
}

int orc_thomas(const orc_ctx *c, double *tt, int op, int dir, int nx, int ny, int nz)
{
  if (op < 0 || op >= ORC_NOPS_ALL) return -1;
  const orc_op *o = &c->ops[op];
  int nl = dir == 0 ? nx : dir == 1 ? ny : nz;
  if (o->n != nl) return -2;
  if (dir == 0) orc_xthomas_0(tt, o, nx, ny, nz);
  else if (dir == 1) orc_ythomas_0(tt, o, nx, ny, nz);
  else orc_zthomas_0(tt, o, nx, ny, nz);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Periodic right-hand-side stencils.                                        */
/* The reference writes the wrapped rows out explicitly; every one of them   */
/* is the interior formula with indices taken modulo n, which is what is     */
/* evaluated here (operand order preserved):                                 */
/*  D   derx_00      src/x3d_derive.f90:128-143                              */
/*  Dvp derxvp       src/x3d_staggered.f90:41-56                             */
/*  Dpv derxpv       src/x3d_staggered.f90:239-254                           */
/*  Ivp interxvp     src/x3d_staggered.f90:128-165                           */
/*  Ipv interxpv     src/x3d_staggered.f90:307-344                           */
/* and likewise for y (:423-481, :561-591, :656-714, :809-839) and z         */
/* (:911-938, :1025-1081, :1167-1195, :1257-1313).                            */
/* ------------------------------------------------------------------------ */
static inline int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }

static double rhs_point(const double *u, long stride, int i, int n, int kind,
                        double a, double b, double cc, double d)
{
#define U(o) u[(long)wrap(i + (o), n) * stride]
  switch (kind) {
  case 0: return a * (U(1) - U(-1)) + b * (U(2) - U(-2));
  case 1: return a * (U(1) - U(0)) + b * (U(2) - U(-1));
  case 2: return a * (U(0) - U(-1)) + b * (U(1) - U(-2));
  case 3: return a * (U(1) + U(0)) + b * (U(2) + U(-1)) + cc * (U(3) + U(-2)) + d * (U(4) + U(-3));
  case 5: return a * ((U(1) + U(-1)) - 2.0 * U(0)) + b * ((U(2) + U(-2)) - 2.0 * U(0));
  default: return a * (U(0) + U(-1)) + b * (U(1) + U(-2)) + cc * (U(2) + U(-3)) + d * (U(3) + U(-4));
  }
#undef U
}

/* The stencil of rhs_point over a run of len points: U[o + 4] points to the
 * operand at offset o of the run's first point (o = -4..4), consecutive
 * points are consecutive in memory.  Same expressions, same operand order. */
static void rhs_run(double *t, const double *const *U, int len, int kind,
                    double a, double b, double cc, double d)
{
  switch (kind) {
  case 0: for (int i = 0; i < len; ++i) t[i] = a * (U[5][i] - U[3][i]) + b * (U[6][i] - U[2][i]); break;
  case 1: for (int i = 0; i < len; ++i) t[i] = a * (U[5][i] - U[4][i]) + b * (U[6][i] - U[3][i]); break;
  case 2: for (int i = 0; i < len; ++i) t[i] = a * (U[4][i] - U[3][i]) + b * (U[5][i] - U[2][i]); break;
  case 3:
    for (int i = 0; i < len; ++i)
      t[i] = a * (U[5][i] + U[4][i]) + b * (U[6][i] + U[3][i]) + cc * (U[7][i] + U[2][i]) + d * (U[8][i] + U[1][i]);
    break;
  case 5:
    for (int i = 0; i < len; ++i)
      t[i] = a * ((U[5][i] + U[3][i]) - 2.0 * U[4][i]) + b * ((U[6][i] + U[2][i]) - 2.0 * U[4][i]);
    break;
  default:
    for (int i = 0; i < len; ++i)
      t[i] = a * (U[4][i] + U[3][i]) + b * (U[5][i] + U[2][i]) + cc * (U[6][i] + U[1][i]) + d * (U[7][i] + U[0][i]);
    break;
  }
}

int orc_set_stretching(orc_ctx *c, int istret, const double *ppy, const double *ppyi)
{
  if (istret && (!ppy || !ppyi)) return -1;
  free(c->ppy); free(c->ppyi);
  c->ppy = c->ppyi = NULL;
  c->istret = istret;
  if (!istret) return 0;
  c->ppy = malloc(sizeof(double) * (size_t)c->ny);
  c->ppyi = malloc(sizeof(double) * (size_t)c->ny);
  memcpy(c->ppy, ppy, sizeof(double) * (size_t)c->ny);
  memcpy(c->ppyi, ppyi, sizeof(double) * (size_t)c->ny);
  return 0;
}

/* The stretching loop after the y solve: dery_00 / dery_ij
 * (src/x3d_derive.f90:331-337, 427-433: ppy), deryvp (src/x3d_staggered.f90:628-634:
 * ppyi), derypv (:874-880: ppy).  Interpolations are not scaled. */
static void stretch_y(const orc_ctx *c, double *t, int kind, int nx, int ny, int nz)
{
  if (!c->istret || !(kind == 0 || kind == 1 || kind == 2)) return;
  const double *s = kind == 1 ? c->ppyi : c->ppy;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) t[IDX(i, j, k, nx, ny)] = t[IDX(i, j, k, nx, ny)] * s[j];
}

int orc_apply(const orc_ctx *c, double *t, const double *u, int op, int dir,
              int kind, int nx, int ny, int nz)
{
  if (op < 0 || op >= ORC_NOPS_ALL || dir < 0 || dir > 2 || kind < 0 || kind > 5) return -1;
  double a, b, cc = 0.0, d = 0.0;
  if (kind == 0) { a = c->afi[dir]; b = c->bfi[dir]; }
  else if (kind == 5) { a = c->asi[dir]; b = c->bsi[dir]; }
  else if (kind <= 2) { a = c->aci[dir]; b = c->bci[dir]; }
  else { a = c->aici[dir]; b = c->bici[dir]; cc = c->cici[dir]; d = c->dici[dir]; }
  int n = dir == 0 ? nx : dir == 1 ? ny : nz;
  if (c->ops[op].n != n) return -2;
  long stride = dir == 0 ? 1 : dir == 1 ? nx : (long)nx * ny;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      size_t q = IDX(0, j, k, nx, ny);
      const double *U[9];
      if (dir == 0) {
        /* interior points of the x-line as one run, the wrapped ends point by point */
        const double *line = u + q;
        for (int o = -4; o <= 4; ++o) U[o + 4] = line + 4 + o;
        if (nx > 8) rhs_run(t + q + 4, U, nx - 8, kind, a, b, cc, d);
        for (int i = 0; i < nx; ++i)
          if (i < 4 || i >= nx - 4 || nx <= 8) t[q + i] = rhs_point(line, 1, i, n, kind, a, b, cc, d);
      } else {
        /* a whole x-row at once: each offset is one wrapped neighbour row */
        int pos = dir == 1 ? j : k;
        const double *row0 = u + q - (size_t)pos * stride;
        for (int o = -4; o <= 4; ++o) U[o + 4] = row0 + (long)wrap(pos + o, n) * stride;
        rhs_run(t + q, U, nx, kind, a, b, cc, d);
      }
    }
  const int rc = orc_thomas(c, t, op, dir, nx, ny, nz);
  if (rc == 0 && dir == 1) stretch_y(c, t, kind, nx, ny, nz);
  return rc;
}

/* derx_00 / dery_00 / derz_00: src/x3d_derive.f90:111-152, 279-339, 503-558
 * (dery's ppy scaling when the context is stretched, in orc_apply). */
void orc_derx(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz)
{ orc_apply(c, t, u, op, 0, 0, nx, ny, nz); }
void orc_dery(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz)
{ orc_apply(c, t, u, op, 1, 0, nx, ny, nz); }
void orc_derz(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz)
{ orc_apply(c, t, u, op, 2, 0, nx, ny, nz); }

/* ------------------------------------------------------------------------ */
/* Non-periodic first derivatives: der?_11/12/21/22                          */
/* (src/x3d_derive.f90:156-275, and the y/z copies :343-499, :562-725).      */
/* Boundary codes: 1 free-slip (npaire chooses the even or odd extension),   */
/* 2 Dirichlet (one-sided compact closure).  The grid spacing is L/(n-1).    */
/* ------------------------------------------------------------------------ */
/* first_deriv_exp_, ifirstder = 4 (src/x3d_operator_1d.f90:325-393) */
void orc_np_consts(double h, double *k /* 15 */)
{
  k[0] = 1.0 / 3.0;            /* alfai */
  k[1] = (7.0 / 9.0) / h;      /* afi   */
  k[2] = (1.0 / 36.0) / h;     /* bfi   */
  k[3] = 2.0;                  /* alfa1 */
  k[4] = -(5.0 / 2.0) / h;     /* af1   */
  k[5] = 2.0 / h;              /* bf1   */
  k[6] = 0.5 / h;              /* cf1   */
  k[7] = 1.0 / 4.0;            /* alfa2 */
  k[8] = (3.0 / 4.0) / h;      /* af2   */
  k[9] = 2.0;                  /* alfan */
  k[10] = -(5.0 / 2.0) / h;    /* afn   */
  k[11] = 2.0 / h;             /* bfn   */
  k[12] = 0.5 / h;             /* cfn   */
  k[13] = 1.0 / 4.0;           /* alfam */
  k[14] = (3.0 / 4.0) / h;     /* afm   */
}

/* first_deriv_imp_ for ncl1, ncln in {1, 2} (src/x3d_operator_1d.f90:395-491):
 * the rows (lower fb, diagonal fc, upper ff) and the Thomas vectors of the
 * npaire = 0 instance (ff, fs, fw) and of the npaire = 1 instance (ffp, fsp,
 * fwp).  rows (3n, optional) receives fb, fc, ff of the first, rowsp those of
 * the second. */
int orc_np_tables(int n, double h, int ncl1, int ncln, double *ff, double *fs, double *fw,
                  double *ffp, double *fsp, double *fwp, double *rows, double *rowsp)
{
  if (n < 8 || ncl1 < 1 || ncl1 > 2 || ncln < 1 || ncln > 2) return -1;
  double k[15];
  orc_np_consts(h, k);
  const double alfai = k[0], alfa1 = k[3], alfa2 = k[7], alfan = k[9], alfam = k[13];
  double *fb = calloc(n, sizeof(double)), *fc = calloc(n, sizeof(double));
  if (!fb || !fc) { free(fb); free(fc); return -2; }
  for (int i = 0; i < n; ++i) { ff[i] = fs[i] = fw[i] = ffp[i] = fsp[i] = fwp[i] = 0.0; }
  if (ncl1 == 1) {
    ff[0] = alfai + alfai; ff[1] = alfai; fc[0] = 1.0; fc[1] = 1.0; fb[0] = alfai; fb[1] = alfai;
  } else {
    ff[0] = alfa1; ff[1] = alfa2; fc[0] = 1.0; fc[1] = 1.0; fb[0] = alfa2; fb[1] = alfai;
  }
  if (ncln == 1) {
    ff[n - 3] = alfai; ff[n - 2] = alfai; ff[n - 1] = 0.0;
    fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0;
    fb[n - 3] = alfai; fb[n - 2] = alfai + alfai; fb[n - 1] = 0.0;
  } else {
    ff[n - 3] = alfai; ff[n - 2] = alfam; ff[n - 1] = 0.0;
    fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0;
    fb[n - 3] = alfam; fb[n - 2] = alfan; fb[n - 1] = 0.0;
  }
  for (int i = 2; i <= n - 4; ++i) { ff[i] = alfai; fc[i] = 1.0; fb[i] = alfai; }
  for (int i = 0; i < n; ++i) ffp[i] = ff[i];
  if (rows) { memcpy(rows, fb, sizeof(double) * n); memcpy(rows + n, fc, sizeof(double) * n); memcpy(rows + 2 * n, ff, sizeof(double) * n); }
  orc_prepare(fb, fc, ff, fs, fw, n);
  if (ncl1 == 1) ffp[0] = 0.0;
  if (ncln == 1) fb[n - 2] = 0.0;
  if (rowsp) { memcpy(rowsp, fb, sizeof(double) * n); memcpy(rowsp + n, fc, sizeof(double) * n); memcpy(rowsp + 2 * n, ffp, sizeof(double) * n); }
  orc_prepare(fb, fc, ffp, fsp, fwp, n);
  free(fb); free(fc);
  return 0;
}

/* The right-hand side of derx_ij at point i of one line u[i * stride]. */
static double np_rhs(const double *u, long stride, int i, int n, int ncl1, int ncln, int npaire,
                     const double *k)
{
#define V(o) u[(long)(i + (o)) * stride]
  const double afi = k[1], bfi = k[2];
  if (i >= 2 && i <= n - 3) return afi * (V(1) - V(-1)) + bfi * (V(2) - V(-2));
  if (i == 0) {
    if (ncl1 == 1) return npaire == 1 ? 0.0 : afi * (V(1) + V(1)) + bfi * (V(2) + V(2));
    return k[4] * V(0) + k[5] * V(1) + k[6] * V(2);
  }
  if (i == 1) {
    if (ncl1 == 1)
      return npaire == 1 ? afi * (V(1) - V(-1)) + bfi * (V(2) - V(0)) : afi * (V(1) - V(-1)) + bfi * (V(2) + V(0));
    return k[8] * (V(1) - V(-1));
  }
  if (i == n - 2) {
    if (ncln == 1)
      return npaire == 1 ? afi * (V(1) - V(-1)) + bfi * (V(0) - V(-2)) : afi * (V(1) - V(-1)) + bfi * ((-V(0)) - V(-2));
    return k[14] * (V(1) - V(-1));
  }
  /* i == n - 1 */
  if (ncln == 1) return npaire == 1 ? 0.0 : afi * ((-V(-1)) - V(-1)) + bfi * ((-V(-2)) - V(-2));
  return -k[10] * V(0) - k[11] * V(-1) - k[12] * V(-2);
#undef V
}

/* derx_ij / dery_ij / derz_ij: the right-hand side above on every line along
 * dir, then the non-periodic Thomas sweeps (xthomas_12, src/thomas.f90:61-84)
 * with the instance's tables (ff, fs, fw). */
int orc_der_np(double *t, const double *u, int dir, int nx, int ny, int nz, double h, int ncl1,
               int ncln, int npaire, const double *ff, const double *fs, const double *fw)
{
  if (dir < 0 || dir > 2) return -1;
  const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
  if (n < 8 || ncl1 < 1 || ncl1 > 2 || ncln < 1 || ncln > 2) return -1;
  double k[15];
  orc_np_consts(h, k);
  const long stride = dir == 0 ? 1 : dir == 1 ? nx : (long)nx * ny;
  const int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < na; ++a) {
      const size_t q = dir == 0 ? IDX(0, a, b, nx, ny) : dir == 1 ? IDX(a, 0, b, nx, ny) : IDX(a, b, 0, nx, ny);
      for (int i = 0; i < n; ++i) t[q + (size_t)i * stride] = np_rhs(u + q, stride, i, n, ncl1, ncln, npaire, k);
    }
  if (dir == 0) orc_xthomas_12(t, ff, fs, fw, nx, ny, nz);
  else if (dir == 1) orc_ythomas_12(t, ff, fs, fw, nx, ny, nz);
  else orc_zthomas_12(t, ff, fs, fw, nx, ny, nz);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT: a mixed-radix Cooley-Tukey DFT (any n, fast for small prime         */
/* factors) used to restate 2decomp-fft's unnormalised PHYSICAL_IN_Z        */
/* r2c/c2r (call sites src/poisson.f90:318,325,412).  Forward sign -1.      */
/* ------------------------------------------------------------------------ */
typedef double complex cplx;

static void fft_rec(const cplx *in, cplx *out, int n, long stride, const cplx *tw, int N)
{
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (n % p) ++p;
  int m = n / p;
  for (int q = 0; q < p; ++q) fft_rec(in + q * stride, out + q * m, m, stride * p, tw, N);
  cplx tmp[64];
  cplx *t = p <= 64 ? tmp : malloc(sizeof(cplx) * p);
  long step = N / n;
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) t[q] = out[q * m + k] * tw[((long)q * k * step) % N];
    for (int r = 0; r < p; ++r) {
      cplx acc = 0;
      for (int q = 0; q < p; ++q) acc += t[q] * tw[((long)q * r * m * step) % N];
      out[k + r * m] = acc;
    }
  }
  if (t != tmp) free(t);
}

/* Power-of-two n: iterative radix-2 decimation in time on a bit-reversed
 * copy, the same butterflies as fft_rec's p = 2 level without its recursion
 * and index arithmetic. */
static void fft_pow2(const cplx *in, cplx *out, int n, const cplx *tw)
{
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  for (int j = 0; j < n; ++j) {
    unsigned r = 0, v = (unsigned)j;
    for (int b = 0; b < lg; ++b) { r = (r << 1) | (v & 1u); v >>= 1; }
    out[r] = in[j];
  }
  for (int h = 1; h < n; h <<= 1) {
    const int step = n / (2 * h);
    for (int s0 = 0; s0 < n; s0 += 2 * h)
      for (int k = 0; k < h; ++k) {
        cplx e = out[s0 + k], o = out[s0 + k + h] * tw[k * step];
        out[s0 + k] = e + o;
        out[s0 + k + h] = e - o;
      }
  }
}

/* y[k] = sum_j x[j*stride] exp(sign 2 pi i j k / n), result written back with stride */
static void fft_line(cplx *x, int n, long stride, int sign, cplx *work, const cplx *tw)
{
  (void)sign;
  for (int j = 0; j < n; ++j) work[n + j] = x[(long)j * stride];
  if ((n & (n - 1)) == 0) fft_pow2(work + n, work, n, tw);
  else fft_rec(work + n, work, n, 1, tw, n);
  for (int j = 0; j < n; ++j) x[(long)j * stride] = work[j];
}

static cplx *twiddles(int n, int sign)
{
  cplx *tw = malloc(sizeof(cplx) * n);
  for (int t = 0; t < n; ++t) {
    double ang = 2.0 * ORC_PI * (double)t / (double)n;
    tw[t] = cos(ang) + (double)sign * sin(ang) * I;
  }
  return tw;
}

/* r2c, physical data in z-pencil: z first (halved), then y, then x.
 * cw1 layout: (nx, ny, nz/2+1) complex, x fastest.  Lines are independent,
 * so they are spread over threads, each with its own work buffers. */
void orc_fft_r2c_z(const orc_ctx *c, const double *rhs, double *cw1d)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2 + 1;
  cplx *cw1 = (cplx *)cw1d;
  int nmax = nx > ny ? nx : ny; nmax = nmax > nz ? nmax : nz;
  cplx *twz = twiddles(nz, -1), *twy = twiddles(ny, -1), *twx = twiddles(nx, -1);
#pragma omp parallel
  {
    cplx *work = malloc(sizeof(cplx) * 2 * nmax);
    cplx *line = malloc(sizeof(cplx) * nz);
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        for (int k = 0; k < nz; ++k) line[k] = rhs[IDX(i, j, k, nx, ny)];
        fft_line(line, nz, 1, -1, work, twz);
        for (int k = 0; k < nzh; ++k) cw1[IDX(i, j, k, nx, ny)] = line[k];
      }
#pragma omp for schedule(static)
    for (int k = 0; k < nzh; ++k)
      for (int i = 0; i < nx; ++i) fft_line(cw1 + IDX(i, 0, k, nx, ny), ny, nx, -1, work, twy);
#pragma omp for schedule(static)
    for (int k = 0; k < nzh; ++k)
      for (int j = 0; j < ny; ++j) fft_line(cw1 + IDX(0, j, k, nx, ny), nx, 1, -1, work, twx);
    free(work); free(line);
  }
  free(twz); free(twy); free(twx);
}

/* c2r: x, y backward (sign +1), then the Hermitian z line back to real.
 * Destroys cw1. */
void orc_fft_c2r_z(const orc_ctx *c, double *cw1d, double *rhs)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2 + 1;
  cplx *cw1 = (cplx *)cw1d;
  int nmax = nx > ny ? nx : ny; nmax = nmax > nz ? nmax : nz;
  cplx *twz = twiddles(nz, 1), *twy = twiddles(ny, 1), *twx = twiddles(nx, 1);
#pragma omp parallel
  {
    cplx *work = malloc(sizeof(cplx) * 2 * nmax);
    cplx *line = malloc(sizeof(cplx) * nz);
#pragma omp for schedule(static)
    for (int k = 0; k < nzh; ++k)
      for (int j = 0; j < ny; ++j) fft_line(cw1 + IDX(0, j, k, nx, ny), nx, 1, 1, work, twx);
#pragma omp for schedule(static)
    for (int k = 0; k < nzh; ++k)
      for (int i = 0; i < nx; ++i) fft_line(cw1 + IDX(i, 0, k, nx, ny), ny, nx, 1, work, twy);
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        for (int k = 0; k < nzh; ++k) line[k] = cw1[IDX(i, j, k, nx, ny)];
        for (int k = nzh; k < nz; ++k) line[k] = conj(line[nz - k]);
        fft_line(line, nz, 1, 1, work, twz);
        for (int k = 0; k < nz; ++k) rhs[IDX(i, j, k, nx, ny)] = creal(line[k]);
      }
    free(work); free(line);
  }
  free(twz); free(twy); free(twx);
}

/* The spectral kernel of poisson_000, statement by statement
 * (src/poisson.f90:339-396): phase shifts, sign flips, division by
 * -kxyz with the 1e-16 null-mode test, inverse shifts. */
void orc_poisson_spectral(const orc_ctx *c, double *cw1d)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2 + 1;
  const double ntot = (double)nx * (double)ny * (double)nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nzh; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t q = IDX(i, j, k, nx, ny);
        double re = cw1d[2 * q], im = cw1d[2 * q + 1];
        double t1 = re / ntot, t2 = im / ntot, r, s;
        /* forward: z */
        r = t1 * c->bz[k] + t2 * c->az[k]; s = t2 * c->bz[k] - t1 * c->az[k];
        /* y */
        t1 = r; t2 = s;
        r = t1 * c->by[j] + t2 * c->ay[j]; s = t2 * c->by[j] - t1 * c->ay[j];
        if (j + 1 > ny / 2 + 1) { r = -r; s = -s; }
        /* x */
        t1 = r; t2 = s;
        r = t1 * c->bx[i] + t2 * c->ax[i]; s = t2 * c->bx[i] - t1 * c->ax[i];
        if (i + 1 > nx / 2 + 1) { r = -r; s = -s; }
        /* solve */
        t1 = c->kxyz[2 * q]; t2 = c->kxyz[2 * q + 1];
        if (t1 < c->eps || t2 < c->eps) { r = 0.0; s = 0.0; }
        else { r = r / (-t1); s = s / (-t2); }
        /* backward: z */
        t1 = r; t2 = s;
        r = t1 * c->bz[k] - t2 * c->az[k]; s = -t2 * c->bz[k] - t1 * c->az[k];
        /* y */
        t1 = r; t2 = s;
        r = t1 * c->by[j] + t2 * c->ay[j]; s = t2 * c->by[j] - t1 * c->ay[j];
        if (j + 1 > ny / 2 + 1) { r = -r; s = -s; }
        /* x */
        t1 = r; t2 = s;
        r = t1 * c->bx[i] + t2 * c->ax[i]; s = -t2 * c->bx[i] + t1 * c->ax[i];
        if (i + 1 > nx / 2 + 1) { r = -r; s = -s; }
        cw1d[2 * q] = r; cw1d[2 * q + 1] = s;
      }
}

/* poisson_000: src/poisson.f90:276-419 (in place on rhs) */
void orc_poisson(const orc_ctx *c, double *rhs)
{
  size_t ns = (size_t)c->nx * c->ny * (c->nz / 2 + 1);
  double *cw1 = malloc(sizeof(double) * 2 * ns);
  orc_fft_r2c_z(c, rhs, cw1);
  orc_poisson_spectral(c, cw1);
  orc_fft_c2r_z(c, cw1, rhs);
  free(cw1);
}

/* ------------------------------------------------------------------------ */
/* Step composition. At one rank every transpose is an identity copy, so    */
/* pencils share storage here and the copies are left out.                  */
/* ------------------------------------------------------------------------ */
static double *field(const orc_ctx *c) { return malloc(sizeof(double) * (size_t)c->nx * c->ny * c->nz); }

/* momentum_rhs_eq: src/transeq.f90:44-355 */
void orc_momentum_rhs(const orc_ctx *c, double *dux, double *duy, double *duz,
                      const double *ux, const double *uy, const double *uz)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t N = (size_t)nx * ny * nz;
  double *ta = field(c), *td = field(c), *tg = field(c), *t3 = field(c);
  const double *u[3] = {ux, uy, uz};
  double *du[3] = {dux, duy, duz};
  /* per component: which operators the reference uses (:106-107, 126-127,
   * 145-146 for dux; :192-193, 211-212, 231-232 for duy; :278-279,
   * 297-298, 316-317 for duz) as {prod-op, field-op} per direction. */
  static const int opx[3][2] = {{ORC_OP_DERXP, ORC_OP_DERX}, {ORC_OP_DERX, ORC_OP_DERXP}, {ORC_OP_DERX, ORC_OP_DERXP}};
  static const int opy[3][2] = {{ORC_OP_DERY, ORC_OP_DERYP}, {ORC_OP_DERYP, ORC_OP_DERY}, {ORC_OP_DERY, ORC_OP_DERYP}};
  static const int opz[3][2] = {{ORC_OP_DERZ, ORC_OP_DERZP}, {ORC_OP_DERZ, ORC_OP_DERZP}, {ORC_OP_DERZP, ORC_OP_DERZ}};
  for (int cpt = 0; cpt < 3; ++cpt) {
    const double *uc = u[cpt];
    /* x-pencil: ta = ux*uc; td = derx(ta); ta = derx(uc); du = td + ux*ta */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) ta[q] = ux[q] * uc[q];
    orc_apply(c, td, ta, opx[cpt][0], 0, 0, nx, ny, nz);
    orc_apply(c, ta, uc, opx[cpt][1], 0, 0, nx, ny, nz);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) du[cpt][q] = td[q] + ux[q] * ta[q];
    /* y-pencil: td = uc*uy; tg = dery(td); td = dery(uc); tg = tg + uy*td */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) td[q] = cpt == 2 ? uz[q] * uy[q] : uc[q] * uy[q];
    orc_apply(c, tg, td, opy[cpt][0], 1, 0, nx, ny, nz);
    orc_apply(c, td, uc, opy[cpt][1], 1, 0, nx, ny, nz);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) tg[q] = tg[q] + uy[q] * td[q];
    /* z-pencil: td = uc*uz; t3 = derz(td); td = derz(uc); t3 = t3 + uz*td */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) td[q] = uc[q] * uz[q];
    orc_apply(c, t3, td, opz[cpt][0], 2, 0, nx, ny, nz);
    orc_apply(c, td, uc, opz[cpt][1], 2, 0, nx, ny, nz);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) t3[q] = t3[q] + uz[q] * td[q];
    /* z->y: tg = tg + t3; y->x: du = du + tg */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) tg[q] = tg[q] + t3[q];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) du[cpt][q] = du[cpt][q] + tg[q];
  }
  free(ta); free(td); free(tg); free(t3);
}

/* intt: src/time_integrators.f90:20-49 (Euler, var += gdt*dvar) */
void orc_int_time(const orc_ctx *c, double gdt, double *ux, double *uy, double *uz,
                  const double *dux, const double *duy, const double *duz)
{
  const size_t N = (size_t)c->nx * c->ny * c->nz;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) ux[q] = gdt * dux[q] + ux[q];
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) uy[q] = gdt * duy[q] + uy[q];
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) uz[q] = gdt * duz[q] + uz[q];
}

/* divergence: src/navier.f90:98-236 (nlock = 1, test-mode prints apart) */
void orc_divergence(const orc_ctx *c, double *pp3, const double *ux, const double *uy,
                    const double *uz)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t N = (size_t)nx * ny * nz;
  double *pp1 = field(c), *pgy1 = field(c), *pgz1 = field(c);
  double *upi2 = field(c), *duydypi2 = field(c), *po3 = field(c);
  orc_apply(c, pp1, ux, ORC_OP_DERXVP, 0, 1, nx, ny, nz);
  orc_apply(c, pgy1, uy, ORC_OP_INTXVP, 0, 3, nx, ny, nz);
  orc_apply(c, pgz1, uz, ORC_OP_INTXVP, 0, 3, nx, ny, nz);
  orc_apply(c, upi2, pp1, ORC_OP_INTYVP, 1, 3, nx, ny, nz);
  orc_apply(c, duydypi2, pgy1, ORC_OP_DERYVP, 1, 1, nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) duydypi2[q] = duydypi2[q] + upi2[q];
  orc_apply(c, upi2, pgz1, ORC_OP_INTYVP, 1, 3, nx, ny, nz);
  orc_apply(c, pp3, duydypi2, ORC_OP_INTZVP, 2, 3, nx, ny, nz);
  orc_apply(c, po3, upi2, ORC_OP_DERZVP, 2, 1, nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) pp3[q] = pp3[q] + po3[q];
  free(pp1); free(pgy1); free(pgz1); free(upi2); free(duydypi2); free(po3);
}

/* gradp: src/navier.f90:249-359 */
void orc_gradp(const orc_ctx *c, double *px, double *py, double *pz, const double *pp3)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  double *ppi3 = field(c), *pgz3 = field(c), *ppi2 = field(c), *pgy2 = field(c), *pgzi2 = field(c);
  orc_apply(c, ppi3, pp3, ORC_OP_INTZPV, 2, 4, nx, ny, nz);
  orc_apply(c, pgz3, pp3, ORC_OP_DERZPV, 2, 2, nx, ny, nz);
  orc_apply(c, ppi2, ppi3, ORC_OP_INTYPV, 1, 4, nx, ny, nz);
  orc_apply(c, pgy2, ppi3, ORC_OP_DERYPV, 1, 2, nx, ny, nz);
  orc_apply(c, pgzi2, pgz3, ORC_OP_INTYPV, 1, 4, nx, ny, nz);
  orc_apply(c, px, ppi2, ORC_OP_DERXPV, 0, 2, nx, ny, nz);
  orc_apply(c, py, pgy2, ORC_OP_INTXPV, 0, 4, nx, ny, nz);
  orc_apply(c, pz, pgzi2, ORC_OP_INTXPV, 0, 4, nx, ny, nz);
  free(ppi3); free(pgz3); free(ppi2); free(pgy2); free(pgzi2);
}

/* cor_vel: src/navier.f90:64-91 */
void orc_cor_vel(const orc_ctx *c, double *ux, double *uy, double *uz,
                 const double *px, const double *py, const double *pz)
{
  const size_t N = (size_t)c->nx * c->ny * c->nz;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) {
    ux[q] = ux[q] - px[q];
    uy[q] = uy[q] - py[q];
    uz[q] = uz[q] - pz[q];
  }
}

/* solve_poisson: src/navier.f90:21-56 */
void orc_solve_poisson(const orc_ctx *c, double *pp3, double *px, double *py, double *pz,
                       const double *ux, const double *uy, const double *uz)
{
  orc_divergence(c, pp3, ux, uy, uz);
  orc_poisson(c, pp3);
  orc_gradp(c, px, py, pz, pp3);
}

/* one iteration of the step loop: src/xcompact3d.f90:80-95 */
void orc_step(const orc_ctx *c, double gdt, double *ux, double *uy, double *uz,
              double *dux, double *duy, double *duz, double *pp3,
              double *px, double *py, double *pz)
{
  orc_momentum_rhs(c, dux, duy, duz, ux, uy, uz);
  orc_int_time(c, gdt, ux, uy, uz, dux, duy, duz);
  orc_solve_poisson(c, pp3, px, py, pz, ux, uy, uz);
  orc_cor_vel(c, ux, uy, uz, px, py, pz);
}

/* ------------------------------------------------------------------------ */
/* Fields and test mode                                                      */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* counter-based U[0,1): identical integer arithmetic on host and device */
double orc_uniform(uint64_t seed, int comp, uint64_t idx)
{
  uint64_t key = orc_splitmix64(seed ^ ((uint64_t)(comp + 1) * 0x9E3779B97F4A7C15ull));
  return (double)(orc_splitmix64(key + idx) >> 11) * 0x1.0p-53;
}

/* kind 0: the reference's own vel() (src/mom.f90:26-57);
 * kind 1: Taylor-Green with uy = -cos sin cos, uz = 0, plus amp*(2U-1);
 * kind 2: sum of 8 seeded Fourier modes per component, amplitude 1/8 each. */
void orc_init_fields(const orc_ctx *c, int kind, uint64_t seed, double amp,
                     double *ux, double *uy, double *uz)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double twopi = 2.0 * ORC_PI;
  int kv[3][8][3]; double ph[3][8];
  if (kind == 2) {
    for (int cpt = 0; cpt < 3; ++cpt)
      for (int m = 0; m < 8; ++m) {
        for (int d = 0; d < 3; ++d)
          kv[cpt][m][d] = (int)floor(orc_uniform(seed, 16 + cpt, (uint64_t)(4 * m + d)) * 9.0) - 4;
        if (!kv[cpt][m][0] && !kv[cpt][m][1] && !kv[cpt][m][2]) kv[cpt][m][0] = 1;
        ph[cpt][m] = twopi * orc_uniform(seed, 16 + cpt, (uint64_t)(4 * m + 3));
      }
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t q = IDX(i, j, k, nx, ny);
        double x = (double)i * c->dx, y = (double)j * c->dy, z = (double)k * c->dz;
        double sx = sin(twopi * (x / c->xlx)), cx = cos(twopi * (x / c->xlx));
        double sy = sin(twopi * (y / c->yly)), cy = cos(twopi * (y / c->yly));
        double sz = sin(twopi * (z / c->zlz)), cz = cos(twopi * (z / c->zlz));
        if (kind == 0) {
          ux[q] = sx * cy * cz;
          uy[q] = cx * sy * cz;
          uz[q] = -2.0 * cx * cy * sz;
        } else if (kind == 1) {
          ux[q] = sx * cy * cz + amp * (2.0 * orc_uniform(seed, 0, q) - 1.0);
          uy[q] = -cx * sy * cz + amp * (2.0 * orc_uniform(seed, 1, q) - 1.0);
          uz[q] = amp * (2.0 * orc_uniform(seed, 2, q) - 1.0);
        } else {
          double *dst[3] = {ux, uy, uz};
          for (int cpt = 0; cpt < 3; ++cpt) {
            double acc = 0.0;
            for (int m = 0; m < 8; ++m) {
              double arg = twopi * ((double)kv[cpt][m][0] * (double)i / (double)nx +
                                    (double)kv[cpt][m][1] * (double)j / (double)ny +
                                    (double)kv[cpt][m][2] * (double)k / (double)nz) + ph[cpt][m];
              acc += 0.125 * sin(arg);
            }
            dst[cpt][q] = acc;
          }
        }
      }
}

/* test_du / test_dv / test_dw: src/mom.f90:59-243 (RMS error vs the
 * analytic derivative of vel()). which = 0, 1, 2. */
double orc_test_err(const orc_ctx *c, const double *d, int which)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double twopi = 2.0 * ORC_PI;
  double err = 0.0;
  for (int k = 0; k < nz; ++k) {
    double z = (double)k * c->dz;
    for (int j = 0; j < ny; ++j) {
      double y = (double)j * c->dy;
      for (int i = 0; i < nx; ++i) {
        double x = (double)i * c->dx;
        double ana = cos(twopi * (x / c->xlx)) * cos(twopi * (y / c->yly)) * cos(twopi * (z / c->zlz));
        if (which == 0) ana = twopi * ana / c->xlx;
        else if (which == 1) ana = twopi * ana / c->yly;
        else { ana = -2.0 * cos(twopi * (x / c->xlx)) * cos(twopi * (y / c->yly)) * cos(twopi * (z / c->zlz));
               ana = twopi * ana / c->zlz; }
        double e = d[IDX(i, j, k, nx, ny)] - ana;
        err = err + e * e;
      }
    }
  }
  return sqrt(err / nx / ny / nz);
}

/* DIV U* statistics: src/navier.f90:213-231 */
void orc_div_stats(const orc_ctx *c, const double *pp3, double *tmax, double *tmean)
{
  const size_t N = (size_t)c->nxm * c->nym * c->nzm;
  double mx = 0.0, s = 0.0;
  for (size_t q = 0; q < N; ++q) {
    double a = fabs(pp3[q]);
    if (a > mx) mx = a;
    s += a;
  }
  *tmax = mx; *tmean = s / (double)N;
}

/* OpenMP threads used by the line loops (the results do not depend on it) */
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_max_threads(void) { return omp_get_max_threads(); }
