/*
 * x3d_oracle.h -- CPU oracle for the periodic x3div Euler-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithm, used as the checker for the HIP library and as the
 * CPU baseline leg of bench.py.  Nothing in the product (x3div_amd/, include/)
 * may link, import or call it.
 *
 * Array convention: every 3-D field is (nx,ny,nz) with x fastest, exactly the
 * Fortran layout of the reference: element (i,j,k) (0-based) sits at
 * i + nx*(j + ny*k).  Complex arrays are interleaved (re,im) doubles.
 */
#ifndef X3D_ORACLE_H
#define X3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Operator instances, same numbering as the HIP library's X3D_OP_* ids
 * (src/x3d_operator_1d.f90:33-47). */
enum {
  ORC_OP_DERX = 0, ORC_OP_DERXP, ORC_OP_DERY, ORC_OP_DERYP, ORC_OP_DERZ, ORC_OP_DERZP,
  ORC_OP_DERXVP, ORC_OP_DERYVP, ORC_OP_DERZVP,
  ORC_OP_DERXPV, ORC_OP_DERYPV, ORC_OP_DERZPV,
  ORC_OP_INTXVP, ORC_OP_INTYVP, ORC_OP_INTZVP,
  ORC_OP_INTXPV, ORC_OP_INTYPV, ORC_OP_INTZPV,
  ORC_NOPS,
  /* second derivatives (derxx/deryy/derzz, declared and never implemented by
   * the reference, src/x3d_derive.f90:54-59): same numbering as the library */
  ORC_OP_DERXX = ORC_NOPS, ORC_OP_DERYY, ORC_OP_DERZZ,
  ORC_NOPS_ALL
};

typedef struct {
  int n, npaire;
  double alfa;
  double *f, *s, *w, *periodic;  /* length n each */
} orc_op;

typedef struct {
  int nx, ny, nz, nxm, nym, nzm;
  double xlx, yly, zlz, dx, dy, dz;
  /* scheme constants per direction d = 0 (x), 1 (y), 2 (z) */
  double alfai[3], afi[3], bfi[3];                       /* 6th-order compact D   */
  double alcai[3], aci[3], bci[3];                       /* staggered Ds          */
  double ailcai[3], aici[3], bici[3], cici[3], dici[3];  /* mid-point interp. I   */
  double alsai[3], asi[3], bsi[3];                       /* compact 2nd deriv.    */
  orc_op ops[ORC_NOPS_ALL];
  /* Poisson tables (src/poisson.f90:421-479, 483-767) */
  double *ax, *bx, *ay, *by, *az, *bz;
  double *xk2, *yk2, *zk2, *exs, *eys, *ezs;  /* real parts (imag parts equal) */
  double *kxyz;   /* complex nx*ny*(nz/2+1), interleaved */
  double eps;     /* 1e-16, src/poisson.f90:24-28 */
  /* y-stretching: istret and ppy, ppyi (ny each), src/module_param.f90:70-71,104 */
  int istret;
  double *ppy, *ppyi;
} orc_ctx;

/* ---- setup --------------------------------------------------------------- */
int  orc_init(orc_ctx *c, int nx, int ny, int nz, double xlx, double yly, double zlz);
void orc_free(orc_ctx *c);
orc_ctx *orc_new(int nx, int ny, int nz, double xlx, double yly, double zlz);
void orc_delete(orc_ctx *c);
void orc_prepare(const double *b, const double *cdiag, const double *f,
                 double *s, double *w, int n);
void orc_thomas1d(double *tt, const double *ff, const double *fs, const double *fw, int n);
int  orc_get_op(const orc_ctx *c, int op, double *f, double *s, double *w,
                double *periodic, double *alfa, int *n, int *npaire);
void orc_get_consts(const orc_ctx *c, double *out /* 3 x 11 */);
/* real and imaginary parts of kxyz, each (nx, ny, nz/2+1) */
void orc_get_kxyz(const orc_ctx *c, double *re, double *im);
/* ax, bx, ay, by, az, bz concatenated */
void orc_get_ab(const orc_ctx *c, double *out);

/* ---- batched Thomas (in place) -------------------------------------------- */
void orc_xthomas_0(double *tt, const orc_op *op, int nx, int ny, int nz);
void orc_ythomas_0(double *tt, const orc_op *op, int nx, int ny, int nz);
void orc_zthomas_0(double *tt, const orc_op *op, int nx, int ny, int nz);
void orc_xthomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz);
void orc_ythomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz);
void orc_zthomas_12(double *tt, const double *ff, const double *fs, const double *fw,
                    int nx, int ny, int nz);
/* generic entry by operator id: dir = 0,1,2 */
int  orc_thomas(const orc_ctx *c, double *tt, int op, int dir, int nx, int ny, int nz);

/* ---- compact operators ---------------------------------------------------- */
/* kind: 0 = D (der?_00), 1 = Ds v->p, 2 = Ds p->v, 3 = I v->p, 4 = I p->v     */
int  orc_apply(const orc_ctx *c, double *t, const double *u, int op, int dir,
               int kind, int nx, int ny, int nz);
/* istret /= 0: y-derivatives are multiplied by ppy(j) (dery, derypv) or
 * ppyi(j) (deryvp) after the solve; vectors of ny values, copied. */
int  orc_set_stretching(orc_ctx *c, int istret, const double *ppy, const double *ppyi);
void orc_derx(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz);
void orc_dery(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz);
void orc_derz(const orc_ctx *c, double *t, const double *u, int op, int nx, int ny, int nz);

/* ---- non-periodic first derivatives (der?_11/12/21/22) --------------------- */
/* boundary constants alfai afi bfi alfa1 af1 bf1 cf1 alfa2 af2 alfan afn bfn cfn
 * alfam afm for spacing h */
void orc_np_consts(double h, double *k);
int  orc_np_tables(int n, double h, int ncl1, int ncln, double *ff, double *fs, double *fw,
                   double *ffp, double *fsp, double *fwp, double *rows, double *rowsp);
int  orc_der_np(double *t, const double *u, int dir, int nx, int ny, int nz, double h, int ncl1,
                int ncln, int npaire, const double *ff, const double *fs, const double *fw);

/* ---- Poisson -------------------------------------------------------------- */
void orc_fft_r2c_z(const orc_ctx *c, const double *rhs, double *cw1);
void orc_fft_c2r_z(const orc_ctx *c, double *cw1, double *rhs);
void orc_poisson_spectral(const orc_ctx *c, double *cw1);
void orc_poisson(const orc_ctx *c, double *rhs);

/* ---- step composition ----------------------------------------------------- */
void orc_momentum_rhs(const orc_ctx *c, double *dux, double *duy, double *duz,
                      const double *ux, const double *uy, const double *uz);
void orc_int_time(const orc_ctx *c, double gdt, double *ux, double *uy, double *uz,
                  const double *dux, const double *duy, const double *duz);
void orc_divergence(const orc_ctx *c, double *pp3, const double *ux, const double *uy,
                    const double *uz);
void orc_gradp(const orc_ctx *c, double *px, double *py, double *pz, const double *pp3);
void orc_cor_vel(const orc_ctx *c, double *ux, double *uy, double *uz,
                 const double *px, const double *py, const double *pz);
void orc_solve_poisson(const orc_ctx *c, double *pp3, double *px, double *py, double *pz,
                       const double *ux, const double *uy, const double *uz);
void orc_step(const orc_ctx *c, double gdt, double *ux, double *uy, double *uz,
              double *dux, double *duy, double *duz, double *pp3,
              double *px, double *py, double *pz);

/* ---- fields and test-mode ------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x);
double   orc_uniform(uint64_t seed, int comp, uint64_t idx);
void orc_init_fields(const orc_ctx *c, int kind, uint64_t seed, double amp,
                     double *ux, double *uy, double *uz);
double orc_test_err(const orc_ctx *c, const double *d, int which);
void   orc_div_stats(const orc_ctx *c, const double *pp3, double *tmax, double *tmean);

/* ---- threading ------------------------------------------------------------ */
void orc_set_threads(int n);
int  orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
