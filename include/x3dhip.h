/*
 * x3dhip.h -- C ABI of libx3dhip, the MI355X (gfx950) implementation of
 * x3div's periodic Euler-timestep hot path.
 *
 * Conventions
 *  - Every field is a device pointer to an (n1,n2,n3) array with the first
 *    index fastest, i.e. the Fortran layout the reference uses in every
 *    pencil.  At one rank the x-, y- and z-pencils are the same array.
 *  - Calls are ordered on the context's HIP stream (x3d_set_stream); only
 *    calls that return host values (reductions, x3d_sync) synchronise.
 *  - Every entry point returns 0 (X3D_OK) or a negative X3D_ERR_* code; the
 *    Fortran binding turns a nonzero code into the decomp_2d_abort path the
 *    reference uses for fatal errors (src/x3d_tools.f90:70-74).
 *  - Outputs are fully overwritten and inputs are not modified (except the
 *    in-place Thomas and Poisson calls), as in the reference.
 *  - Boundary conditions are triply periodic (ncl?1 = ncl?n = 0), the only
 *    case the reference's Poisson solver accepts (src/poisson.f90:104-108),
 *    except on a context made by x3d_create_bc, whose non-periodic
 *    directions take the first derivatives der?_11/12/21/22 only.
 *  - Suffix _f32 marks the precision form: FP32 fields, FP32 line solves.
 *
 * Reference citations are relative to the x3div source tree (src/...).
 */
#ifndef X3DHIP_H
#define X3DHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define X3D_OK 0
#define X3D_ERR_ARG (-1)         /* bad pointer, id or enum value            */
#define X3D_ERR_SIZE (-2)        /* sizes do not match the context / n < 8   */
#define X3D_ERR_HIP (-3)         /* a HIP runtime call failed                */
#define X3D_ERR_FFT (-4)         /* a hipFFT call failed                     */
#define X3D_ERR_ALLOC (-5)       /* device allocation failed                 */
#define X3D_ERR_UNSUPPORTED (-6) /* line length not factorable for the kernels, or a
                                    field of the requested precision of 4 GiB or more */

/* Operator instances (src/x3d_operator_1d.f90:33-47), in this order. */
enum {
  X3D_OP_DERX = 0, X3D_OP_DERXP, X3D_OP_DERY, X3D_OP_DERYP, X3D_OP_DERZ, X3D_OP_DERZP,
  X3D_OP_DERXVP, X3D_OP_DERYVP, X3D_OP_DERZVP,
  X3D_OP_DERXPV, X3D_OP_DERYPV, X3D_OP_DERZPV,
  X3D_OP_INTXVP, X3D_OP_INTYVP, X3D_OP_INTZVP,
  X3D_OP_INTXPV, X3D_OP_INTYPV, X3D_OP_INTZPV,
  X3D_NOPS,
  /* Second derivatives of derxx / deryy / derzz (procedure pointers declared
   * at src/x3d_derive.f90:54-59 and never associated there): this build's
   * own operator instances, after the reference's 18. */
  X3D_OP_DERXX = X3D_NOPS, X3D_OP_DERYY, X3D_OP_DERZZ,
  X3D_NOPS_ALL
};

/* Initial-field kinds for x3d_init_fields. */
enum {
  X3D_FIELD_REFERENCE_VEL = 0, /* u=sin cos cos, v=cos sin cos, w=-2 cos cos sin (src/mom.f90:47-52) */
  X3D_FIELD_TG_NOISE = 1,      /* u=sin cos cos, v=-cos sin cos, w=0, each + amp*(2U-1) */
  X3D_FIELD_FOURIER8 = 2       /* 8 seeded Fourier modes per component, amplitude 1/8 each */
};

typedef struct x3d_ctx x3d_ctx;

/* ---- library / context -------------------------------------------------- */
int         x3d_version(void);
const char *x3d_strerror(int code);

/* Grid set-up: parameter() with periodic defaults (src/parameters.f90:24-90),
 * x3d_operator_{x,y,z}_data_init + x3d_operator_1d_init
 * (src/x3d_tools.f90:178-182, src/x3d_operator_1d.f90:59-195),
 * x3d_derive_init (src/x3d_derive.f90:66-93) and decomp_2d_poisson_init
 * (src/poisson.f90:73-231), on the current HIP device. */
int x3d_create(int nx, int ny, int nz, double xlx, double yly, double zlz, x3d_ctx **ctx);
/* The same with the boundary codes of parameter() (nclx1/nclxn, ncly1/nclyn,
 * nclz1/nclzn, src/module_param.f90:88; src/parameters.f90): per direction
 * 0 periodic at both ends, or 1 (free-slip) / 2 (Dirichlet) at each end, and
 * dX = L / (nX - 1) in a non-periodic direction.  NULL arrays mean periodic.
 * In a non-periodic direction the first-derivative instances (derX: npaire
 * 0, derXp: npaire 1) get first_deriv_imp_'s non-periodic tables
 * (src/x3d_operator_1d.f90:395-491) and x3d_der{x,y,z} runs der?_11/12/21/22
 * (src/x3d_derive.f90:156-275, 343-499, 562-725); the Thomas entry points run
 * the _12 sweeps for them.  That direction's staggered, interpolation and
 * second-derivative instances, and on such a grid the Poisson solver, the
 * divergence, gradp, the momentum right-hand side and the step, return
 * X3D_ERR_UNSUPPORTED (the reference's Poisson stops there,
 * src/poisson.f90:104-108). */
int x3d_create_bc(int nx, int ny, int nz, double xlx, double yly, double zlz,
                  const int *ncl1, const int *ncln, x3d_ctx **ctx);
/* Grid stretching in y: istret and the vectors ppy, ppyi of module param
 * (src/module_param.f90:70-71,104; allocated at src/var.f90:138-140).  With
 * istret /= 0 the reference multiplies the results of dery by ppy(j)
 * (src/x3d_derive.f90:331-337, 427-433), of deryvp by ppyi(j) and of derypv
 * by ppy(j) (src/x3d_staggered.f90:628-634, 874-880).  ppy and ppyi are host
 * vectors of ny values each, copied to the device here; istret = 0 turns the
 * scaling off (the vectors may then be NULL).  x3d_dery, x3d_deryvp,
 * x3d_derypv and x3d_dery_f32 scale by the vector they are given (one that
 * differs from the context's replaces it; NULL keeps the context's);
 * x3d_apply_f32, x3d_apply_bfp, the steps and their pieces scale by the
 * context's.  Interpolations and Poisson are unchanged (the reference's
 * kxyz uses yk2, which istret leaves alone, src/poisson.f90:562-565).  On a
 * stretched context the second derivative in y (a stretched grid needs
 * pp2y/pp4y), the matrix-core y-derivatives and batches return
 * X3D_ERR_UNSUPPORTED.  A captured step is released.  The reference's x3div
 * driver fixes istret = 0 (src/parameters.f90:117). */
int x3d_set_stretching(x3d_ctx *ctx, int istret, const double *ppy, const double *ppyi);
/* x3d_operator_1d_finalize + decomp_2d_poisson_finalize (src/poisson.f90:237-271) */
int x3d_destroy(x3d_ctx *ctx);
int x3d_set_stream(x3d_ctx *ctx, void *hip_stream);
int x3d_sync(x3d_ctx *ctx);
/* MPI_Wtime replacement for the step timer (src/xcompact3d.f90:78,98) */
double x3d_wtime(void);
int x3d_get_sizes(const x3d_ctx *ctx, int *nx, int *ny, int *nz);

/* Device field memory owned by the library; the Fortran side maps it with
 * c_f_pointer (the reference's module allocatables, src/var.f90:39-201). */
int x3d_field_alloc(x3d_ctx *ctx, size_t nbytes, void **dptr);
int x3d_field_free(x3d_ctx *ctx, void *dptr);
int x3d_memcpy_h2d(x3d_ctx *ctx, void *dst, const void *src, size_t nbytes);
int x3d_memcpy_d2h(x3d_ctx *ctx, void *dst, const void *src, size_t nbytes);
/* n FP64 values of a device array set to v (a kernel, not a memset, so it
 * can be recorded in a graph); the pressure and gradient zeroing of case's
 * init (src/case.f90:36-44). */
int x3d_fill(x3d_ctx *ctx, double *p, size_t n, double v);

/* ---- operator tables (type x3doperator1d, src/x3d_operator_1d.f90:12-30) -- */
/* Host copies of one operator's f, s, w, periodic vectors (length n) and alfa. */
int x3d_operator_get(const x3d_ctx *ctx, int op, double *f, double *s, double *w,
                     double *periodic, double *alfa, int *n, int *npaire);
/* Replace one operator's tables (drop-in for x3d_operator_1d%init,
 * src/x3d_operator_1d.f90:234-266, when the caller builds its own).  A step
 * captured earlier (x3d_graph_capture_step*) is released: x3d_graph_launch
 * then returns X3D_ERR_UNSUPPORTED, as does a new capture, since steps on
 * caller tables run the unfused composition, which is not captured. */
int x3d_operator_set(x3d_ctx *ctx, int op, const double *f, const double *s, const double *w,
                     const double *periodic, double alfa);
/* Scheme constants per direction, 11 each: alfai afi bfi alcai aci bci ailcai aici bici cici dici
 * (module x3d_operator_{x,y,z}_data, src/x3d_operator_x_data.f90:15-18). */
int x3d_scheme_consts(const x3d_ctx *ctx, double *out33);

/* ---- batched Thomas, in place (src/thomas.f90:15-28 generics) ------------ */
/* xthomas_0 / ythomas_0 / zthomas_0 (src/thomas.f90:34-58, 87-111, 139-163):
 * forward/backward sweeps with op's f,s,w then the periodic correction. */
int x3d_xthomas(x3d_ctx *ctx, double *tt, int op, int nx, int ny, int nz);
int x3d_ythomas(x3d_ctx *ctx, double *tt, int op, int nx, int ny, int nz);
int x3d_zthomas(x3d_ctx *ctx, double *tt, int op, int nx, int ny, int nz);
/* The same generics called with the caller's own coefficient vectors, the
 * reference's argument lists: xthomas_0(tt, ff, fs, fw, perio, alfa, ...) and,
 * with perio = NULL, xthomas_12(tt, ff, fs, fw, ...) (src/thomas.f90:34-84,
 * 87-136, 139-188).  ff, fs, fw, perio are host arrays of the line length,
 * copied to the device for the call (the call waits for that copy). */
int x3d_xthomas_tab(x3d_ctx *ctx, double *tt, const double *ff, const double *fs, const double *fw,
                    const double *perio, double alfa, int nx, int ny, int nz);
int x3d_ythomas_tab(x3d_ctx *ctx, double *tt, const double *ff, const double *fs, const double *fw,
                    const double *perio, double alfa, int nx, int ny, int nz);
int x3d_zthomas_tab(x3d_ctx *ctx, double *tt, const double *ff, const double *fs, const double *fw,
                    const double *perio, double alfa, int nx, int ny, int nz);
/* thomas1d (src/thomas.f90:200-218): one tridiagonal line on the host, in
 * place, as x3doperator1d%init uses it for the periodic vector. */
int x3d_thomas1d(double *tt, const double *ff, const double *fs, const double *fw, int n);

/* ---- collocated first derivatives (procedure pointers derx/dery/derz,
 *      abstract interfaces src/x3d_derive.f90:18-51; periodic derx_00,
 *      dery_00, derz_00 at :111-152, :279-339, :503-558) ------------------ */
int x3d_derx(x3d_ctx *ctx, double *tx, const double *ux, int op, int nx, int ny, int nz);
/* ppy: the stretching vector of DERIVATIVE_Y (host, ny values), used when
 * the context is stretched (x3d_set_stretching); may be NULL. */
int x3d_dery(x3d_ctx *ctx, double *ty, const double *uy, int op, const double *ppy,
             int nx, int ny, int nz);
int x3d_derz(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nz);

/* ---- second derivatives (procedure pointers derxx/deryy/derzz with the
 *      DERIVATIVE_X / DERIVATIVE_YY / DERIVATIVE_Z interfaces,
 *      src/x3d_derive.f90:36-59; the reference declares them and keeps the
 *      coefficient names alsaix, asix, bsix, src/x3d_operator_x_data.f90:24,
 *      but implements no scheme).  This build's scheme is the sixth-order
 *      compact one of isecondder = 4 (src/parameters.f90:107; Lele 1992):
 *      alfa t_{i-1} + t_i + alfa t_{i+1}
 *        = a (u_{i+1} - 2u_i + u_{i-1}) / h^2 + b (u_{i+2} - 2u_i + u_{i-2}) / (4 h^2),
 *      alfa = 2/11, a = 12/11, b = 3/11, periodic; op is X3D_OP_DERXX/YY/ZZ. */
int x3d_derxx(x3d_ctx *ctx, double *tx, const double *ux, int op, int nx, int ny, int nz);
int x3d_deryy(x3d_ctx *ctx, double *ty, const double *uy, int op, int nx, int ny, int nz);
int x3d_derzz(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nz);

/* ---- staggered derivatives and interpolations (src/x3d_staggered.f90) ---- */
int x3d_derxvp(x3d_ctx *ctx, double *tx, const double *ux, int op, int nx, int nxm, int ny, int nz);    /* :23-106  */
int x3d_interxvp(x3d_ctx *ctx, double *tx, const double *ux, int op, int nx, int nxm, int ny, int nz);  /* :110-217 */
int x3d_derxpv(x3d_ctx *ctx, double *tx, const double *ux, int op, int nxm, int nx, int ny, int nz);    /* :221-285 */
int x3d_interxpv(x3d_ctx *ctx, double *tx, const double *ux, int op, int nxm, int nx, int ny, int nz);  /* :289-404 */
int x3d_interyvp(x3d_ctx *ctx, double *ty, const double *uy, int op, int nx, int ny, int nym, int nz);  /* :408-541 */
int x3d_deryvp(x3d_ctx *ctx, double *ty, const double *uy, int op, const double *ppyi,
               int nx, int ny, int nym, int nz);                                                       /* :545-637 */
int x3d_interypv(x3d_ctx *ctx, double *ty, const double *uy, int op, int nx, int nym, int ny, int nz);  /* :641-786 */
int x3d_derypv(x3d_ctx *ctx, double *ty, const double *uy, int op, const double *ppy,
               int nx, int nym, int ny, int nz);                                                       /* :790-883 */
int x3d_derzvp(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nz, int nzm);    /* :887-997 */
int x3d_interzvp(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nz, int nzm);  /* :1001-1139 */
int x3d_derzpv(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nzm, int nz);    /* :1143-1229 */
int x3d_interzpv(x3d_ctx *ctx, double *tz, const double *uz, int op, int nx, int ny, int nzm, int nz);  /* :1233-1383 */

/* ---- pencil transposes (src/x3d_transpose.f90:16-47 generics; 2decomp's
 *      transpose_* called directly at src/navier.f90:171-173,189-190).
 *      At one rank every pencil holds the whole array in the same order, so
 *      each is a device copy of n1*n2*n3 elements of elem_bytes (8 real,
 *      16 complex FP64); in == out is a no-op. ---------------------------- */
int x3d_transpose_x_to_y(x3d_ctx *ctx, const void *in, void *out, size_t elem_bytes, int n1, int n2, int n3);
int x3d_transpose_y_to_z(x3d_ctx *ctx, const void *in, void *out, size_t elem_bytes, int n1, int n2, int n3);
int x3d_transpose_z_to_y(x3d_ctx *ctx, const void *in, void *out, size_t elem_bytes, int n1, int n2, int n3);
int x3d_transpose_y_to_x(x3d_ctx *ctx, const void *in, void *out, size_t elem_bytes, int n1, int n2, int n3);

/* ---- spectral Poisson (procedure pointer poisson => poisson_000,
 *      src/poisson.f90:57-63, 276-419): r2c FFT, division by the modified
 *      wavenumber with null modes zeroed, c2r FFT; in place on rhs. -------- */
int x3d_poisson(x3d_ctx *ctx, double *rhs);
/* Host copies of the public module data ax,bx,ay,by,az,bz (src/poisson.f90:40-42)
 * and of kxyz's real part on the reference's (nx, ny, nz/2+1) spectral grid
 * (src/poisson.f90:45, 660-699).  Any pointer may be NULL. */
int x3d_poisson_tables(const x3d_ctx *ctx, double *ax, double *bx, double *ay, double *by,
                       double *az, double *bz, double *kxyz_re);

/* ---- timestep pieces (callers of the operators) ----------------------- */
/* calculate_transeq_rhs / momentum_rhs_eq (src/transeq.f90:18-355) */
int x3d_momentum_rhs(x3d_ctx *ctx, double *dux, double *duy, double *duz,
                     const double *ux, const double *uy, const double *uz);
/* int_time / intt, Euler: u += gdt*du (src/time_integrators.f90:20-106) */
int x3d_int_time(x3d_ctx *ctx, double gdt, double *ux, double *uy, double *uz,
                 const double *dux, const double *duy, const double *duz);
/* divergence, nlock = 1 (src/navier.f90:98-236) */
int x3d_divergence(x3d_ctx *ctx, double *pp3, const double *ux, const double *uy, const double *uz);
/* gradp (src/navier.f90:249-359) */
int x3d_gradp(x3d_ctx *ctx, double *px, double *py, double *pz, const double *pp3);
/* cor_vel (src/navier.f90:64-91) */
int x3d_cor_vel(x3d_ctx *ctx, double *ux, double *uy, double *uz,
                const double *px, const double *py, const double *pz);
/* solve_poisson = divergence + poisson + gradp (src/navier.f90:21-56) */
int x3d_solve_poisson(x3d_ctx *ctx, double *pp3, double *px, double *py, double *pz,
                      const double *ux, const double *uy, const double *uz);
/* One iteration of the step loop (src/xcompact3d.f90:80-95): transeq, Euler
 * update, solve_poisson, cor_vel, with the elementwise glue fused into the
 * line kernels.  Same results as the four calls above in sequence. */
int x3d_step(x3d_ctx *ctx, double gdt, double *ux, double *uy, double *uz,
             double *dux, double *duy, double *duz, double *pp3,
             double *px, double *py, double *pz);

/* Per-pass timing of one step, the role of the reference's NVTX ranges
 * around each call of the step loop (src/xcompact3d.f90:79-93).  Runs
 * x3d_step with a HIP event after each of its X3D_NPASS passes on the
 * context's stream, waits for it, and writes each pass's milliseconds to
 * pass_ms[X3D_NPASS].  Pass i is named by x3d_pass_name(i); its algorithmic
 * HBM bytes per grid point (every field it reads or writes touched once) are
 * x3d_pass_bytes_per_point(i), in FP64.  The transeq passes run z, y, x: z
 * reads u and writes its part (6 fields), x adds the sum and the Euler update
 * (12 fields).  The FP32 form's single grids run them x, y, z: there
 * transeq_x moves 6 fields and transeq_z 12 (at 4 B each). */
#define X3D_NPASS 10
int x3d_step_timed(x3d_ctx *ctx, double gdt, double *ux, double *uy, double *uz,
                   double *dux, double *duy, double *duz, double *pp3,
                   double *px, double *py, double *pz, float *pass_ms);
const char *x3d_pass_name(int pass);
double x3d_pass_bytes_per_point(int pass);

/* ---- precision form: FP32 fields, FP32 line solves ------------------------
 * The same operators and step as above on FP32 arrays.  The line solves run
 * in FP32 (circulant solver on the library's periodic tables; an operator
 * replaced with x3d_operator_set gives X3D_ERR_UNSUPPORTED).  Poisson runs
 * FP32 FFTs around an FP64 modified-wavenumber division.  Results differ from
 * the FP64 entry points by FP32 rounding (about 1e-7 relative for one
 * derivative). */
int x3d_derx_f32(x3d_ctx *ctx, float *tx, const float *ux, int op, int nx, int ny, int nz);
int x3d_dery_f32(x3d_ctx *ctx, float *ty, const float *uy, int op, const float *ppy, int nx, int ny, int nz);
int x3d_derz_f32(x3d_ctx *ctx, float *tz, const float *uz, int op, int nx, int ny, int nz);
/* Any operator along its own direction: derivatives (op 0-5), the staggered
 * derivatives and interpolations (op 6-17), as the FP64 routines of
 * src/x3d_derive.f90 and src/x3d_staggered.f90 with periodic BCs, and the
 * second derivatives (X3D_OP_DERXX..ZZ). */
int x3d_apply_f32(x3d_ctx *ctx, int op, float *out, const float *in);
int x3d_poisson_f32(x3d_ctx *ctx, float *rhs);
/* Matrix-core variant of x3d_apply_f32: with periodic boundaries each
 * operator is the dense circulant C = A^-1 B (built once per operator in
 * FP64 from the reference's own periodic solve), applied to every line as a
 * GEMM on v_mfma_f32_32x32x16_f16 in three split-precision terms
 * (C_hi u_hi + C_hi u_lo + C_lo u_hi, fp16 halves after exact power-of-two
 * scaling, FP32 accumulation); a few 1e-6 relative to the FP64 operator.  Needs the line length a
 * multiple of 128 (at most 1024) and the line count a multiple of 32;
 * out != in. */
int x3d_apply_circ_mfma(x3d_ctx *ctx, int op, float *out, const float *in);
int x3d_step_f32(x3d_ctx *ctx, double gdt, float *ux, float *uy, float *uz,
                 float *dux, float *duy, float *duz, float *pp3,
                 float *px, float *py, float *pz);
int x3d_step_timed_f32(x3d_ctx *ctx, double gdt, float *ux, float *uy, float *uz,
                       float *dux, float *duy, float *duz, float *pp3,
                       float *px, float *py, float *pz, float *pass_ms);

/* ---- BFP32 form: block-floating-point fields, FP64 line solves -------------
 * A BFP32 field is an int32 array of the grid's shape (4 bytes per point, as
 * FP32).  x is cut into 16-point segments (nx must be a multiple of 16; the
 * field below 4 GiB); the values of a segment share the exponent e, the
 * smallest with |v| < 2^e over the segment, clamped to [-126, 127].  Word k
 * of a segment is (q << 1) | b with q = rint(v 2^(30-e)) clamped to
 * [-2^30, 2^30 - 1] (so v = q 2^(e-30), within 2^-29 of the segment's largest
 * value, 2^-30 but for values within half a unit of 2^e) and b = bit k of
 * e + 128 for k < 8, 0 for k >= 8.  The operators and the step read and write
 * BFP32 and compute in FP64, so they differ from the FP64 entry points on the
 * decoded input by the rounding of their output; against the FP64 entry
 * points on the unrounded field the rounding of the input is amplified as in
 * the FP32 form, from a 64-128 times smaller start.  These entry points have
 * no counterpart in the reference (real(mytype) fields, src/decomp_2d.f90);
 * they give callers the FP32 form's traffic without FP32's rounding. */
int x3d_bfp_encode(x3d_ctx *ctx, int32_t *q, const double *u);
int x3d_bfp_decode(x3d_ctx *ctx, double *u, const int32_t *q);
/* any operator along its own direction, ids as x3d_apply_f32; out != in */
int x3d_apply_bfp(x3d_ctx *ctx, int op, int32_t *out, const int32_t *in);
/* the step of x3d_step on BFP32 fields (Poisson in FP64 on the decoded
 * right-hand side, the pressure stored back as BFP32) */
int x3d_step_bfp(x3d_ctx *ctx, double gdt, int32_t *ux, int32_t *uy, int32_t *uz,
                 int32_t *dux, int32_t *duy, int32_t *duz, int32_t *pp3,
                 int32_t *px, int32_t *py, int32_t *pz);
int x3d_step_timed_bfp(x3d_ctx *ctx, double gdt, int32_t *ux, int32_t *uy, int32_t *uz,
                       int32_t *dux, int32_t *duy, int32_t *duz, int32_t *pp3,
                       int32_t *px, int32_t *py, int32_t *pz, float *pass_ms);

/* ---- captured step --------------------------------------------------------
 * Record one x3d_step (or x3d_step_f32) with these arrays into a HIP graph
 * held by the context (replacing any earlier one); x3d_graph_launch replays
 * it on the context's stream.  Scratch memory and FFT plans are allocated
 * before the capture, so replays allocate nothing.  The reference drives the
 * step loop from the host, one call at a time (src/xcompact3d.f90:74-104). */
int x3d_graph_capture_step(x3d_ctx *ctx, double gdt, double *ux, double *uy, double *uz,
                           double *dux, double *duy, double *duz, double *pp3,
                           double *px, double *py, double *pz);
int x3d_graph_capture_step_f32(x3d_ctx *ctx, double gdt, float *ux, float *uy, float *uz,
                               float *dux, float *duy, float *duz, float *pp3,
                               float *px, float *py, float *pz);
int x3d_graph_launch(x3d_ctx *ctx);
/* Record any sequence of this context's calls into its graph instead: the
 * calls between x3d_graph_begin and x3d_graph_end are recorded, not run
 * (the reference-shaped driver records its transeq, int_time, solve_poisson
 * and cor_vel calls, src/xcompact3d.f90:80-100, once, then replays them per
 * step).  The FP64 step's scratch fields are allocated by x3d_graph_begin;
 * recorded calls must not synchronise (reductions, x3d_sync), and a call
 * that fails makes x3d_graph_end fail and discards the graph. */
int x3d_graph_begin(x3d_ctx *ctx);
int x3d_graph_end(x3d_ctx *ctx);

/* ---- the step without allocation ------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  x3d_step, bit for bit, with its 8 scratch fields taken
 * from a static device array of the library (X3D_NOALLOC_MAX_POINTS * 8
 * doubles, 128 MiB, which every device the library runs on holds until the
 * process ends) instead of the context's own work fields, which x3d_step
 * allocates on its first call.  It calls no hipMalloc or hipFree: the FFT
 * plans, spectrum and tables it uses are made by x3d_create.  Grids of more
 * than X3D_NOALLOC_MAX_POINTS points give X3D_ERR_SIZE.  All calls on one
 * device share the array and it keeps the last call's contents: calls that
 * overlap on one device (two streams, threads or contexts) race. */
#define X3D_NOALLOC_MAX_POINTS (1 << 21) /* 128^3 */
int x3d_step_noalloc(x3d_ctx *ctx, double gdt, double *ux, double *uy, double *uz,
                     double *dux, double *duy, double *duz, double *pp3,
                     double *px, double *py, double *pz);

/* ---- batching form ---------------------------------------------------------
 * Many independent triply periodic grids of uneven sizes, each the
 * reference's problem on its own (nx, ny, nz) (src/xcompact3d.f90:80-95),
 * stepped together.  The batch owns the fields: per grid the 10 arrays
 * ux uy uz dux duy duz pp3 px py pz (field index 0-9), of FP64 (precision
 * 8) or FP32 (precision 4) values.  Each step pass is one persistent launch
 * per group of grids sharing a line tiling; Poisson is one launch per stage
 * and line length over the grids with lines of up to 384 points (the others
 * solve one by one), and the launches of one pass or stage run side by side
 * on side streams (X3D_BATCH_PSTREAMS, X3D_BATCH_OVERLAP=0).  A grid's
 * own context (x3d_batch_context) serves the single-grid entry points; while
 * any grid's context holds y-stretching (x3d_set_stretching) or operator
 * tables of the caller's own (x3d_operator_set), x3d_batch_step and
 * x3d_batch_graph_launch return X3D_ERR_UNSUPPORTED (the batched passes run
 * the circulant solvers of the library's tables, unstretched).
 * dims: nx, ny, nz per grid; lengths: xlx, yly, zlz per grid, or NULL for
 * 2*pi.  Seeds of x3d_batch_init_fields are seed + grid index. */
typedef struct x3d_batch x3d_batch;
int x3d_batch_create(int ngrids, const int *dims, const double *lengths, int precision, x3d_batch **batch);
int x3d_batch_destroy(x3d_batch *batch);
int x3d_batch_set_stream(x3d_batch *batch, void *hip_stream);
int x3d_batch_sizes(const x3d_batch *batch, int *ngrids, long long *total_points);
int x3d_batch_field(const x3d_batch *batch, int grid, int field, void **dptr);
int x3d_batch_context(const x3d_batch *batch, int grid, x3d_ctx **ctx);
int x3d_batch_init_fields(x3d_batch *batch, int kind, uint64_t seed, double amp);
int x3d_batch_step(x3d_batch *batch, double gdt);
/* Record one batched step into a HIP graph (gdt fixed at capture); replay it. */
int x3d_batch_graph_capture(x3d_batch *batch, double gdt);
int x3d_batch_graph_launch(x3d_batch *batch);
/* Time the batched step (reps steps per trial) for every tiling that fits
 * each direction's line lengths, keep the fastest, and write a text report. */
int x3d_batch_autotune(x3d_batch *batch, int reps, char *report, size_t report_len);

/* ---- test mode and fields --------------------------------------------- */
/* test_du/test_dv/test_dw RMS error vs the analytic derivative of vel()
 * (src/mom.f90:59-243); which = 0,1,2.  Synchronises. */
int x3d_test_err(x3d_ctx *ctx, const double *d, int which, double *err);
/* "DIV U* max mean" statistics (src/navier.f90:213-231).  Synchronises. */
int x3d_div_stats(x3d_ctx *ctx, const double *pp3, double *tmax, double *tmean);
/* Initial fields on the device (case::init -> vel, src/case.f90:20-58 and
 * src/mom.f90:26-57 for kind 0; seeded benchmark fields for kinds 1, 2). */
int x3d_init_fields(x3d_ctx *ctx, int kind, uint64_t seed, double amp,
                    double *ux, double *uy, double *uz);

#ifdef __cplusplus
}
#endif
#endif /* X3DHIP_H */
