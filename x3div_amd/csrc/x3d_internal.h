// Internal definitions of libx3dhip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "../../include/x3dhip.h"

namespace x3d {

// Stencil kinds of the periodic right-hand sides.
enum Kind : int {
  K_D = 0,    // collocated first derivative (der?_00)
  K_DVP = 1,  // staggered derivative, velocity -> pressure grid
  K_DPV = 2,  // staggered derivative, pressure -> velocity grid
  K_IVP = 3,  // mid-point interpolation, velocity -> pressure grid
  K_IPV = 4,  // mid-point interpolation, pressure -> velocity grid
  K_NONE = 5,  // no stencil: the Thomas solve alone (in place)
  K_DD = 6     // collocated second derivative (derxx)
};

// Scheme constants of one direction (module x3d_operator_?_data).
struct Scheme {
  double alfai, afi, bfi;                    // compact first derivative
  double alcai, aci, bci;                    // staggered derivative
  double ailcai, aici, bici, cici, dici;     // interpolation (ipinter = 3)
  double alsai, asi, bsi;                    // compact second derivative (isecondder = 4)
};

// Host copy of one x3doperator1d instance.
struct OpHost {
  int n = 0, npaire = 0;
  double alfa = 0.0;
  bool circulant = false;  // tables are the library's periodic [alfa, 1, alfa] ones
  bool np = false;         // non-periodic first derivative (der?_ij): boundary rows, no periodic correction
  size_t npoff = 0;        // np: its padded f, s, w in x3d_ctx::nptab (each npplan's C*M long)
  bool bc_unsupported = false;  // an operator of a non-periodic direction this library does not build
  std::vector<double> f, s, w, p;
  std::vector<double> f0, s0, w0, p0;  // the library's own tables, kept to recognise them
  double alfa0 = 0.0;
};

// Device view of one operator: four length-n vectors and alfa.
template <typename T>
struct OpDev {
  const T* f;
  const T* s;
  const T* w;
  const T* p;
  T alfa;
};

// How the lines of one direction are tiled onto a work-group.
struct LinePlan {
  int n = 0;      // line length
  int M = 0;      // points per chunk (per thread)
  int C = 0;      // chunks per line
  int G = 0;      // chunks of one line inside one wavefront
  int Wc = 0;     // wavefronts cooperating on the same lines
  int Wl = 0;     // independent line groups per work-group
  int LW = 0;     // lines per wavefront (64 / G)
  bool line_fast = false;  // lanes: consecutive lines (y/z) or consecutive chunks (x)
};

}  // namespace x3d

struct x3d_ctx {
  int nx = 0, ny = 0, nz = 0;
  double xlx = 1.0, yly = 1.0, zlz = 1.0, dx = 0.0, dy = 0.0, dz = 0.0;
  int device = 0;
  bool f64_lines = true;  // FP64 fields below 4 GiB (32-bit offsets of the line kernels)
  hipStream_t stream = nullptr;
  x3d::Scheme sc[3];
  int ncl1[3] = {0, 0, 0}, ncln[3] = {0, 0, 0};  // boundary codes per direction (0 periodic)
  x3d::LinePlan npplan[3];  // non-periodic directions: lines of any length, padded to C*M
  double* nptab = nullptr;  // padded Thomas tables of the non-periodic instances
  x3d::OpHost op[X3D_NOPS_ALL];
  // device operator tables: for op o, f,s,w,p at tab64 + off[o] + {0,1,2,3}*n
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  size_t off[X3D_NOPS_ALL] = {};
  size_t tab_len = 0;
  x3d::LinePlan plan[3];   // single-operator kernels
  x3d::LinePlan fplan[3];  // fused step passes (at most 512 threads per work-group)
  x3d::LinePlan fplan32[3];  // fused step passes of the FP32 precision form
  // Poisson
  hipfftHandle fft_r2c = 0, fft_c2r = 0;
  bool fft_ready = false;
  hipfftHandle fft_r2c32 = 0, fft_c2r32 = 0;  // FP32 plans, made on first use
  bool fft32_ready = false;
  double2* spec = nullptr;          // (nx/2+1) * ny * nz complex
  double* ktab = nullptr;           // device 1-D tables, see poisson source
  bool zsplit = false;              // fused z-stage kernel (nz a line-kernel length up to 1024)
  void* ztw = nullptr;              // e^{-2 pi i t / nz}, t < nz (double2)
  bool ysplit = false;              // and y-stage kernels (ny a line-kernel length up to 1024)
  void* ytw = nullptr;              // e^{-2 pi i t / ny}, t < ny (double2)
  bool xsplit = false;              // and the forward x r2c and y DFT as line kernels (X3D_XSPLIT=0: hipFFT 2-D)
  void* xtw = nullptr;              // e^{-2 pi i t / nx}, t < nx (double2)
  int ncu = 0;                      // compute units of the context's device
  int xcd_remap = 1;                // line blocks regrouped per XCD (X3D_XCD_REMAP=0 turns it off)
  int spl_wg[2][3] = {{0, 0, 0}, {0, 0, 0}};  // resident work-groups/CU of the Poisson line kernels [FP64/FP32][mode]
  int rfx_wg[3][2] = {{0, 0}, {0, 0}, {0, 0}};   // and of the x r2c / c2r kernels [FP64/FP32/BFP32][forward/inverse]
  std::vector<double> ax, bx, ay, by, az, bz;
  std::vector<double> xk2, yk2, zk2, exs, eys, ezs;
  // step scratch fields (allocated on first use)
  double* work[8] = {};
  size_t work_elems = 0;
  double* red = nullptr;            // reduction scratch (host-visible results)
  // per-pass timing of x3d_step_timed: event i is recorded after pass i
  hipEvent_t pass_ev[X3D_NPASS + 1] = {};
  bool timing = false;
  // dense-circulant fragments of the matrix-core variant, per operator (made on first use)
  void* circ_tab[X3D_NOPS] = {};
  float circ_scale[X3D_NOPS] = {};  // 2^-s of the power-of-two scaling of each table
  // captured step (x3d_graph_capture_step*, or the calls between
  // x3d_graph_begin and x3d_graph_end), replayed by x3d_graph_launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t cap_stream = nullptr;  // private capture stream while capturing
  hipStream_t cap_user = nullptr;    // the caller's stream, restored at the end
  // device copy of caller-supplied Thomas tables (f, s, w, periodic; 4 * max(nx, ny, nz))
  double* usertab = nullptr;
  // y-stretching (istret /= 0): ppy then ppyi, ny values each, on the device
  // in FP64 and FP32; sy_host is the host copy of the vectors last uploaded
  int istret = 0;
  double* sy64 = nullptr;
  float* sy32 = nullptr;
  std::vector<double> sy_host;
};

// A batch of independent grids (the batching form).  Each grid keeps its own
// context (plans, operator tables, FFT plans, scratch fields); the batch owns
// the fields of all grids, 10 blocks (ux uy uz dux duy duz pp3 px py pz) of
// all grids' points, grid after grid, and launches each step pass once per
// group of grids that share a line tiling.
struct x3d_batch {
  int ng = 0;
  int prec = 8;  // bytes per value: 8 FP64, 4 FP32
  std::vector<x3d_ctx*> g;
  std::vector<size_t> off;  // first point of each grid inside a field block
  size_t total = 0;         // points of all grids
  void* store = nullptr;
  hipStream_t stream = nullptr;
  int ncu = 256;
  // device argument records of the grouped launches, rebuilt when the plans
  // or gdt change
  struct Launch {
    int pass;       // 0-2 transeq z y x, 3-5 divergence x y z, 6-8 gradp z y x
    int M;
    dim3 grid, block;
    size_t smem;
    size_t args;    // byte offset of the group's records in dargs
    size_t block0;  // byte offset of its block prefix (ngroup + 1 ints)
    int ngroup;
  };
  std::vector<Launch> launches;
  void* dargs = nullptr;
  size_t dargs_bytes = 0;
  bool built = false;
  double built_gdt = 0.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // Poisson of the grids the line kernels take, grouped: one launch per
  // stage (0 x r2c, 1 forward y, 2 z, 3 inverse y, 4 x c2r) and line length
  // over all of them, whose per-grid records sit in pargs; the other grids
  // (psolo) solve one by one
  struct PLaunch {
    const void* fn;
    unsigned blocks, threads;
    size_t smem;
    int stage;
    const void* tw;    // twiddles of the group's line length
    size_t grids;      // byte offset of its grid records in pargs
    size_t bgrid;      // byte offset of its first grid per work-group
    long long tiles;   // tiles of all its grids
  };
  std::vector<PLaunch> plaunch;
  std::vector<int> psolo;
  void* pargs = nullptr;
  // side streams (forked from and joined back to the batch stream): the
  // launches of one pass or grouped Poisson stage, which touch disjoint grids,
  // spread over them so that each overlaps the others' ramp-up and drain
  // (overlap; X3D_BATCH_OVERLAP=0 keeps them in order), and the per-grid
  // Poisson solves; their fork / join events
  bool overlap = true;
  std::vector<hipStream_t> pstreams;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  void* field(int f, int i) const { return (char*)store + ((size_t)f * total + off[i]) * (size_t)prec; }
};

namespace x3d {

int hip_status(hipError_t e);

// Batched step passes (x3d_fused.hip): build the grouped launch records, and
// run the three directional passes of transeq (which = 0), divergence (1)
// or gradp with cor_vel (2) for every grid of the batch.
template <typename T>
int batch_build(x3d_batch* b, double gdt);
template <typename T>
int batch_passes(x3d_batch* b, int which);
// Order of the single-grid transeq passes for storage S: z, y, x (true), or
// x, y, z for FP32 (measured faster there).  The batch always runs z, y, x.
// The timed steps report the passes under the names of the z-first order.
template <typename S>
constexpr bool transeq_zfirst() { return !std::is_same<S, float>::value; }
// The side streams wait for the batch stream's work so far (fork), the batch
// stream for theirs (join; also after an error, returning the first one).
int batch_fork(x3d_batch* b);
int batch_join(x3d_batch* b, int rc);
int line_plan(int n, bool line_fast, LinePlan* out);
int line_plan_fused(int n, bool line_fast, LinePlan* out, int variant);

// Launch one compact operator (or Thomas alone) along direction dir.
int launch_line_f32(x3d_ctx* c, int dir, int kind, int op, const float* in, float* out);
// One word of a BFP32 field (block floating point: 31-bit mantissa, the
// exponent of its 16-point x-segment spread over the segment's low bits).
struct bfp32 {
  int32_t w;
};
int launch_line_bfp(x3d_ctx* c, int dir, int kind, int op, const bfp32* in, bfp32* out);

// Stretched y-derivatives (src/x3d_derive.f90:331-337, 427-433,
// src/x3d_staggered.f90:628-634, 874-880): with istret /= 0 the results of
// dery / derypv are multiplied by ppy(j), those of deryvp by ppyi(j).
// stretch_which: the vector a y-line operator kind takes (0 ppy, 1 ppyi), -1
// none, -2 refused (second derivatives: a stretched grid needs pp2y/pp4y).
inline int stretch_which(const x3d_ctx* c, int dir, int kind)
{
  if (!c->istret || dir != 1) return -1;
  if (kind == K_D || kind == K_DPV) return 0;
  if (kind == K_DVP) return 1;
  return kind == K_DD ? -2 : -1;
}
// out(i, j, k) *= s(j), s = ppy (which = 0) or ppyi (which = 1)
int stretch_y(x3d_ctx* c, int which, double* out);
int stretch_y_f32(x3d_ctx* c, int which, float* out);
int stretch_y_bfp(x3d_ctx* c, int which, bfp32* out);
// device vector for the fused y passes: nullptr unless stretched
template <typename T>
inline const T* stretch_vec(const x3d_ctx* c, int which)
{
  if (!c->istret) return nullptr;
  if constexpr (sizeof(T) == 8) return c->sy64 + (size_t)which * c->ny;
  else return c->sy32 + (size_t)which * c->ny;
}
// Boundary rows of the non-periodic first derivative der?_ij
// (src/x3d_derive.f90:156-217): codes 1 free-slip, 2 Dirichlet at each end,
// the instance's npaire, and first_deriv_exp_'s constants for the spacing.
struct NpRows {
  int n, ncl1, ncln, npaire;
  double afi, bfi, af1, bf1, cf1, af2, afm, afn, bfn, cfn;
};
NpRows np_rows(const x3d_ctx* c, int dir, int op);
int np_plan(int n, bool line_fast, LinePlan* P);
// rhs: der?_ij (boundary rows, then the sweeps); else the _12 sweeps alone on
// the input (in may equal out)
int launch_line_np(x3d_ctx* c, int dir, int op, const double* in, double* out, bool rhs);
int launch_line(x3d_ctx* c, int dir, int kind, int op, const double* in, double* out,
                bool periodic_fix);
// The Thomas sweeps alone (in place) with device tables f, s, w, p at
// tab + {0, 1, 2, 3} * n; periodic_fix adds the periodic correction.
int launch_thomas_tables(x3d_ctx* c, int dir, const double* tab, double alfa, double* tt, bool periodic_fix);

struct Geom;
struct Tiling;
void launch_dims(const x3d_ctx* c, int dir, Geom* geo, Tiling* tl, dim3* grid, dim3* block, size_t* smem,
                 size_t tbytes, bool fused = false);
void launch_dims_plan(const x3d_ctx* c, int dir, const LinePlan& P, Geom* geo, Tiling* tl, dim3* grid,
                      dim3* block, size_t* smem, size_t tbytes);
int line_plan_fused32(int n, bool line_fast, LinePlan* out);

int poisson_setup(x3d_ctx* c);
void poisson_release(x3d_ctx* c);
int poisson_solve(x3d_ctx* c, double* rhs);
// BFP32 right-hand side, solved in FP64 and stored back as BFP32 words, the
// decode in the x r2c and the encode in the x c2r (contexts whose Poisson
// runs the line kernels in x, y and z; else X3D_ERR_UNSUPPORTED)
int poisson_solve_bfp(x3d_ctx* c, int32_t* rhs);
// Grouped Poisson of a batch (x3d_poisson.hip): build the launches and their
// device records (outside any capture; X3D_BATCH_PGROUP=0 solves every grid
// on its own), and run them: on the batch stream, or with ns > 0 each stage's
// launches spread over the first ns side streams.
int poisson_batch_build(x3d_batch* b);
int poisson_batch_run(x3d_batch* b, int ns);

int ensure_work(x3d_ctx* c);

// Record the end of step pass i (1-based) when x3d_step_timed is running.
inline int mark_pass(x3d_ctx* c, int i)
{
  if (c->timing && hipEventRecord(c->pass_ev[i], c->stream) != hipSuccess) return X3D_ERR_HIP;
  return X3D_OK;
}

// Triply periodic: the reference's Poisson solver stops otherwise
// (src/poisson.f90:104-108), and so do the step and its Poisson entry points.
inline bool all_periodic(const x3d_ctx* c)
{
  for (int d = 0; d < 3; ++d)
    if (c->ncl1[d] || c->ncln[d]) return false;
  return true;
}

// Fused step passes (all operators must carry the library's periodic tables).
bool fused_ok(const x3d_ctx* c);
template <typename T>
int fused_momentum(x3d_ctx* c, double gdt, int euler, T* ux, T* uy, T* uz, T* dux, T* duy, T* duz);
template <typename T>
int fused_divergence(x3d_ctx* c, T* pp3, const T* ux, const T* uy, const T* uz);
template <typename T>
int fused_gradp(x3d_ctx* c, T* px, T* py, T* pz, const T* pp3, T* ux, T* uy, T* uz, int corr);

// FP32 Poisson: FP32 FFTs around an FP64 spectral division.
int poisson_prepare_f32(x3d_ctx* c);
int poisson_solve_f32(x3d_ctx* c, float* rhs);

}  // namespace x3d

#define X3D_CHECK_HIP(expr)                        \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return x3d::hip_status(_e); \
  } while (0)
