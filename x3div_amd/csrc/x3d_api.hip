// C-ABI entry points of libx3dhip: context and operator-table set-up,
// derivative / staggered / Thomas / transpose entries, the step pieces and
// test-mode reductions.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>

#include "x3d_internal.h"

namespace x3d {

int hip_status(hipError_t e)
{
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return X3D_ERR_ALLOC;
  return X3D_ERR_HIP;
}

static int hip_status_or_ok(hipError_t e) { return e == hipSuccess ? X3D_OK : hip_status(e); }

// ---------------------------------------------------------------------------
// Operator coefficients (src/x3d_operator_1d.f90).
// ---------------------------------------------------------------------------

// prepare(b, c, f, s, w, n) (:301-323): forward multipliers s and inverted
// pivots w of the tridiagonal matrix with lower b, diagonal c, upper f.
static void prepare(const std::vector<double>& b, const std::vector<double>& cd,
                    const std::vector<double>& f, std::vector<double>& s, std::vector<double>& w)
{
  const size_t n = cd.size();
  s.assign(n, 0.0);
  w = cd;
  for (size_t i = 1; i < n; ++i) {
    s[i] = b[i - 1] / w[i - 1];
    w[i] = w[i] - f[i - 1] * s[i];
  }
  for (size_t i = 0; i < n; ++i) w[i] = 1.0 / w[i];
}

// Periodic rows of first_deriv_imp_ (:414-471) and interpol_imp_ (:576-599):
// off-diagonals a except the last row/column, diagonal (2, 1, ..., 1, 1+a^2).
static void periodic_rows(double a, int n, std::vector<double>& f, std::vector<double>& cd,
                          std::vector<double>& b)
{
  f.assign(n, 0.0); cd.assign(n, 0.0); b.assign(n, 0.0);
  f[0] = a; f[1] = a; cd[0] = 2.0; cd[1] = 1.0; b[0] = a; b[1] = a;
  f[n - 3] = a; f[n - 2] = a; f[n - 1] = 0.0;
  cd[n - 3] = 1.0; cd[n - 2] = 1.0; cd[n - 1] = 1.0 + a * a;
  b[n - 3] = a; b[n - 2] = a; b[n - 1] = 0.0;
  for (int i = 2; i <= n - 4; ++i) { f[i] = a; cd[i] = 1.0; b[i] = a; }
}

// x3doperator1d%init (:234-266): periodic = thomas1d((-1, 0, ..., 0, alfa)).
static void op_init(OpHost& o, const std::vector<double>& f, const std::vector<double>& s,
                    const std::vector<double>& w, double alfa, int paire)
{
  const int n = (int)f.size();
  o.n = n; o.npaire = paire; o.alfa = alfa; o.f = f; o.s = s; o.w = w;
  o.p.assign(n, 0.0);
  o.p[0] = -1.0;
  o.p[n - 1] = alfa;
  o.circulant = true;
  for (int k = 1; k < n; ++k) o.p[k] = o.p[k] - o.p[k - 1] * s[k];
  o.p[n - 1] = o.p[n - 1] * w[n - 1];
  for (int k = n - 2; k >= 0; --k) o.p[k] = (o.p[k] - f[k] * o.p[k + 1]) * w[k];
  o.f0 = o.f; o.s0 = o.s; o.w0 = o.w; o.p0 = o.p; o.alfa0 = alfa;
}

static void scheme(Scheme& s, double h)
{
  // first_deriv_exp_ ifirstder = 4 (:350-353)
  s.alfai = 1.0 / 3.0;
  s.afi = (7.0 / 9.0) / h;
  s.bfi = (1.0 / 36.0) / h;
  // interpol_exp_ ifirstder = 4, ipinter = 3 (:509-545)
  s.alcai = 9.0 / 62.0;
  s.aci = (63.0 / 62.0) / h;
  s.bci = (17.0 / 62.0) / 3.0 / h;
  s.ailcai = 0.49;
  s.aici = 1.0 / 128.0 * (75.0 + 70.0 * s.ailcai);
  s.bici = 1.0 / 256.0 * (126.0 * s.ailcai - 25.0);
  s.cici = 1.0 / 256.0 * (-10.0 * s.ailcai + 3.0);
  s.dici = 0.0;
  // second derivative, isecondder = 4 (src/parameters.f90:107): the
  // sixth-order tridiagonal scheme of Lele (1992), alfa = 2/11, a = 12/11,
  // b = 3/11 on the stencils (u+1 - 2u + u-1)/h^2 and (u+2 - 2u + u-2)/(4h^2)
  s.alsai = 2.0 / 11.0;
  s.asi = (12.0 / 11.0) / (h * h);
  s.bsi = (3.0 / 44.0) / (h * h);
}

// The six operators of one direction (x3d_operator_1d_init, :76-189, periodic).
static void direction_ops(x3d_ctx* c, int d, int n)
{
  const Scheme& sc = c->sc[d];
  std::vector<double> f, cd, b, s, w;
  periodic_rows(sc.alfai, n, f, cd, b);
  prepare(b, cd, f, s, w);
  op_init(c->op[X3D_OP_DERX + 2 * d], f, s, w, sc.alfai, 0);
  op_init(c->op[X3D_OP_DERXP + 2 * d], f, s, w, sc.alfai, 1);  // ffp = ff when ncl1 /= 1
  periodic_rows(sc.alcai, n, f, cd, b);
  prepare(b, cd, f, s, w);
  op_init(c->op[X3D_OP_DERXVP + d], f, s, w, sc.alcai, 1);     // cfx6, csx6, cwx6
  op_init(c->op[X3D_OP_DERXPV + d], f, s, w, sc.alcai, 0);
  periodic_rows(sc.ailcai, n, f, cd, b);
  prepare(b, cd, f, s, w);
  op_init(c->op[X3D_OP_INTXVP + d], f, s, w, sc.ailcai, 1);    // cifxp6 == cifx6 here
  op_init(c->op[X3D_OP_INTXPV + d], f, s, w, sc.ailcai, 1);
  periodic_rows(sc.alsai, n, f, cd, b);
  prepare(b, cd, f, s, w);
  op_init(c->op[X3D_OP_DERXX + d], f, s, w, sc.alsai, 0);
}

// A non-periodic direction (codes ncl1, ncln in {1, 2}; x3d_operator_1d_init
// with first_deriv_imp_, src/x3d_operator_1d.f90:395-491): the two first
// derivative instances take the non-periodic rows, prepared twice as the
// reference does (derX: ff; derXp: ffp, which zeroes the free-slip corner
// entries).  The staggered, interpolation and second-derivative instances
// of that direction live on grids of n - 1 points in the reference and are
// not built: they keep periodic tables of the right length and report
// X3D_ERR_UNSUPPORTED.
static void np_direction_ops(x3d_ctx* c, int d, int n, int ncl1, int ncln)
{
  direction_ops(c, d, n);
  for (int op = 0; op < X3D_NOPS_ALL; ++op) {
    const bool dd = op < 6 ? op / 2 == d : (op - 6) % 3 == d;
    if (dd) {
      c->op[op].bc_unsupported = true;
      c->op[op].circulant = false;  // no form may take them for periodic
    }
  }
  const double h = d == 0 ? c->dx : d == 1 ? c->dy : c->dz;
  const double alfai = c->sc[d].alfai, alfa1 = 2.0, alfa2 = 1.0 / 4.0, alfan = 2.0, alfam = 1.0 / 4.0;
  (void)h;
  std::vector<double> ff(n, 0.0), fc(n, 0.0), fb(n, 0.0);
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
  std::vector<double> ffp = ff, s, w, sp, wp;
  prepare(fb, fc, ff, s, w);
  if (ncl1 == 1) ffp[0] = 0.0;
  if (ncln == 1) fb[n - 2] = 0.0;
  prepare(fb, fc, ffp, sp, wp);
  for (int k = 0; k < 2; ++k) {
    OpHost& o = c->op[X3D_OP_DERX + 2 * d + k];
    op_init(o, k ? ffp : ff, k ? sp : s, k ? wp : w, alfai, k);
    o.p.assign(n, 0.0);  // no periodic correction (x3doperator1d%init with ncl false)
    o.p0 = o.p;
    o.circulant = false;
    o.np = true;
    o.bc_unsupported = false;
  }
}

// The non-periodic instances' tables padded to their direction's C*M points
// with identity rows (f = s = 0, w = 1), in one device array.
static int np_upload(x3d_ctx* c)
{
  std::vector<double> host;
  for (int op = 0; op < X3D_NOPS_ALL; ++op) {
    OpHost& o = c->op[op];
    if (!o.np) continue;
    const int d = op / 2;  // the derivative instances: 2d, 2d + 1
    const size_t np = (size_t)c->npplan[d].C * c->npplan[d].M;
    o.npoff = host.size();
    host.resize(host.size() + 3 * np, 0.0);
    double* f = host.data() + o.npoff;
    for (size_t j = 0; j < np; ++j) {
      const bool real = j < (size_t)o.n;
      f[j] = real ? o.f[j] : 0.0;
      f[np + j] = real ? o.s[j] : 0.0;
      f[2 * np + j] = real ? o.w[j] : 1.0;
    }
  }
  if (host.empty()) return X3D_OK;
  if (!c->nptab) X3D_CHECK_HIP(hipMalloc(&c->nptab, host.size() * sizeof(double)));
  X3D_CHECK_HIP(hipMemcpy(c->nptab, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  return X3D_OK;
}

static int upload_tables(x3d_ctx* c)
{
  size_t len = 0;
  for (int o = 0; o < X3D_NOPS_ALL; ++o) { c->off[o] = len; len += 4 * (size_t)c->op[o].n; }
  if (!c->tab64 || c->tab_len != len) {
    if (c->tab64) (void)hipFree(c->tab64);
    if (c->tab32) (void)hipFree(c->tab32);
    c->tab64 = nullptr; c->tab32 = nullptr;
    X3D_CHECK_HIP(hipMalloc(&c->tab64, len * sizeof(double)));
    X3D_CHECK_HIP(hipMalloc(&c->tab32, len * sizeof(float)));
    c->tab_len = len;
  }
  std::vector<double> h(len);
  std::vector<float> hf(len);
  for (int o = 0; o < X3D_NOPS_ALL; ++o) {
    const OpHost& op = c->op[o];
    const size_t n = op.n;
    std::memcpy(&h[c->off[o]], op.f.data(), n * sizeof(double));
    std::memcpy(&h[c->off[o] + n], op.s.data(), n * sizeof(double));
    std::memcpy(&h[c->off[o] + 2 * n], op.w.data(), n * sizeof(double));
    std::memcpy(&h[c->off[o] + 3 * n], op.p.data(), n * sizeof(double));
  }
  for (size_t i = 0; i < len; ++i) hf[i] = (float)h[i];
  X3D_CHECK_HIP(hipMemcpy(c->tab64, h.data(), len * sizeof(double), hipMemcpyHostToDevice));
  X3D_CHECK_HIP(hipMemcpy(c->tab32, hf.data(), len * sizeof(float), hipMemcpyHostToDevice));
  return X3D_OK;
}

int ensure_work(x3d_ctx* c)
{
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  if (c->work[0] && c->work_elems == N) return X3D_OK;
  for (double*& w : c->work) {
    if (w) (void)hipFree(w);
    w = nullptr;
  }
  for (double*& w : c->work) X3D_CHECK_HIP(hipMalloc(&w, N * sizeof(double)));
  c->work_elems = N;
  return X3D_OK;
}

// ---------------------------------------------------------------------------
// Elementwise glue kernels.
// ---------------------------------------------------------------------------
enum EwOp { EW_MUL = 0, EW_FMA_ADD = 1, EW_ADD = 2, EW_AXPY = 3, EW_SUB = 4 };

// EW_MUL:     o = a * b
// EW_FMA_ADD: o = t + a * b         (src/transeq.f90:114, 133, 152)
// EW_ADD:     o = a + b             (src/transeq.f90:164, 176; navier.f90:183, 201)
// EW_AXPY:    o = g * a + b         (src/time_integrators.f90:42)
// EW_SUB:     o = a - b             (src/navier.f90:84-86)
__global__ void k_ew(int opk, double* __restrict__ o, const double* a, const double* b,
                     const double* t, double gsc, long long N)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
    double r;
    switch (opk) {
      case EW_MUL: r = a[q] * b[q]; break;
      case EW_FMA_ADD: r = t[q] + a[q] * b[q]; break;
      case EW_ADD: r = a[q] + b[q]; break;
      case EW_AXPY: r = gsc * a[q] + b[q]; break;
      default: r = a[q] - b[q]; break;
    }
    o[q] = r;
  }
}

static int ew(x3d_ctx* c, int opk, double* o, const double* a, const double* b, const double* t,
              double gsc)
{
  const long long N = (long long)c->nx * c->ny * c->nz;
  long long blocks = (N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_ew, dim3((unsigned)blocks), dim3(256), 0, c->stream, opk, o, a, b, t, gsc, N);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

// ---------------------------------------------------------------------------
// Reductions (deterministic two-level sums in FP64).
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024;

__device__ inline uint64_t splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ inline double uniform01(uint64_t seed, int comp, uint64_t idx)
{
  const uint64_t key = splitmix64(seed ^ ((uint64_t)(comp + 1) * 0x9E3779B97F4A7C15ull));
  return (double)(splitmix64(key + idx) >> 11) * 0x1.0p-53;
}

struct GridInfo {
  int nx, ny, nz;
  double dx, dy, dz, xlx, yly, zlz;
};

// which: 0 dudx, 1 dvdy, 2 dwdz (src/mom.f90:95-96, 158-159, 217-218);
// 3 = max/sum of |x| (src/navier.f90:215-217).  Partial results per block.
__global__ void k_reduce_partial(const double* __restrict__ d, int which, GridInfo g,
                                 double* __restrict__ part)
{
  const double twopi = 2.0 * 3.14159265358979323846;
  const long long N = (long long)g.nx * g.ny * g.nz;
  double acc = 0.0, mx = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (long long)gridDim.x * blockDim.x) {
    if (which == 3) {
      const double a = fabs(d[q]);
      acc += a;
      mx = a > mx ? a : mx;
    } else {
      const int i = (int)(q % g.nx);
      const int j = (int)((q / g.nx) % g.ny);
      const int k = (int)(q / ((long long)g.nx * g.ny));
      const double x = i * g.dx, y = j * g.dy, z = k * g.dz;
      double ana = cos(twopi * (x / g.xlx)) * cos(twopi * (y / g.yly)) * cos(twopi * (z / g.zlz));
      if (which == 0) ana = twopi * ana / g.xlx;
      else if (which == 1) ana = twopi * ana / g.yly;
      else ana = twopi * (-2.0 * ana) / g.zlz;
      const double e = d[q] - ana;
      acc += e * e;
    }
  }
  __shared__ double ssum[256], smax[256];
  ssum[threadIdx.x] = acc;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      ssum[threadIdx.x] += ssum[threadIdx.x + s];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = ssum[0]; part[2 * blockIdx.x + 1] = smax[0]; }
}

__global__ void k_reduce_final(const double* __restrict__ part, int nparts, double* __restrict__ res)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0, m = 0.0;
    for (int b = 0; b < nparts; ++b) { s += part[2 * b]; m = fmax(m, part[2 * b + 1]); }
    res[0] = s; res[1] = m;
  }
}

static int reduce(x3d_ctx* c, const double* d, int which, double* sum, double* mx)
{
  if (!c->red) X3D_CHECK_HIP(hipMalloc(&c->red, sizeof(double) * (2 * RED_BLOCKS + 2)));
  GridInfo g{c->nx, c->ny, c->nz, c->dx, c->dy, c->dz, c->xlx, c->yly, c->zlz};
  hipLaunchKernelGGL(k_reduce_partial, dim3(RED_BLOCKS), dim3(256), 0, c->stream, d, which, g, c->red);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, c->stream, c->red, RED_BLOCKS,
                     c->red + 2 * RED_BLOCKS);
  X3D_CHECK_HIP(hipGetLastError());
  double res[2];
  X3D_CHECK_HIP(hipMemcpyAsync(res, c->red + 2 * RED_BLOCKS, sizeof(res), hipMemcpyDeviceToHost, c->stream));
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  *sum = res[0];
  *mx = res[1];
  return X3D_OK;
}

// Initial fields; same integer generator as the host reference so both sides
// draw identical noise.
__global__ void k_init_fields(int kind, uint64_t seed, double amp, GridInfo g, const int* kv,
                              const double* ph, double* __restrict__ ux, double* __restrict__ uy,
                              double* __restrict__ uz)
{
  const double twopi = 2.0 * 3.14159265358979323846;
  const long long N = (long long)g.nx * g.ny * g.nz;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % g.nx);
    const int j = (int)((q / g.nx) % g.ny);
    const int k = (int)(q / ((long long)g.nx * g.ny));
    if (kind == 2) {
      double* dst[3] = {ux, uy, uz};
      for (int cpt = 0; cpt < 3; ++cpt) {
        double acc = 0.0;
        for (int m = 0; m < 8; ++m) {
          const int* kk = kv + 3 * (8 * cpt + m);
          const double arg = twopi * ((double)kk[0] * (double)i / (double)g.nx +
                                      (double)kk[1] * (double)j / (double)g.ny +
                                      (double)kk[2] * (double)k / (double)g.nz) + ph[8 * cpt + m];
          acc += 0.125 * sin(arg);
        }
        dst[cpt][q] = acc;
      }
      continue;
    }
    const double x = i * g.dx, y = j * g.dy, z = k * g.dz;
    const double sx = sin(twopi * (x / g.xlx)), cx = cos(twopi * (x / g.xlx));
    const double sy = sin(twopi * (y / g.yly)), cy = cos(twopi * (y / g.yly));
    const double sz = sin(twopi * (z / g.zlz)), cz = cos(twopi * (z / g.zlz));
    if (kind == 0) {
      ux[q] = sx * cy * cz;
      uy[q] = cx * sy * cz;
      uz[q] = -2.0 * cx * cy * sz;
    } else {
      ux[q] = sx * cy * cz + amp * (2.0 * uniform01(seed, 0, q) - 1.0);
      uy[q] = -cx * sy * cz + amp * (2.0 * uniform01(seed, 1, q) - 1.0);
      uz[q] = amp * (2.0 * uniform01(seed, 2, q) - 1.0);
    }
  }
}

static uint64_t host_splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

static double host_uniform01(uint64_t seed, int comp, uint64_t idx)
{
  const uint64_t key = host_splitmix64(seed ^ ((uint64_t)(comp + 1) * 0x9E3779B97F4A7C15ull));
  return (double)(host_splitmix64(key + idx) >> 11) * 0x1.0p-53;
}

}  // namespace x3d

using namespace x3d;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int x3d_version(void) { return 100; }

const char* x3d_strerror(int code)
{
  switch (code) {
    case X3D_OK: return "ok";
    case X3D_ERR_ARG: return "invalid argument";
    case X3D_ERR_SIZE: return "array sizes do not match the context (or n < 8)";
    case X3D_ERR_HIP: return "HIP runtime error";
    case X3D_ERR_FFT: return "hipFFT error";
    case X3D_ERR_ALLOC: return "device allocation failed";
    case X3D_ERR_UNSUPPORTED: return "not supported: line length not tiled by the line kernels, FP64 field of 4 GiB or more, or replaced operator tables";
    default: return "unknown error";
  }
}

double x3d_wtime(void)
{
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

// ---------------------------------------------------------------------------
// Scratch of x3d_step_noalloc: the step's 8 work fields for grids of up to
// X3D_NOALLOC_MAX_POINTS points, as a static device array.  Each device the
// library is loaded on holds its own copy (128 MiB) until the process ends.
// Host code never uses the array's name as an address: its device address is
// looked up once per device with hipGetSymbolAddress (at x3d_create) and kept
// in scratch_addr.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchElems = (size_t)8 * X3D_NOALLOC_MAX_POINTS;
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_addr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kScratchElems];

// Device address of `scratch` on the current device `dev`, looked up once.
static int scratch_setup(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return X3D_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_addr[dev]) return X3D_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return X3D_ERR_HIP;
  scratch_addr[dev] = static_cast<double*>(p);
  return X3D_OK;
}

static double* scratch_of(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_addr[dev];
}

int x3d_create(int nx, int ny, int nz, double xlx, double yly, double zlz, x3d_ctx** out)
{
  return x3d_create_bc(nx, ny, nz, xlx, yly, zlz, nullptr, nullptr, out);
}

int x3d_create_bc(int nx, int ny, int nz, double xlx, double yly, double zlz, const int* ncl1, const int* ncln,
                  x3d_ctx** out)
{
  if (!out) return X3D_ERR_ARG;
  *out = nullptr;
  if (nx < 8 || ny < 8 || nz < 8) return X3D_ERR_SIZE;
  if (!(xlx > 0) || !(yly > 0) || !(zlz > 0)) return X3D_ERR_ARG;
  int b1[3] = {0, 0, 0}, bn[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    if (ncl1) b1[d] = ncl1[d];
    if (ncln) bn[d] = ncln[d];
    // periodic at both ends or at neither (src/parameters.f90: nclx = nclx1 == 0 .and. nclxn == 0)
    if (b1[d] < 0 || b1[d] > 2 || bn[d] < 0 || bn[d] > 2 || ((b1[d] == 0) != (bn[d] == 0))) return X3D_ERR_ARG;
  }
  // the line kernels address a field with 32-bit byte offsets: FP32 fields
  // must stay below 4 GiB, and FP64 entry points need FP64 fields below it
  const unsigned long long npts = (unsigned long long)nx * ny * nz;
  if (npts * sizeof(float) > 0xFFFFFFFFull) return X3D_ERR_UNSUPPORTED;
  x3d_ctx* c = new x3d_ctx();
  c->f64_lines = npts * sizeof(double) <= 0xFFFFFFFFull;
  c->nx = nx; c->ny = ny; c->nz = nz;
  c->xlx = xlx; c->yly = yly; c->zlz = zlz;
  // dX = L / nXm with nXm = nX periodic, nX - 1 otherwise (src/parameters.f90)
  for (int d = 0; d < 3; ++d) { c->ncl1[d] = b1[d]; c->ncln[d] = bn[d]; }
  c->dx = xlx / (double)(b1[0] ? nx - 1 : nx);
  c->dy = yly / (double)(b1[1] ? ny - 1 : ny);
  c->dz = zlz / (double)(b1[2] ? nz - 1 : nz);
  int rc = X3D_OK;
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return X3D_ERR_HIP; }
  if ((rc = scratch_setup(c->device)) != X3D_OK) { delete c; return rc; }
  scheme(c->sc[0], c->dx);
  scheme(c->sc[1], c->dy);
  scheme(c->sc[2], c->dz);
  const int ns[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    if (b1[d]) np_direction_ops(c, d, ns[d], b1[d], bn[d]);
    else direction_ops(c, d, ns[d]);
  }
  if (const char* xr = std::getenv("X3D_XCD_REMAP")) c->xcd_remap = std::atoi(xr) != 0;
  // tiling variant of the fused y/z passes (X3D_FUSED_TILING=0|1, default 0)
  const char* tv = std::getenv("X3D_FUSED_TILING");
  const int variant = tv ? std::atoi(tv) : 0;
  // FP32 precision-form plans; a line they cannot tile leaves M = 0 and the
  // FP32 step reports X3D_ERR_UNSUPPORTED
  if (line_plan_fused32(nx, false, &c->fplan32[0]) || line_plan_fused32(ny, true, &c->fplan32[1]) ||
      line_plan_fused32(nz, true, &c->fplan32[2]))
    c->fplan32[0].M = c->fplan32[1].M = c->fplan32[2].M = 0;
  // a non-periodic direction runs its own kernels on its own plan (lines of
  // any length); the periodic plans there may not exist, and are not used
  for (int d = 0; d < 3; ++d) {
    if (!b1[d]) continue;
    if ((rc = np_plan(ns[d], d != 0, &c->npplan[d]))) { x3d_destroy(c); return rc; }
    if (line_plan(ns[d], d != 0, &c->plan[d])) c->plan[d] = LinePlan{};
    if (line_plan_fused(ns[d], d != 0, &c->fplan[d], variant)) c->fplan[d] = LinePlan{};
  }
  auto plans = [&]() -> int {
    int r = X3D_OK;
    if (!b1[0] && ((r = line_plan(nx, false, &c->plan[0])) || (r = line_plan_fused(nx, false, &c->fplan[0], variant))))
      return r;
    if (!b1[1] && ((r = line_plan(ny, true, &c->plan[1])) || (r = line_plan_fused(ny, true, &c->fplan[1], variant))))
      return r;
    if (!b1[2] && ((r = line_plan(nz, true, &c->plan[2])) || (r = line_plan_fused(nz, true, &c->fplan[2], variant))))
      return r;
    return r;
  };
  if ((rc = plans()) || (rc = upload_tables(c)) || (rc = np_upload(c)) ||
      (all_periodic(c) && (rc = poisson_setup(c))) ||
      (rc = hip_status_or_ok(hipMalloc(&c->usertab, 4 * sizeof(double) * (size_t)std::max(nx, std::max(ny, nz)))))) {
    x3d_destroy(c);
    return rc;
  }
  *out = c;
  return X3D_OK;
}

static void graph_release(x3d_ctx* c);

int x3d_destroy(x3d_ctx* c)
{
  if (!c) return X3D_ERR_ARG;
  poisson_release(c);
  if (c->tab64) (void)hipFree(c->tab64);
  if (c->tab32) (void)hipFree(c->tab32);
  for (double* w : c->work)
    if (w) (void)hipFree(w);
  if (c->red) (void)hipFree(c->red);
  for (hipEvent_t e : c->pass_ev)
    if (e) (void)hipEventDestroy(e);
  graph_release(c);
  for (void* t : c->circ_tab)
    if (t) (void)hipFree(t);
  if (c->usertab) (void)hipFree(c->usertab);
  if (c->nptab) (void)hipFree(c->nptab);
  if (c->sy64) (void)hipFree(c->sy64);
  if (c->sy32) (void)hipFree(c->sy32);
  if (c->cap_stream) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->cap_stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(c->cap_stream);
  }
  delete c;
  return X3D_OK;
}

int x3d_set_stream(x3d_ctx* c, void* stream)
{
  if (!c) return X3D_ERR_ARG;
  c->stream = (hipStream_t)stream;
  if (c->fft_ready) {
    if (hipfftSetStream(c->fft_r2c, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
    if (hipfftSetStream(c->fft_c2r, c->stream) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  }
  return X3D_OK;
}

int x3d_sync(x3d_ctx* c)
{
  if (!c) return X3D_ERR_ARG;
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_get_sizes(const x3d_ctx* c, int* nx, int* ny, int* nz)
{
  if (!c || !nx || !ny || !nz) return X3D_ERR_ARG;
  *nx = c->nx; *ny = c->ny; *nz = c->nz;
  return X3D_OK;
}

int x3d_field_alloc(x3d_ctx* c, size_t nbytes, void** p)
{
  if (!c || !p) return X3D_ERR_ARG;
  X3D_CHECK_HIP(hipMalloc(p, nbytes));
  return X3D_OK;
}

int x3d_field_free(x3d_ctx* c, void* p)
{
  if (!c) return X3D_ERR_ARG;
  if (p) X3D_CHECK_HIP(hipFree(p));
  return X3D_OK;
}

int x3d_memcpy_h2d(x3d_ctx* c, void* dst, const void* src, size_t nbytes)
{
  if (!c || !dst || !src) return X3D_ERR_ARG;
  X3D_CHECK_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, c->stream));
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_memcpy_d2h(x3d_ctx* c, void* dst, const void* src, size_t nbytes)
{
  if (!c || !dst || !src) return X3D_ERR_ARG;
  X3D_CHECK_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, c->stream));
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_operator_get(const x3d_ctx* c, int op, double* f, double* s, double* w, double* p,
                     double* alfa, int* n, int* npaire)
{
  if (!c || op < 0 || op >= X3D_NOPS_ALL) return X3D_ERR_ARG;
  const OpHost& o = c->op[op];
  const size_t nb = sizeof(double) * o.n;
  if (f) std::memcpy(f, o.f.data(), nb);
  if (s) std::memcpy(s, o.s.data(), nb);
  if (w) std::memcpy(w, o.w.data(), nb);
  if (p) std::memcpy(p, o.p.data(), nb);
  if (alfa) *alfa = o.alfa;
  if (n) *n = o.n;
  if (npaire) *npaire = o.npaire;
  return X3D_OK;
}

int x3d_operator_set(x3d_ctx* c, int op, const double* f, const double* s, const double* w,
                     const double* p, double alfa)
{
  if (!c || op < 0 || op >= X3D_NOPS_ALL || !f || !s || !w || !p) return X3D_ERR_ARG;
  OpHost& o = c->op[op];
  const int n = o.n;
  const size_t nb = (size_t)n * sizeof(double);
  // the library's own tables (e.g. a caller rebuilding them with
  // x3doperator1d%init): the operator keeps the circulant solver
  const bool own = alfa == o.alfa0 && !std::memcmp(f, o.f0.data(), nb) && !std::memcmp(s, o.s0.data(), nb) &&
                   !std::memcmp(w, o.w0.data(), nb) && !std::memcmp(p, o.p0.data(), nb);
  if (own && o.circulant) return X3D_OK;
  o.f.assign(f, f + n); o.s.assign(s, s + n); o.w.assign(w, w + n); o.p.assign(p, p + n);
  o.alfa = alfa;
  o.circulant = own && !o.np;  // caller tables: solved with the table sweeps
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  // a step captured earlier ran the circulant kernels on the old device
  // tables, which upload_tables replaces: it must not be replayed
  graph_release(c);
  int rc = upload_tables(c);
  return rc ? rc : (o.np ? np_upload(c) : X3D_OK);
}

int x3d_scheme_consts(const x3d_ctx* c, double* out)
{
  if (!c || !out) return X3D_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    const Scheme& s = c->sc[d];
    double* o = out + 11 * d;
    o[0] = s.alfai; o[1] = s.afi; o[2] = s.bfi; o[3] = s.alcai; o[4] = s.aci; o[5] = s.bci;
    o[6] = s.ailcai; o[7] = s.aici; o[8] = s.bici; o[9] = s.cici; o[10] = s.dici;
  }
  return X3D_OK;
}

// ---- Thomas ---------------------------------------------------------------
static int check3(const x3d_ctx* c, int nx, int ny, int nz)
{
  if (!c) return X3D_ERR_ARG;
  return (nx == c->nx && ny == c->ny && nz == c->nz) ? X3D_OK : X3D_ERR_SIZE;
}

// ---- y-stretching (istret, ppy, ppyi: src/parameters.f90:117, src/var.f90:138-140)
// Host vectors, ny values each, kept on the device in FP64 and FP32.  The
// upload waits for the stream first: earlier work may still read the vectors.
static int stretch_upload(x3d_ctx* c, const double* ppy, const double* ppyi)
{
  const size_t ny = (size_t)c->ny;
  std::vector<double> h(2 * ny);
  std::memcpy(h.data(), ppy, ny * sizeof(double));
  std::memcpy(h.data() + ny, ppyi, ny * sizeof(double));
  std::vector<float> h32(h.begin(), h.end());
  if (!c->sy64) X3D_CHECK_HIP(hipMalloc(&c->sy64, 2 * ny * sizeof(double)));
  if (!c->sy32) X3D_CHECK_HIP(hipMalloc(&c->sy32, 2 * ny * sizeof(float)));
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  X3D_CHECK_HIP(hipMemcpy(c->sy64, h.data(), 2 * ny * sizeof(double), hipMemcpyHostToDevice));
  X3D_CHECK_HIP(hipMemcpy(c->sy32, h32.data(), 2 * ny * sizeof(float), hipMemcpyHostToDevice));
  c->sy_host = std::move(h);
  return X3D_OK;
}

// The vector a y-derivative entry point was given (which = 0 ppy, 1 ppyi):
// with istret /= 0 the reference multiplies by the vector passed, so one
// that differs from the context's replaces it (both vectors stay in use by
// the step passes).  NULL keeps the context's; istret = 0 ignores it.
static int stretch_arg(x3d_ctx* c, int which, const double* v)
{
  if (!c->istret || !v) return X3D_OK;
  const size_t ny = (size_t)c->ny;
  const double* cur = c->sy_host.data() + which * ny;
  if (std::memcmp(cur, v, ny * sizeof(double)) == 0) return X3D_OK;
  const std::vector<double> keep(c->sy_host);
  graph_release(c);
  return which == 0 ? stretch_upload(c, v, keep.data() + ny) : stretch_upload(c, keep.data(), v);
}

int x3d_set_stretching(x3d_ctx* c, int istret, const double* ppy, const double* ppyi)
{
  if (!c || (istret && (!ppy || !ppyi))) return X3D_ERR_ARG;
  // a captured step holds the previous state
  graph_release(c);
  if (istret) {
    const int rc = stretch_upload(c, ppy, ppyi);
    if (rc) return rc;
  }
  c->istret = istret;
  return X3D_OK;
}

int x3d_xthomas(x3d_ctx* c, double* tt, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 0, K_NONE, op, tt, tt, true);
}
int x3d_ythomas(x3d_ctx* c, double* tt, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 1, K_NONE, op, tt, tt, true);
}
int x3d_zthomas(x3d_ctx* c, double* tt, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 2, K_NONE, op, tt, tt, true);
}

// The generics' members with the caller's own host tables: copied into the
// context's device table (ordered on the stream after earlier calls that
// read it), then the sweeps, and the periodic correction when perio is given.
static int thomas_tab(x3d_ctx* c, int dir, double* tt, const double* ff, const double* fs, const double* fw,
                      const double* perio, double alfa, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  if (rc) return rc;
  if (!tt || !ff || !fs || !fw || !c->usertab) return X3D_ERR_ARG;
  const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
  const size_t nb = (size_t)n * sizeof(double);
  std::vector<double> h(4 * (size_t)n, 0.0);
  std::memcpy(h.data(), ff, nb);
  std::memcpy(h.data() + n, fs, nb);
  std::memcpy(h.data() + 2 * n, fw, nb);
  if (perio) std::memcpy(h.data() + 3 * n, perio, nb);
  X3D_CHECK_HIP(hipMemcpyAsync(c->usertab, h.data(), 4 * nb, hipMemcpyHostToDevice, c->stream));
  // the host staging vector must outlive the copy
  X3D_CHECK_HIP(hipStreamSynchronize(c->stream));
  return launch_thomas_tables(c, dir, c->usertab, alfa, tt, perio != nullptr);
}

int x3d_xthomas_tab(x3d_ctx* c, double* tt, const double* ff, const double* fs, const double* fw,
                    const double* perio, double alfa, int nx, int ny, int nz)
{
  return c ? thomas_tab(c, 0, tt, ff, fs, fw, perio, alfa, nx, ny, nz) : X3D_ERR_ARG;
}
int x3d_ythomas_tab(x3d_ctx* c, double* tt, const double* ff, const double* fs, const double* fw,
                    const double* perio, double alfa, int nx, int ny, int nz)
{
  return c ? thomas_tab(c, 1, tt, ff, fs, fw, perio, alfa, nx, ny, nz) : X3D_ERR_ARG;
}
int x3d_zthomas_tab(x3d_ctx* c, double* tt, const double* ff, const double* fs, const double* fw,
                    const double* perio, double alfa, int nx, int ny, int nz)
{
  return c ? thomas_tab(c, 2, tt, ff, fs, fw, perio, alfa, nx, ny, nz) : X3D_ERR_ARG;
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int x3d_fill(x3d_ctx* c, double* p, size_t n, double v)
{
  if (!c || (!p && n)) return X3D_ERR_ARG;
  if (!n) return X3D_OK;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, c->stream, p, n, v);
  X3D_CHECK_HIP(hipGetLastError());
  return X3D_OK;
}

// thomas1d (src/thomas.f90:200-218): one line on the host.
int x3d_thomas1d(double* tt, const double* ff, const double* fs, const double* fw, int n)
{
  if (!tt || !ff || !fs || !fw || n < 1) return X3D_ERR_ARG;
  for (int k = 1; k < n; ++k) tt[k] = tt[k] - tt[k - 1] * fs[k];
  tt[n - 1] = tt[n - 1] * fw[n - 1];
  for (int k = n - 2; k >= 0; --k) tt[k] = (tt[k] - ff[k] * tt[k + 1]) * fw[k];
  return X3D_OK;
}

// ---- collocated derivatives ----------------------------------------------
int x3d_derx(x3d_ctx* c, double* t, const double* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 0, K_D, op, u, t, true);
}
int x3d_dery(x3d_ctx* c, double* t, const double* u, int op, const double* ppy, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  if (!rc) rc = stretch_arg(c, 0, ppy);
  return rc ? rc : launch_line(c, 1, K_D, op, u, t, true);
}
int x3d_derz(x3d_ctx* c, double* t, const double* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 2, K_D, op, u, t, true);
}

// derxx / deryy / derzz (declared at src/x3d_derive.f90:54-59): the compact
// second derivative, periodic, through the same line kernels.
int x3d_derxx(x3d_ctx* c, double* t, const double* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 0, K_DD, op, u, t, true);
}
int x3d_deryy(x3d_ctx* c, double* t, const double* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 1, K_DD, op, u, t, true);
}
int x3d_derzz(x3d_ctx* c, double* t, const double* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line(c, 2, K_DD, op, u, t, true);
}

// ---- staggered (periodic: nXm = nX) --------------------------------------
#define X3D_STAG(name, dir, kind, a1, a2, a3, a4, cond)                                        \
  int name(x3d_ctx* c, double* t, const double* u, int op, int a1, int a2, int a3, int a4)    \
  {                                                                                           \
    if (!c) return X3D_ERR_ARG;                                                               \
    if (!(cond)) return X3D_ERR_SIZE;                                                         \
    return launch_line(c, dir, kind, op, u, t, true);                                         \
  }

X3D_STAG(x3d_derxvp, 0, K_DVP, nx, nxm, ny, nz, nx == c->nx && nxm == nx && ny == c->ny && nz == c->nz)
X3D_STAG(x3d_interxvp, 0, K_IVP, nx, nxm, ny, nz, nx == c->nx && nxm == nx && ny == c->ny && nz == c->nz)
X3D_STAG(x3d_derxpv, 0, K_DPV, nxm, nx, ny, nz, nx == c->nx && nxm == nx && ny == c->ny && nz == c->nz)
X3D_STAG(x3d_interxpv, 0, K_IPV, nxm, nx, ny, nz, nx == c->nx && nxm == nx && ny == c->ny && nz == c->nz)
X3D_STAG(x3d_interyvp, 1, K_IVP, nx, ny, nym, nz, nx == c->nx && ny == c->ny && nym == ny && nz == c->nz)
X3D_STAG(x3d_interypv, 1, K_IPV, nx, nym, ny, nz, nx == c->nx && ny == c->ny && nym == ny && nz == c->nz)
X3D_STAG(x3d_derzvp, 2, K_DVP, nx, ny, nz, nzm, nx == c->nx && ny == c->ny && nz == c->nz && nzm == nz)
X3D_STAG(x3d_interzvp, 2, K_IVP, nx, ny, nz, nzm, nx == c->nx && ny == c->ny && nz == c->nz && nzm == nz)
X3D_STAG(x3d_derzpv, 2, K_DPV, nx, ny, nzm, nz, nx == c->nx && ny == c->ny && nz == c->nz && nzm == nz)
X3D_STAG(x3d_interzpv, 2, K_IPV, nx, ny, nzm, nz, nx == c->nx && ny == c->ny && nz == c->nz && nzm == nz)
#undef X3D_STAG

int x3d_deryvp(x3d_ctx* c, double* t, const double* u, int op, const double* ppyi, int nx, int ny,
               int nym, int nz)
{
  if (!c) return X3D_ERR_ARG;
  if (!(nx == c->nx && ny == c->ny && nym == ny && nz == c->nz)) return X3D_ERR_SIZE;
  const int rc = stretch_arg(c, 1, ppyi);
  return rc ? rc : launch_line(c, 1, K_DVP, op, u, t, true);
}

int x3d_derypv(x3d_ctx* c, double* t, const double* u, int op, const double* ppy, int nx, int nym,
               int ny, int nz)
{
  if (!c) return X3D_ERR_ARG;
  if (!(nx == c->nx && ny == c->ny && nym == ny && nz == c->nz)) return X3D_ERR_SIZE;
  const int rc = stretch_arg(c, 0, ppy);
  return rc ? rc : launch_line(c, 1, K_DPV, op, u, t, true);
}

// ---- transposes ------------------------------------------------------------
static int tcopy(x3d_ctx* c, const void* in, void* out, size_t eb, int n1, int n2, int n3)
{
  if (!c || !in || !out || (eb != 4 && eb != 8 && eb != 16) || n1 <= 0 || n2 <= 0 || n3 <= 0)
    return X3D_ERR_ARG;
  if (in == out) return X3D_OK;
  X3D_CHECK_HIP(hipMemcpyAsync(out, in, eb * (size_t)n1 * n2 * n3, hipMemcpyDeviceToDevice, c->stream));
  return X3D_OK;
}
int x3d_transpose_x_to_y(x3d_ctx* c, const void* in, void* out, size_t eb, int n1, int n2, int n3)
{ return tcopy(c, in, out, eb, n1, n2, n3); }
int x3d_transpose_y_to_z(x3d_ctx* c, const void* in, void* out, size_t eb, int n1, int n2, int n3)
{ return tcopy(c, in, out, eb, n1, n2, n3); }
int x3d_transpose_z_to_y(x3d_ctx* c, const void* in, void* out, size_t eb, int n1, int n2, int n3)
{ return tcopy(c, in, out, eb, n1, n2, n3); }
int x3d_transpose_y_to_x(x3d_ctx* c, const void* in, void* out, size_t eb, int n1, int n2, int n3)
{ return tcopy(c, in, out, eb, n1, n2, n3); }

// ---- Poisson -----------------------------------------------------------------
int x3d_poisson(x3d_ctx* c, double* rhs)
{
  if (!c || !rhs) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  return poisson_solve(c, rhs);
}

// ---- step pieces (reference-shaped composition) ------------------------------
#define X3D_TRY(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

}  // extern "C"

namespace x3d {
// Reference-shaped compositions: one line kernel per operator call and one
// elementwise kernel per glue loop, in the reference's statement order.  Used
// when a caller replaced operator tables (then the line kernels run the table
// sweeps); otherwise the fused passes of x3d_fused.hip run.
static int ref_momentum_rhs(x3d_ctx* c, double* dux, double* duy, double* duz, const double* ux,
                            const double* uy, const double* uz)
{
  X3D_TRY(ensure_work(c));
  double *ta = c->work[0], *td = c->work[1], *tg = c->work[2], *t3 = c->work[3];
  const double* u[3] = {ux, uy, uz};
  double* du[3] = {dux, duy, duz};
  // operator instances per component and direction, {product, field}
  // (src/transeq.f90:106-107, 126-127, 145-146 / 192-193, 211-212, 231-232 /
  //  278-279, 297-298, 316-317)
  static const int opx[3][2] = {{X3D_OP_DERXP, X3D_OP_DERX}, {X3D_OP_DERX, X3D_OP_DERXP}, {X3D_OP_DERX, X3D_OP_DERXP}};
  static const int opy[3][2] = {{X3D_OP_DERY, X3D_OP_DERYP}, {X3D_OP_DERYP, X3D_OP_DERY}, {X3D_OP_DERY, X3D_OP_DERYP}};
  static const int opz[3][2] = {{X3D_OP_DERZ, X3D_OP_DERZP}, {X3D_OP_DERZ, X3D_OP_DERZP}, {X3D_OP_DERZP, X3D_OP_DERZ}};
  for (int k = 0; k < 3; ++k) {
    const double* uc = u[k];
    X3D_TRY(ew(c, EW_MUL, ta, ux, uc, nullptr, 0));
    X3D_TRY(launch_line(c, 0, K_D, opx[k][0], ta, td, true));
    X3D_TRY(launch_line(c, 0, K_D, opx[k][1], uc, ta, true));
    X3D_TRY(ew(c, EW_FMA_ADD, du[k], ux, ta, td, 0));
    X3D_TRY(ew(c, EW_MUL, td, uc, uy, nullptr, 0));
    X3D_TRY(launch_line(c, 1, K_D, opy[k][0], td, tg, true));
    X3D_TRY(launch_line(c, 1, K_D, opy[k][1], uc, td, true));
    X3D_TRY(ew(c, EW_FMA_ADD, tg, uy, td, tg, 0));
    X3D_TRY(ew(c, EW_MUL, td, uc, uz, nullptr, 0));
    X3D_TRY(launch_line(c, 2, K_D, opz[k][0], td, t3, true));
    X3D_TRY(launch_line(c, 2, K_D, opz[k][1], uc, td, true));
    X3D_TRY(ew(c, EW_FMA_ADD, t3, uz, td, t3, 0));
    X3D_TRY(ew(c, EW_ADD, tg, tg, t3, nullptr, 0));
    X3D_TRY(ew(c, EW_ADD, du[k], du[k], tg, nullptr, 0));
  }
  return X3D_OK;
}

static int ref_divergence(x3d_ctx* c, double* pp3, const double* ux, const double* uy, const double* uz)
{
  X3D_TRY(ensure_work(c));
  double *pp1 = c->work[0], *pgy1 = c->work[1], *pgz1 = c->work[2];
  double *upi2 = c->work[3], *duydypi2 = c->work[4], *po3 = c->work[5];
  X3D_TRY(launch_line(c, 0, K_DVP, X3D_OP_DERXVP, ux, pp1, true));
  X3D_TRY(launch_line(c, 0, K_IVP, X3D_OP_INTXVP, uy, pgy1, true));
  X3D_TRY(launch_line(c, 0, K_IVP, X3D_OP_INTXVP, uz, pgz1, true));
  X3D_TRY(launch_line(c, 1, K_IVP, X3D_OP_INTYVP, pp1, upi2, true));
  X3D_TRY(launch_line(c, 1, K_DVP, X3D_OP_DERYVP, pgy1, duydypi2, true));
  X3D_TRY(ew(c, EW_ADD, duydypi2, duydypi2, upi2, nullptr, 0));
  X3D_TRY(launch_line(c, 1, K_IVP, X3D_OP_INTYVP, pgz1, upi2, true));
  X3D_TRY(launch_line(c, 2, K_IVP, X3D_OP_INTZVP, duydypi2, pp3, true));
  X3D_TRY(launch_line(c, 2, K_DVP, X3D_OP_DERZVP, upi2, po3, true));
  X3D_TRY(ew(c, EW_ADD, pp3, pp3, po3, nullptr, 0));
  return X3D_OK;
}

static int ref_gradp(x3d_ctx* c, double* px, double* py, double* pz, const double* pp3)
{
  X3D_TRY(ensure_work(c));
  double *ppi3 = c->work[0], *pgz3 = c->work[1], *ppi2 = c->work[2], *pgy2 = c->work[3];
  double* pgzi2 = c->work[4];
  X3D_TRY(launch_line(c, 2, K_IPV, X3D_OP_INTZPV, pp3, ppi3, true));
  X3D_TRY(launch_line(c, 2, K_DPV, X3D_OP_DERZPV, pp3, pgz3, true));
  X3D_TRY(launch_line(c, 1, K_IPV, X3D_OP_INTYPV, ppi3, ppi2, true));
  X3D_TRY(launch_line(c, 1, K_DPV, X3D_OP_DERYPV, ppi3, pgy2, true));
  X3D_TRY(launch_line(c, 1, K_IPV, X3D_OP_INTYPV, pgz3, pgzi2, true));
  X3D_TRY(launch_line(c, 0, K_DPV, X3D_OP_DERXPV, ppi2, px, true));
  X3D_TRY(launch_line(c, 0, K_IPV, X3D_OP_INTXPV, pgy2, py, true));
  X3D_TRY(launch_line(c, 0, K_IPV, X3D_OP_INTXPV, pgzi2, pz, true));
  return X3D_OK;
}
}  // namespace x3d

extern "C" {

int x3d_momentum_rhs(x3d_ctx* c, double* dux, double* duy, double* duz, const double* ux,
                     const double* uy, const double* uz)
{
  if (!c || !dux || !duy || !duz || !ux || !uy || !uz) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  if (!fused_ok(c)) return ref_momentum_rhs(c, dux, duy, duz, ux, uy, uz);
  // the fused passes do not write u when euler = 0
  return fused_momentum<double>(c, 0.0, 0, const_cast<double*>(ux), const_cast<double*>(uy),
                        const_cast<double*>(uz), dux, duy, duz);
}

int x3d_int_time(x3d_ctx* c, double gdt, double* ux, double* uy, double* uz, const double* dux,
                 const double* duy, const double* duz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz) return X3D_ERR_ARG;
  X3D_TRY(ew(c, EW_AXPY, ux, dux, ux, nullptr, gdt));
  X3D_TRY(ew(c, EW_AXPY, uy, duy, uy, nullptr, gdt));
  X3D_TRY(ew(c, EW_AXPY, uz, duz, uz, nullptr, gdt));
  return X3D_OK;
}

int x3d_divergence(x3d_ctx* c, double* pp3, const double* ux, const double* uy, const double* uz)
{
  if (!c || !pp3 || !ux || !uy || !uz) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  if (!fused_ok(c)) return ref_divergence(c, pp3, ux, uy, uz);
  return fused_divergence<double>(c, pp3, ux, uy, uz);
}

int x3d_gradp(x3d_ctx* c, double* px, double* py, double* pz, const double* pp3)
{
  if (!c || !px || !py || !pz || !pp3) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  if (!fused_ok(c)) return ref_gradp(c, px, py, pz, pp3);
  return fused_gradp<double>(c, px, py, pz, pp3, nullptr, nullptr, nullptr, 0);
}

int x3d_cor_vel(x3d_ctx* c, double* ux, double* uy, double* uz, const double* px, const double* py,
                const double* pz)
{
  if (!c || !ux || !uy || !uz || !px || !py || !pz) return X3D_ERR_ARG;
  X3D_TRY(ew(c, EW_SUB, ux, ux, px, nullptr, 0));
  X3D_TRY(ew(c, EW_SUB, uy, uy, py, nullptr, 0));
  X3D_TRY(ew(c, EW_SUB, uz, uz, pz, nullptr, 0));
  return X3D_OK;
}

int x3d_solve_poisson(x3d_ctx* c, double* pp3, double* px, double* py, double* pz, const double* ux,
                      const double* uy, const double* uz)
{
  X3D_TRY(x3d_divergence(c, pp3, ux, uy, uz));
  X3D_TRY(x3d_poisson(c, pp3));
  X3D_TRY(x3d_gradp(c, px, py, pz, pp3));
  return X3D_OK;
}

int x3d_step(x3d_ctx* c, double gdt, double* ux, double* uy, double* uz, double* dux, double* duy,
             double* duz, double* pp3, double* px, double* py, double* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  if (!fused_ok(c)) {
    X3D_TRY(x3d_momentum_rhs(c, dux, duy, duz, ux, uy, uz));
    X3D_TRY(x3d_int_time(c, gdt, ux, uy, uz, dux, duy, duz));
    X3D_TRY(x3d_solve_poisson(c, pp3, px, py, pz, ux, uy, uz));
    X3D_TRY(x3d_cor_vel(c, ux, uy, uz, px, py, pz));
    return X3D_OK;
  }
  // transeq with the Euler update folded into its x pass, divergence, Poisson,
  // gradp with cor_vel folded into its x pass
  X3D_TRY(fused_momentum<double>(c, gdt, 1, ux, uy, uz, dux, duy, duz));
  X3D_TRY(fused_divergence<double>(c, pp3, ux, uy, uz));
  X3D_TRY(poisson_solve(c, pp3));
  X3D_TRY(mark_pass(c, 7));
  X3D_TRY(fused_gradp<double>(c, px, py, pz, pp3, ux, uy, uz, 1));
  return X3D_OK;
}

namespace {
// Pass names and algorithmic HBM bytes per grid point (FP64: each field read or
// written once per pass), in launch order.  transeq Z reads u and writes du =
// its part; transeq Y reads u, du and writes tmp = du + its part; transeq X
// reads u, tmp and writes du, u; gradp X
// reads three gradients and u and writes p gradients and u.  The Poisson figure
// is one r2c read/write, one in-place spectral update and one c2r read/write.
struct PassInfo {
  const char* name;
  int fields;
};
constexpr PassInfo kPass[X3D_NPASS] = {
    {"transeq_z", 6}, {"transeq_y", 9}, {"transeq_x", 12}, {"divergence_x", 6}, {"divergence_y", 5},
    {"divergence_z", 3}, {"poisson", 6}, {"gradp_z", 3}, {"gradp_y", 5}, {"gradp_x", 12}};
}  // namespace

// ---- precision form (FP32 fields, FP32 line solves) -------------------------
int x3d_derx_f32(x3d_ctx* c, float* t, const float* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line_f32(c, 0, K_D, op, u, t);
}
int x3d_dery_f32(x3d_ctx* c, float* t, const float* u, int op, const float* ppy, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  if (!rc && c->istret && ppy) {
    const std::vector<double> v(ppy, ppy + c->ny);
    rc = stretch_arg(c, 0, v.data());
  }
  return rc ? rc : launch_line_f32(c, 1, K_D, op, u, t);
}
int x3d_derz_f32(x3d_ctx* c, float* t, const float* u, int op, int nx, int ny, int nz)
{
  int rc = check3(c, nx, ny, nz);
  return rc ? rc : launch_line_f32(c, 2, K_D, op, u, t);
}

int x3d_apply_f32(x3d_ctx* c, int op, float* out, const float* in)
{
  if (!c || op < 0 || op >= X3D_NOPS_ALL) return X3D_ERR_ARG;
  // operator ids: 0-5 D (x, x, y, y, z, z); then Dvp, Dpv, Ivp, Ipv, each
  // x, y, z; then the second derivatives x, y, z
  const int dir = op < 6 ? op / 2 : (op - 6) % 3;
  static const int kinds[5] = {K_DVP, K_DPV, K_IVP, K_IPV, K_DD};
  const int kind = op < 6 ? K_D : kinds[(op - 6) / 3];
  return launch_line_f32(c, dir, kind, op, in, out);
}

int x3d_poisson_f32(x3d_ctx* c, float* rhs)
{
  if (!c || !rhs) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  return poisson_solve_f32(c, rhs);
}

static bool fp32_ok(const x3d_ctx* c)
{
  return fused_ok(c) && c->fplan32[0].M && c->fplan32[1].M && c->fplan32[2].M;
}

int x3d_step_f32(x3d_ctx* c, double gdt, float* ux, float* uy, float* uz, float* dux, float* duy, float* duz,
                 float* pp3, float* px, float* py, float* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  if (!fp32_ok(c)) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(fused_momentum<float>(c, (float)gdt, 1, ux, uy, uz, dux, duy, duz));
  X3D_TRY(fused_divergence<float>(c, pp3, ux, uy, uz));
  X3D_TRY(poisson_solve_f32(c, pp3));
  X3D_TRY(mark_pass(c, 7));
  X3D_TRY(fused_gradp<float>(c, px, py, pz, pp3, ux, uy, uz, 1));
  return X3D_OK;
}

int x3d_step_timed_f32(x3d_ctx* c, double gdt, float* ux, float* uy, float* uz, float* dux, float* duy,
                       float* duz, float* pp3, float* px, float* py, float* pz, float* pass_ms)
{
  if (!c || !pass_ms) return X3D_ERR_ARG;
  if (!fp32_ok(c)) return X3D_ERR_UNSUPPORTED;
  for (int i = 0; i <= X3D_NPASS; ++i)
    if (!c->pass_ev[i]) X3D_CHECK_HIP(hipEventCreate(&c->pass_ev[i]));
  X3D_CHECK_HIP(hipEventRecord(c->pass_ev[0], c->stream));
  c->timing = true;
  const int rc = x3d_step_f32(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz);
  c->timing = false;
  if (rc != X3D_OK) return rc;
  X3D_CHECK_HIP(hipEventSynchronize(c->pass_ev[X3D_NPASS]));
  for (int i = 0; i < X3D_NPASS; ++i)
    X3D_CHECK_HIP(hipEventElapsedTime(&pass_ms[i], c->pass_ev[i], c->pass_ev[i + 1]));
  // the pass names follow the z-first transeq order
  if (!transeq_zfirst<float>()) std::swap(pass_ms[0], pass_ms[2]);
  return X3D_OK;
}

// ---- captured step ----------------------------------------------------------
static void graph_release(x3d_ctx* c)
{
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

static int fft_streams(x3d_ctx* c, hipStream_t s)
{
  if (c->fft_ready && (hipfftSetStream(c->fft_r2c, s) != HIPFFT_SUCCESS ||
                       hipfftSetStream(c->fft_c2r, s) != HIPFFT_SUCCESS))
    return X3D_ERR_FFT;
  if (c->fft32_ready && (hipfftSetStream(c->fft_r2c32, s) != HIPFFT_SUCCESS ||
                         hipfftSetStream(c->fft_c2r32, s) != HIPFFT_SUCCESS))
    return X3D_ERR_FFT;
  return X3D_OK;
}

// Record body's launches on a private stream into a graph; every allocation
// (scratch fields, FFT plans) is made before the capture starts.
// Start recording the context's launches on a private capture stream (the
// caller's stream is put back at the end).  Every allocation the recorded
// calls need (the FP64 step's scratch fields, FFT plans) exists by now.
static int graph_begin(x3d_ctx* c)
{
  if (c->cap_stream) return X3D_ERR_ARG;  // already capturing
  graph_release(c);
  X3D_TRY(ensure_work(c));
  hipStream_t cs = nullptr;
  X3D_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  const hipStream_t user = c->stream;
  c->stream = cs;
  int rc = fft_streams(c, cs);
  if (!rc && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = X3D_ERR_HIP;
  if (rc) {
    c->stream = user;
    (void)fft_streams(c, user);
    (void)hipStreamDestroy(cs);
    return rc;
  }
  c->cap_stream = cs;
  c->cap_user = user;
  return X3D_OK;
}

// Stop recording and instantiate the graph; body_rc is the status of the
// recorded calls (a failed one discards the graph).
static int graph_end(x3d_ctx* c, int body_rc)
{
  if (!c->cap_stream) return X3D_ERR_ARG;
  hipStream_t cs = c->cap_stream;
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(cs, &g);
  int rc = body_rc;
  if (!rc && ee != hipSuccess) rc = hip_status(ee);
  c->stream = c->cap_user;
  const int rs = fft_streams(c, c->cap_user);
  (void)hipStreamDestroy(cs);
  c->cap_stream = nullptr;
  c->cap_user = nullptr;
  if (!rc) rc = rs;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    // a call refused during the capture left its error as the runtime's last
    // error; clear it so that the next launch check does not report it
    (void)hipGetLastError();
    return rc;
  }
  const hipError_t e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    c->graph_exec = nullptr;
    return hip_status(e);
  }
  c->graph = g;
  return X3D_OK;
}

static int capture_graph(x3d_ctx* c, const std::function<int()>& body)
{
  X3D_TRY(graph_begin(c));
  return graph_end(c, body());
}

int x3d_graph_begin(x3d_ctx* c)
{
  if (!c) return X3D_ERR_ARG;
  return graph_begin(c);
}

int x3d_graph_end(x3d_ctx* c)
{
  if (!c) return X3D_ERR_ARG;
  return graph_end(c, X3D_OK);
}

int x3d_graph_capture_step(x3d_ctx* c, double gdt, double* ux, double* uy, double* uz, double* dux,
                           double* duy, double* duz, double* pp3, double* px, double* py, double* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  if (!fused_ok(c)) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(ensure_work(c));
  return capture_graph(c, [&] { return x3d_step(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz); });
}

int x3d_graph_capture_step_f32(x3d_ctx* c, double gdt, float* ux, float* uy, float* uz, float* dux,
                               float* duy, float* duz, float* pp3, float* px, float* py, float* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  if (!fp32_ok(c)) return X3D_ERR_UNSUPPORTED;
  X3D_TRY(ensure_work(c));
  X3D_TRY(poisson_prepare_f32(c));
  return capture_graph(c, [&] { return x3d_step_f32(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz); });
}

int x3d_graph_launch(x3d_ctx* c)
{
  if (!c) return X3D_ERR_ARG;
  if (!c->graph_exec) return X3D_ERR_UNSUPPORTED;
  X3D_CHECK_HIP(hipGraphLaunch(c->graph_exec, c->stream));
  return X3D_OK;
}

// x3d_step with the 8 work fields in `scratch` (at its device address from
// scratch_setup) instead of the context's own, which are swapped back after
// the call.  The context's FFT plans, spectrum and tables come from
// x3d_create, so nothing here allocates.  Every work field is written by a
// pass before a later pass reads it, so the array needs no clearing.
int x3d_step_noalloc(x3d_ctx* c, double gdt, double* ux, double* uy, double* uz, double* dux, double* duy,
                     double* duz, double* pp3, double* px, double* py, double* pz)
{
  if (!c || !ux || !uy || !uz || !dux || !duy || !duz || !pp3 || !px || !py || !pz) return X3D_ERR_ARG;
  if (!all_periodic(c)) return X3D_ERR_UNSUPPORTED;
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  if (N > (size_t)X3D_NOALLOC_MAX_POINTS) return X3D_ERR_SIZE;
  double* base = scratch_of(c->device);
  if (!base) return X3D_ERR_HIP;
  double* own[8];
  std::memcpy(own, c->work, sizeof(own));
  const size_t own_elems = c->work_elems;
  for (int i = 0; i < 8; ++i) c->work[i] = base + (size_t)i * N;
  c->work_elems = N;
  const int rc = x3d_step(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz);
  std::memcpy(c->work, own, sizeof(own));
  c->work_elems = own_elems;
  return rc;
}

const char* x3d_pass_name(int i) { return (i >= 0 && i < X3D_NPASS) ? kPass[i].name : nullptr; }

double x3d_pass_bytes_per_point(int i) { return (i >= 0 && i < X3D_NPASS) ? 8.0 * kPass[i].fields : 0.0; }

int x3d_step_timed(x3d_ctx* c, double gdt, double* ux, double* uy, double* uz, double* dux, double* duy,
                   double* duz, double* pp3, double* px, double* py, double* pz, float* pass_ms)
{
  if (!c || !pass_ms) return X3D_ERR_ARG;
  if (!fused_ok(c)) return X3D_ERR_UNSUPPORTED;
  for (int i = 0; i <= X3D_NPASS; ++i)
    if (!c->pass_ev[i]) X3D_CHECK_HIP(hipEventCreate(&c->pass_ev[i]));
  X3D_CHECK_HIP(hipEventRecord(c->pass_ev[0], c->stream));
  c->timing = true;
  const int rc = x3d_step(c, gdt, ux, uy, uz, dux, duy, duz, pp3, px, py, pz);
  c->timing = false;
  if (rc != X3D_OK) return rc;
  X3D_CHECK_HIP(hipEventSynchronize(c->pass_ev[X3D_NPASS]));
  for (int i = 0; i < X3D_NPASS; ++i)
    X3D_CHECK_HIP(hipEventElapsedTime(&pass_ms[i], c->pass_ev[i], c->pass_ev[i + 1]));
  return X3D_OK;
}

// ---- test mode and fields ----------------------------------------------------
int x3d_test_err(x3d_ctx* c, const double* d, int which, double* err)
{
  if (!c || !d || !err || which < 0 || which > 2) return X3D_ERR_ARG;
  double s, m;
  X3D_TRY(reduce(c, d, which, &s, &m));
  *err = std::sqrt(s / c->nx / c->ny / c->nz);
  return X3D_OK;
}

int x3d_div_stats(x3d_ctx* c, const double* pp3, double* tmax, double* tmean)
{
  if (!c || !pp3 || !tmax || !tmean) return X3D_ERR_ARG;
  double s, m;
  X3D_TRY(reduce(c, pp3, 3, &s, &m));
  *tmax = m;
  *tmean = s / ((double)c->nx * c->ny * c->nz);
  return X3D_OK;
}

int x3d_init_fields(x3d_ctx* c, int kind, uint64_t seed, double amp, double* ux, double* uy,
                    double* uz)
{
  if (!c || !ux || !uy || !uz || kind < 0 || kind > 2) return X3D_ERR_ARG;
  int hk[3 * 8 * 3];
  double hph[3 * 8];
  const double twopi = 2.0 * 3.14159265358979323846;
  for (int cpt = 0; cpt < 3; ++cpt)
    for (int m = 0; m < 8; ++m) {
      int* kk = hk + 3 * (8 * cpt + m);
      for (int d = 0; d < 3; ++d)
        kk[d] = (int)std::floor(host_uniform01(seed, 16 + cpt, (uint64_t)(4 * m + d)) * 9.0) - 4;
      if (!kk[0] && !kk[1] && !kk[2]) kk[0] = 1;
      hph[8 * cpt + m] = twopi * host_uniform01(seed, 16 + cpt, (uint64_t)(4 * m + 3));
    }
  int* dk = nullptr;
  double* dph = nullptr;
  X3D_CHECK_HIP(hipMalloc(&dk, sizeof(hk)));
  X3D_CHECK_HIP(hipMalloc(&dph, sizeof(hph)));
  X3D_CHECK_HIP(hipMemcpy(dk, hk, sizeof(hk), hipMemcpyHostToDevice));
  X3D_CHECK_HIP(hipMemcpy(dph, hph, sizeof(hph), hipMemcpyHostToDevice));
  GridInfo g{c->nx, c->ny, c->nz, c->dx, c->dy, c->dz, c->xlx, c->yly, c->zlz};
  hipLaunchKernelGGL(k_init_fields, dim3(4096), dim3(256), 0, c->stream, kind, seed, amp, g, dk, dph, ux, uy, uz);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dk);
  (void)hipFree(dph);
  X3D_CHECK_HIP(e);
  return X3D_OK;
}

}  // extern "C"
