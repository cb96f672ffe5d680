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
// free: here hipFFT halves x (the contiguous axis).
//
// K = xk2 (Ty Tz)^2 + yk2 (Tx Tz)^2 + zk2 (Tx Ty)^2 is evaluated per mode
// from 1-D tables in the reference's operation order with contraction off,
// so it is bit-identical to kxyz (src/poisson.f90:673-695) and the null-mode
// test makes the same decisions.
#include <cmath>

#include "x3d_internal.h"

namespace x3d {

static const double PI = 3.14159265358979323846264338327950288;

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
  if (hipfftPlan3d(&c->fft_r2c, c->nz, c->ny, c->nx, HIPFFT_D2Z) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
  if (hipfftPlan3d(&c->fft_c2r, c->nz, c->ny, c->nx, HIPFFT_Z2D) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
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
  c->spec = nullptr;
  c->ktab = nullptr;
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
    if (hipfftPlan3d(&c->fft_r2c32, c->nz, c->ny, c->nx, HIPFFT_R2C) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
    if (hipfftPlan3d(&c->fft_c2r32, c->nz, c->ny, c->nx, HIPFFT_C2R) != HIPFFT_SUCCESS) {
      hipfftDestroy(c->fft_r2c32);
      return X3D_ERR_FFT;
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
  if (hipfftExecR2C(c->fft_r2c32, rhs, (hipfftComplex*)spec) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
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
  if (hipfftExecD2Z(c->fft_r2c, rhs, (hipfftDoubleComplex*)c->spec) != HIPFFT_SUCCESS) return X3D_ERR_FFT;
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

}  // namespace x3d

extern "C" int x3d_poisson_tables(const x3d_ctx* c, double* ax, double* bx, double* ay, double* by,
                                  double* az, double* bz, double* kre)
{
  if (!c) return X3D_ERR_ARG;
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
