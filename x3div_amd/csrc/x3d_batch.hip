// The batching form: many independent periodic grids of uneven sizes stepped
// together.  Every grid is one instance of the reference's problem
// (src/xcompact3d.f90:80-95 on its own nx, ny, nz); the batch owns their
// fields, groups the grids by line tiling so that each step pass is one
// persistent launch per group, and can record the whole batched step in a
// HIP graph.  An autotuner picks the tiling of each direction and line length
// by timing the batched step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>

#include "x3d_internal.h"

using namespace x3d;

#define X3D_TRY_B(expr) do { const int _r = (expr); if (_r) return _r; } while (0)

namespace {

__global__ void k_to_f32(float* __restrict__ out, const double* __restrict__ in, long long n)
{
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    out[q] = (float)in[q];
}

void graph_release(x3d_batch* b)
{
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

int set_streams(x3d_batch* b, hipStream_t s)
{
  for (x3d_ctx* c : b->g) {
    const int rc = x3d_set_stream(c, s);
    if (rc) return rc;
  }
  return X3D_OK;
}

}  // namespace

int x3d::batch_fork(x3d_batch* b)
{
  X3D_CHECK_HIP(hipEventRecord(b->fork_ev, b->stream));
  for (hipStream_t s : b->pstreams) X3D_CHECK_HIP(hipStreamWaitEvent(s, b->fork_ev, 0));
  return X3D_OK;
}

int x3d::batch_join(x3d_batch* b, int rc)
{
  for (size_t k = 0; k < b->pstreams.size(); ++k) {
    if (hipEventRecord(b->join_ev[k], b->pstreams[k]) != hipSuccess && !rc) rc = X3D_ERR_HIP;
    if (hipStreamWaitEvent(b->stream, b->join_ev[k], 0) != hipSuccess && !rc) rc = X3D_ERR_HIP;
  }
  return rc;
}

namespace {

// Poisson of every grid.  The grids the line kernels take run grouped: one
// persistent launch per stage and line length over all of them.  With no
// other grids, a stage's launches run side by side on the side streams.  The
// others (psolo) solve one by one; a grid's solve is a few small kernels that
// leave most compute units idle, so they go round-robin onto the side
// streams, forked from the batch stream and joined back to it, and run beside
// each other and the grouped launches (then on the batch stream).  Each such
// context runs on its side stream for the call and is given the batch stream
// back afterwards.
int batch_poisson(x3d_batch* b)
{
  if (b->psolo.empty()) {
    const int ns = b->overlap ? (int)b->pstreams.size() : 0;
    return poisson_batch_run(b, ns);
  }
  const int ns = (int)b->pstreams.size();
  if (ns > 0) X3D_TRY_B(batch_fork(b));
  int rc = poisson_batch_run(b, 0);
  for (size_t k = 0; !rc && k < b->psolo.size(); ++k) {
    const int i = b->psolo[k];
    x3d_ctx* c = b->g[i];
    if (ns > 0) rc = x3d_set_stream(c, b->pstreams[k % ns]);
    void* pp = b->field(6, i);
    if (!rc) rc = b->prec == 8 ? x3d_poisson(c, static_cast<double*>(pp)) : x3d_poisson_f32(c, static_cast<float*>(pp));
  }
  if (ns > 0) {
    // join every side stream, also after an error, so that no capture or
    // later batch work is left unordered against them
    rc = batch_join(b, rc);
    const int rs = set_streams(b, b->stream);
    if (!rc) rc = rs;
  }
  return rc;
}

// One step of every grid: transeq (Euler folded in), divergence, Poisson,
// gradp with cor_vel.
// The batched passes run the circulant solvers of the library's own operator
// tables and carry no y-stretching vectors: a grid whose context was given
// stretching (x3d_set_stretching) or tables of the caller's own
// (x3d_operator_set), through x3d_batch_context, is refused on every step and
// graph launch, not only when the launches are built.
bool batch_refused(const x3d_batch* b)
{
  for (const x3d_ctx* c : b->g)
    if (c->istret || !fused_ok(c)) return true;
  return false;
}

int batch_step_body(x3d_batch* b, double gdt)
{
  if (batch_refused(b)) return X3D_ERR_UNSUPPORTED;
  if (!b->built || b->built_gdt != gdt) {
    const int rc = b->prec == 8 ? batch_build<double>(b, gdt) : batch_build<float>(b, gdt);
    if (rc) return rc;
  }
  auto passes = [b](int which) { return b->prec == 8 ? batch_passes<double>(b, which) : batch_passes<float>(b, which); };
  int rc = passes(0);
  if (!rc) rc = passes(1);
  if (!rc) rc = batch_poisson(b);
  if (!rc) rc = passes(2);
  return rc;
}

// Line lengths that this y/z tiling candidate fits, for the autotuner: the
// chunk length M, chunk groups per wavefront G, line groups per work-group Wl.
// LDS bytes of the largest pass (transeq, with its keep slots) for a y/z
// tiling: solver aggregates (4 Wc + 2 lines per work-group when Wc > 1) plus
// (2 M + 4) slots per thread.
size_t plan_lds_bytes(const LinePlan& p, int prec)
{
  const size_t lpw = (size_t)p.Wl * p.LW, threads = 64u * p.Wc * p.Wl;
  const size_t solver = p.Wc > 1 ? 4 * p.Wc * lpw + 2 * lpw : 0;
  return (size_t)prec * (solver + threads * (size_t)(2 * p.M + 4));
}

// x-line tilings: chunks of M points, C = n / M chunks per line (a power of
// two up to 64, so that a wavefront holds 64 / C whole lines), Wl wavefronts
// per work-group, each with its LDS staging tile.
std::vector<LinePlan> candidates_x(int n, int prec, size_t lds_max)
{
  static const int ms64[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};
  static const int ms32[] = {1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 32};
  std::vector<LinePlan> out;
  const int* ms = prec == 8 ? ms64 : ms32;
  const int nm = prec == 8 ? 9 : 11;
  for (int k = 0; k < nm; ++k) {
    const int m = ms[k];
    if (n % m) continue;
    const int c = n / m;
    if (c > 64 || (c & (c - 1))) continue;
    for (int wl : {1, 2, 4, 8}) {
      LinePlan p;
      p.n = n; p.M = m; p.C = c; p.G = c; p.Wc = 1; p.Wl = wl; p.LW = 64 / c; p.line_fast = false;
      const size_t pitch = (m & 1) ? m : m + 1;
      if ((size_t)prec * wl * 64 * pitch <= lds_max) out.push_back(p);
    }
  }
  return out;
}

std::vector<LinePlan> candidates(int n, int prec, size_t lds_max)
{
  static const int ms64[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};
  static const int ms32[] = {1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 32};
  std::vector<LinePlan> out;
  const int* ms = prec == 8 ? ms64 : ms32;
  const int nm = prec == 8 ? 9 : 11;
  for (int k = 0; k < nm; ++k) {
    const int m = ms[k];
    if (n % m) continue;
    const int c = n / m;
    if (c > 64) continue;
    for (int g : {1, 2, 4, 8, 16}) {
      if (c % g) continue;
      const int wc = c / g;
      for (int wl : {1, 2, 4}) {
        if (wc * wl > 8) continue;
        LinePlan p;
        p.n = n; p.M = m; p.C = c; p.G = g; p.Wc = wc; p.Wl = wl; p.LW = 64 / g; p.line_fast = true;
        if (plan_lds_bytes(p, prec) <= lds_max) out.push_back(p);
      }
    }
  }
  return out;
}

}  // namespace

extern "C" {

int x3d_batch_create(int ngrids, const int* dims, const double* lengths, int precision, x3d_batch** out)
{
  if (!out || !dims || ngrids <= 0 || (precision != 8 && precision != 4)) return X3D_ERR_ARG;
  *out = nullptr;
  x3d_batch* b = new x3d_batch;
  b->ng = ngrids;
  b->prec = precision;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  for (int i = 0; i < ngrids; ++i) {
    const double tau = 2.0 * M_PI;
    const double lx = lengths ? lengths[3 * i] : tau, ly = lengths ? lengths[3 * i + 1] : tau,
                 lz = lengths ? lengths[3 * i + 2] : tau;
    x3d_ctx* c = nullptr;
    const int rc = x3d_create(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], lx, ly, lz, &c);
    if (rc) {
      x3d_batch_destroy(b);
      return rc;
    }
    b->g.push_back(c);
    b->off.push_back(b->total);
    b->total += (size_t)dims[3 * i] * dims[3 * i + 1] * dims[3 * i + 2];
  }
  if (hipMalloc(&b->store, 10 * b->total * (size_t)precision) != hipSuccess) {
    x3d_batch_destroy(b);
    return X3D_ERR_ALLOC;
  }
  // the grouped Poisson launches, and side streams for the grids they do not
  // take (default priority; X3D_BATCH_PSTREAMS sets their number, 0 runs
  // those solves in order on the batch stream)
  if (const int rc = poisson_batch_build(b)) {
    x3d_batch_destroy(b);
    return rc;
  }
  int nps = 4;
  if (const char* ps = std::getenv("X3D_BATCH_PSTREAMS")) nps = std::max(0, std::min(16, std::atoi(ps)));
  if (const char* ov = std::getenv("X3D_BATCH_OVERLAP")) b->overlap = std::atoi(ov) != 0;
  if (nps > 0) {
    if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
      x3d_batch_destroy(b);
      return X3D_ERR_HIP;
    }
    for (int k = 0; k < nps; ++k) {
      hipStream_t s = nullptr;
      hipEvent_t e = nullptr;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        x3d_batch_destroy(b);
        return X3D_ERR_HIP;
      }
      b->pstreams.push_back(s);
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        x3d_batch_destroy(b);
        return X3D_ERR_HIP;
      }
      b->join_ev.push_back(e);
    }
  }
  (void)hipMemset(b->store, 0, 10 * b->total * (size_t)precision);
  *out = b;
  return X3D_OK;
}

int x3d_batch_destroy(x3d_batch* b)
{
  if (!b) return X3D_ERR_ARG;
  graph_release(b);
  for (hipStream_t s : b->pstreams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  for (hipEvent_t e : b->join_ev) (void)hipEventDestroy(e);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (x3d_ctx* c : b->g) x3d_destroy(c);
  if (b->store) (void)hipFree(b->store);
  if (b->dargs) (void)hipFree(b->dargs);
  if (b->pargs) (void)hipFree(b->pargs);
  delete b;
  return X3D_OK;
}

int x3d_batch_set_stream(x3d_batch* b, void* stream)
{
  if (!b) return X3D_ERR_ARG;
  b->stream = (hipStream_t)stream;
  return set_streams(b, b->stream);
}

int x3d_batch_sizes(const x3d_batch* b, int* ngrids, long long* total_points)
{
  if (!b) return X3D_ERR_ARG;
  if (ngrids) *ngrids = b->ng;
  if (total_points) *total_points = (long long)b->total;
  return X3D_OK;
}

int x3d_batch_field(const x3d_batch* b, int grid, int field, void** dptr)
{
  if (!b || !dptr || grid < 0 || grid >= b->ng || field < 0 || field >= 10) return X3D_ERR_ARG;
  *dptr = b->field(field, grid);
  return X3D_OK;
}

int x3d_batch_context(const x3d_batch* b, int grid, x3d_ctx** ctx)
{
  if (!b || !ctx || grid < 0 || grid >= b->ng) return X3D_ERR_ARG;
  *ctx = b->g[grid];
  return X3D_OK;
}

int x3d_batch_init_fields(x3d_batch* b, int kind, uint64_t seed, double amp)
{
  if (!b) return X3D_ERR_ARG;
  // FP32 batches: each grid's FP64 fields go through one staging buffer
  double* tmp = nullptr;
  if (b->prec == 4) {
    size_t nmax = 0;
    for (x3d_ctx* c : b->g) nmax = std::max(nmax, (size_t)c->nx * c->ny * c->nz);
    if (hipMalloc(&tmp, 3 * nmax * sizeof(double)) != hipSuccess) return X3D_ERR_ALLOC;
  }
  int rc = X3D_OK;
  for (int i = 0; i < b->ng && !rc; ++i) {
    x3d_ctx* c = b->g[i];
    const size_t n = (size_t)c->nx * c->ny * c->nz;
    if (b->prec == 8) {
      rc = x3d_init_fields(c, kind, seed + (uint64_t)i, amp, static_cast<double*>(b->field(0, i)),
                           static_cast<double*>(b->field(1, i)), static_cast<double*>(b->field(2, i)));
      continue;
    }
    rc = x3d_init_fields(c, kind, seed + (uint64_t)i, amp, tmp, tmp + n, tmp + 2 * n);
    for (int k = 0; k < 3 && !rc; ++k) {
      hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, b->stream, static_cast<float*>(b->field(k, i)),
                         tmp + k * n, (long long)n);
      if (hipGetLastError() != hipSuccess) rc = X3D_ERR_HIP;
    }
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess && !rc) rc = X3D_ERR_HIP;
  if (tmp) (void)hipFree(tmp);
  return rc;
}

int x3d_batch_step(x3d_batch* b, double gdt)
{
  if (!b) return X3D_ERR_ARG;
  return batch_step_body(b, gdt);
}

int x3d_batch_graph_capture(x3d_batch* b, double gdt)
{
  if (!b) return X3D_ERR_ARG;
  graph_release(b);
  // everything the step allocates or uploads happens before the capture
  int rc = b->prec == 8 ? batch_build<double>(b, gdt) : batch_build<float>(b, gdt);
  for (int i = 0; !rc && i < b->ng; ++i)
    if (b->prec == 4) rc = poisson_prepare_f32(b->g[i]);
  if (rc) return rc;
  hipStream_t cs = nullptr;
  X3D_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipStream_t user = b->stream;
  b->stream = cs;
  rc = set_streams(b, cs);
  hipGraph_t g = nullptr;
  if (!rc) {
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = X3D_ERR_HIP;
    } else {
      rc = batch_step_body(b, gdt);
      const hipError_t e = hipStreamEndCapture(cs, &g);
      if (!rc && e != hipSuccess) rc = hip_status(e);
    }
  }
  b->stream = user;
  const int rs = set_streams(b, user);
  (void)hipStreamDestroy(cs);
  if (!rc) rc = rs;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  const hipError_t e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    b->graph_exec = nullptr;
    return hip_status(e);
  }
  b->graph = g;
  return X3D_OK;
}

int x3d_batch_graph_launch(x3d_batch* b)
{
  if (!b) return X3D_ERR_ARG;
  if (!b->graph_exec || batch_refused(b)) return X3D_ERR_UNSUPPORTED;
  X3D_CHECK_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  return X3D_OK;
}

int x3d_batch_autotune(x3d_batch* b, int reps, char* report, size_t report_len)
{
  if (!b || reps <= 0) return X3D_ERR_ARG;
  graph_release(b);
  std::string log;
  auto time_step = [b, reps](double* ms) -> int {
    int rc = batch_step_body(b, 0.0);  // warm-up, and rebuild after a plan change
    if (rc) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return X3D_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return X3D_ERR_HIP; }
    (void)hipEventRecord(e0, b->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = batch_step_body(b, 0.0);
    (void)hipEventRecord(e1, b->stream);
    float t = 0.f;
    if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = X3D_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
  };
  int dev = 0, lds = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const size_t lds_max = (size_t)lds;
  double best_ms = 0.0;
  int rc = time_step(&best_ms);
  if (rc) return rc;
  char line[256];
  std::snprintf(line, sizeof line, "start %.4f ms/step\n", best_ms);
  log += line;
  // coordinate descent: one (direction, line length) at a time, every grid of
  // that length switched together
  for (int dir = 0; dir <= 2; ++dir) {
    std::map<int, std::vector<int>> byn;
    for (int i = 0; i < b->ng; ++i) byn[dir == 0 ? b->g[i]->nx : dir == 1 ? b->g[i]->ny : b->g[i]->nz].push_back(i);
    for (auto& kv : byn) {
      auto plan_of = [&](int i) -> LinePlan& { return b->prec == 8 ? b->g[i]->fplan[dir] : b->g[i]->fplan32[dir]; };
      const LinePlan keep = plan_of(kv.second[0]);
      LinePlan best = keep;
      const std::vector<LinePlan> cands =
          dir == 0 ? candidates_x(kv.first, b->prec, lds_max) : candidates(kv.first, b->prec, lds_max);
      for (const LinePlan& cand : cands) {
        if (cand.M == keep.M && cand.G == keep.G && cand.Wc == keep.Wc && cand.Wl == keep.Wl) continue;
        for (int i : kv.second) plan_of(i) = cand;
        b->built = false;
        double ms = 0.0;
        rc = time_step(&ms);
        if (rc) break;
        if (ms < best_ms) {
          best_ms = ms;
          best = cand;
        }
      }
      for (int i : kv.second) plan_of(i) = best;
      b->built = false;
      if (rc) return rc;
      std::snprintf(line, sizeof line, "%c n=%d: M=%d G=%d Wc=%d Wl=%d -> %.4f ms/step\n", "xyz"[dir],
                    kv.first, best.M, best.G, best.Wc, best.Wl, best_ms);
      log += line;
    }
  }
  if (report && report_len) {
    std::strncpy(report, log.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return X3D_OK;
}

}  // extern "C"
