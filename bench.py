#!/usr/bin/env python3
"""Benchmark: x3div Euler-timestep throughput (Mpts/s) on one MI355X.

A step is one iteration of the reference's step loop (src/xcompact3d.f90:80-95):
momentum RHS, Euler update, divergence, FFT Poisson, pressure gradient and
velocity correction, on a triply periodic grid (default 512^3, FP64) with
seeded Taylor-Green + 1e-3 noise fields resident in HBM.

Prints ONE JSON line on stdout.  Besides the throughput it carries
  roofline:     the dominant kernel's algorithmic bytes per launch divided by
                its average launch time (HIP events on the library's stream),
                against the 8 TB/s HBM peak;
  cpu_baseline: the CPU oracle (C restatement of the reference, built with
                -O3 and OpenMP over lines) timed on the host's threads and
                on one thread, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_FP16_PEAK_TFLOPS = 2500.0  # MI355X dense fp16 matrix peak


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="grid points per direction")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-n", type=int, default=256, help="grid of the all-threads CPU-baseline steps")
    ap.add_argument("--cpu-n1", type=int, default=128, help="grid of the one-thread CPU-baseline steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="fp32,bfp32,batch,large,nonperiodic",
                    help="comma list of extra forms to measure beside plain FP64: fp32 (precision "
                         "form at n^3), bfp32 (block-floating-point fields, FP64 arithmetic, at "
                         "n^3), batch (96 uneven grids, FP64 and FP32), large (FP32 at "
                         "1024x768x768 beside FP64 at 768^3), nonperiodic (der?_11/22 at 513^3); "
                         "empty for none")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch a plain device copy of one field (counter calibration) "
                         "and whole steps, for rocprofv3 --pmc runs")
    return ap.parse_args()


def profile_kernels(n: int, reps: int = 3):
    """Launches for hardware-counter runs: a device-to-device copy of one n^3 FP64
    field (known bytes: 8 n^3 read + 8 n^3 written, the counter calibration)
    and whole steps of the three forms, whose line passes are then read per kernel."""
    import numpy as np
    import torch

    from x3div_amd import X3D
    ctx = X3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    f = [ctx.field() for _ in range(10)]
    ctx.init_fields(f[0], f[1], f[2], kind=1, seed=1234, amp=1e-3)
    t = ctx.field()
    with torch.cuda.stream(ctx.stream):
        for _ in range(reps):
            t.copy_(f[0])
    for _ in range(reps):
        ctx.step(*f, gdt=0.0)
    f32 = [t.float() for t in f[:3]] + [ctx.field32() for _ in range(7)]
    for _ in range(reps):
        ctx.step_f32(*f32, gdt=0.0)
    del f32
    fb = [ctx.bfp_empty() for _ in range(10)]
    for i in range(3):
        ctx.bfp_encode(fb[i], f[i])
    for _ in range(reps):
        ctx.step_bfp(*fb, gdt=0.0)
    torch.cuda.synchronize()
    ctx.close()
    print(f"profile-kernels: {n}^3, {reps} copies of {8 * n ** 3} B, {reps} FP64, {reps} FP32 and {reps} BFP32 steps",
          flush=True)


# the kernel template each line pass of the step launches
KERNEL_OF_PASS = {
    "transeq_x": "k_transeq<double, M, 0>", "transeq_y": "k_transeq<double, M, 1>",
    "transeq_z": "k_transeq<double, M, 2>", "divergence_x": "k_div<double, M, 0>",
    "divergence_y": "k_div<double, M, 1>", "divergence_z": "k_div<double, M, 2>",
    "gradp_z": "k_grad<double, M, 2>", "gradp_y": "k_grad<double, M, 1>", "gradp_x": "k_grad<double, M, 0>",
}


def traffic_from_profile(pass_name: str, n: int):
    """HBM bytes per launch of one pass's kernel from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if one exists for this grid."""
    path = os.path.join(ROOT, "profiles", f"step_passes_{n}cubed_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^hbm_bytes_per_launch\s+{pass_name}:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_steps(n: int, threads: int, steps: int):
    """Per-step seconds of the oracle's speed build (-O3, OpenMP over lines) on
    `threads` host threads: steps + 1 Euler steps on n^3 seeded TG + 1e-3
    fields (L = 2pi), the first dropped (page faults, first touch)."""
    import numpy as np

    import oracle
    oracle.set_threads(threads, fast=True)
    o = oracle.Oracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, fast=True)
    u = o.fields(1, 1234, 1e-3)
    du = [o.zeros() for _ in range(3)]
    pp = o.zeros()
    g = [o.zeros() for _ in range(3)]
    times = []
    for _ in range(steps + 1):
        t0 = time.perf_counter()
        o.step_inplace(u, du, pp, g, gdt=0.0)
        times.append(time.perf_counter() - t0)
    return times[1:]


def cpu_baseline(n: int, n1: int, steps: int = 5):
    """The C oracle's speed build timed on this host: on every thread the
    process may use (OMP_NUM_THREADS) at n^3, and on one thread at n1^3;
    median of `steps` steps after a dropped first one.  The reference's own
    per-step timing is tend - tstart around one step (src/xcompact3d.f90:78,98)."""
    import statistics

    import oracle
    cores = oracle.max_threads(fast=True)
    t_all = cpu_steps(n, cores, steps)
    t_one = cpu_steps(n1, 1, steps)
    med_all = statistics.median(t_all)
    med_one = statistics.median(t_one)
    return {
        "value": round(n ** 3 / med_all / 1e6, 3),
        "unit": "Mpts/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"median of {steps} Euler steps (after one dropped) of the C oracle's -O3 OpenMP build "
                  f"at {n}^3, seeded TG + 1e-3 noise, L=2pi, {cores} threads: {med_all:.2f} s/step",
        "single_thread": {"value": round(n1 ** 3 / med_one / 1e6, 3), "grid": n1,
                          "s_per_step": round(med_one, 3)},
    }


def log(msg: str) -> None:
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def rel_errs(a, b):
    """max and L2 norms of a - b relative to b, in FP64."""
    a = a.double()
    d = (a - b).abs()
    return {"max": float(d.max() / b.abs().max()), "l2": float((a - b).norm() / b.norm())}


def time_form(ctx, f, step_fn, timed_fn, args):
    """W untimed steps, K steps timed between device synchronisations (wall
    clock and HIP events on the library's stream), then per-pass times of
    kernel_reps more steps from events recorded between the passes."""
    import torch
    for _ in range(args.warmup):
        step_fn(*f, gdt=0.0)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(ctx.stream)
    for _ in range(args.steps):
        step_fn(*f, gdt=0.0)
    ev1.record(ctx.stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    gpu_ms_step = ev0.elapsed_time(ev1) / args.steps
    acc = {}
    for _ in range(args.kernel_reps):
        for name, (ms, bpp) in timed_fn(*f, gdt=0.0).items():
            acc.setdefault(name, [0.0, bpp])[0] += ms
    passes = {}
    n3 = ctx.nx * ctx.ny * ctx.nz
    for name, (ms, bpp) in acc.items():
        ms /= args.kernel_reps
        passes[name] = {"ms": round(ms, 4), "bytes_per_point": bpp,
                        "GBps": round(bpp * n3 / (ms / 1e3) / 1e9, 1)}
    return ms_step, gpu_ms_step, passes


def measure_operators(ctx, src, dst, reps=10):
    """Each of the 18 FP64 drop-in operators (x3d_apply: one RHS + periodic
    line solve pass, 8 B read + 8 B written per point) timed alone: average
    of `reps` launches between HIP events on the library's stream, against
    the 8 TB/s HBM peak."""
    import torch
    from x3div_amd import OPS  # the reference's 18 operator instances, in order
    n3 = ctx.nx * ctx.ny * ctx.nz
    out = {}
    for name in OPS:
        for _ in range(2):
            ctx.apply(name, dst, src)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(ctx.stream)
        for _ in range(reps):
            ctx.apply(name, dst, src)
        ev1.record(ctx.stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        gbs = 16 * n3 / (ms / 1e3) / 1e9
        out[name] = {"ms": round(ms, 4), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def roofline_of(passes, n, gpu_ms_step, suffix, npts=None, stype=None):
    """The slowest line pass (the Poisson pass is several hipFFT kernels and is
    reported beside it) against the HBM peak, plus the whole step's fraction."""
    npts = npts or n ** 3
    dom = max((k for k in passes if k != "poisson"), key=lambda k: passes[k]["ms"])
    kms = passes[dom]["ms"]
    alg_bytes = int(passes[dom]["bytes_per_point"] * npts)
    achieved = alg_bytes / (kms / 1e3) / 1e9
    step_bytes = sum(p["bytes_per_point"] for p in passes.values()) * npts
    kern = KERNEL_OF_PASS[dom].replace("double", stype or ("float" if suffix else "double"))
    return {
        "bound": "hbm",
        "kernel": f"{dom} ({kern})",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile(dom + suffix, n),
        "avg_launch_ms": round(kms, 4),
        "algorithmic_bytes_per_launch": alg_bytes,
        "step_algorithmic_bytes": int(step_bytes),
        "step_frac": round(step_bytes / (gpu_ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "passes": passes,
    }


def measure_fp32(ctx, u0, args, ms_step):
    """The precision form on the same grid, from the same fields rounded to FP32."""
    import torch
    n = ctx.nx
    f32 = [t.float() for t in u0] + [ctx.field32() for _ in range(7)]
    ms32, gpu_ms32, passes32 = time_form(ctx, f32, ctx.step_f32, ctx.step_timed_f32, args)
    # accuracy of one step and of the 18 operators against the plain form
    # (itself within 1e-12 of the CPU oracle)
    g64 = [t.clone() for t in u0] + [ctx.field() for _ in range(7)]
    g32 = [t.float() for t in u0] + [ctx.field32() for _ in range(7)]
    ctx.step(*g64, gdt=0.0)
    ctx.step_f32(*g32, gdt=0.0)
    torch.cuda.synchronize()
    step_err = {name: rel_errs(g32[i], g64[i]) for i, name in ((0, "ux"), (3, "dux"), (6, "pp3"))}
    del g64, g32
    # per operator: the FP32 arithmetic alone (FP64 applied to the same
    # FP32-rounded field), and with the rounding of the field to FP32
    op_err, op_err_in = {}, {}
    o64, r64, o32, i32 = ctx.field(), ctx.field(), ctx.field32(), u0[0].float()
    i64 = i32.double()
    from x3div_amd import OPS
    for name in OPS:
        ctx.apply(name, o64, i64)
        ctx.apply(name, r64, u0[0])
        ctx.apply_f32(name, o32, i32)
        torch.cuda.synchronize()
        op_err[name] = rel_errs(o32, o64)["max"]
        op_err_in[name] = rel_errs(o32, r64)["max"]
    # the dense-circulant matrix-core variant of the three derivatives against
    # the FP32 line solves: time per application, MFMA rate, error vs FP64
    mf = {}
    if n % 128 == 0 and n <= 1024:
        for name, fn in (("derx", ctx.derx_f32), ("dery", ctx.dery_f32), ("derz", ctx.derz_f32)):
            ctx.apply_circ_mfma(name, o32, i32)
            fn(o32, i32)
            torch.cuda.synchronize()
            t = {}
            for tag, call in (("mfma", lambda: ctx.apply_circ_mfma(name, o32, i32)), ("thomas", lambda: fn(o32, i32))):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(ctx.stream)
                for _ in range(args.kernel_reps):
                    call()
                e1.record(ctx.stream)
                torch.cuda.synchronize()
                t[tag] = e0.elapsed_time(e1) / args.kernel_reps
            ctx.apply_circ_mfma(name, o32, i32)
            ctx.apply(name, o64, i64)
            torch.cuda.synchronize()
            flops = 3 * 2 * n * n ** 3
            mf[name] = {"mfma_ms": round(t["mfma"], 4), "thomas_f32_ms": round(t["thomas"], 4),
                        "mfma_TFLOPs": round(flops / (t["mfma"] / 1e3) / 1e12, 1),
                        "mfma_frac_of_fp16_peak": round(flops / (t["mfma"] / 1e3) / 1e12 / MFMA_FP16_PEAK_TFLOPS, 4),
                        "mfma_hbm_GBps": round(8 * n ** 3 / (t["mfma"] / 1e3) / 1e9, 1),
                        "max_rel_err_vs_fp64": rel_errs(o32, o64)["max"]}
    del o64, r64, o32, i32, i64, f32
    return {"value": round(n ** 3 / (ms32 / 1e3) / 1e6, 2), "ms_per_step": round(ms32, 4),
            "circulant_mfma": mf,
            "dtype": "f32 line solves and FFTs, f64 spectral division",
            "speedup_vs_fp64": round(ms_step / ms32, 3),
            "roofline": roofline_of(passes32, n, gpu_ms32, "_f32"),
            "step_rel_err_vs_fp64": step_err,
            "operator_max_rel_err_vs_fp64_same_input": op_err,
            "operator_max_rel_err_vs_fp64_incl_input_rounding": op_err_in}


def measure_bfp(ctx, u0, args, ms_step):
    """The BFP32 form on the same grid: the fields' BFP32 words (4 B per
    point, exponent per 16-point x-segment), FP64 stencils and solves.  Its
    errors are taken against the plain form on the unrounded FP64 fields, so
    they include the rounding of the input."""
    import torch
    from x3div_amd import OPS, OPS2
    n = ctx.nx
    q0 = []
    for t in u0:
        q = ctx.bfp_empty()
        ctx.bfp_encode(q, t)
        q0.append(q)
    fb = [q.clone() for q in q0] + [ctx.bfp_empty() for _ in range(7)]
    msb, gpu_msb, passesb = time_form(ctx, fb, ctx.step_bfp, ctx.step_timed_bfp, args)
    del fb
    # one step against the plain form's step on the unrounded fields
    g64 = [t.clone() for t in u0] + [ctx.field() for _ in range(7)]
    gb = [q.clone() for q in q0] + [ctx.bfp_empty() for _ in range(7)]
    ctx.step(*g64, gdt=0.0)
    ctx.step_bfp(*gb, gdt=0.0)
    dec = ctx.field()
    step_err = {}
    for i, name in ((0, "ux"), (3, "dux"), (6, "pp3"), (7, "px")):
        ctx.bfp_decode(dec, gb[i])
        torch.cuda.synchronize()
        step_err[name] = rel_errs(dec, g64[i])
    del g64, gb
    # every operator: error against FP64 on the unrounded field, and time
    # (4 B read + 4 B written per point)
    ref, outq = ctx.field(), ctx.bfp_empty()
    op_err, op_ms = {}, {}
    for name in OPS + OPS2:
        ctx.apply(name, ref, u0[0])
        ctx.apply_bfp(name, outq, q0[0])
        ctx.bfp_decode(dec, outq)
        torch.cuda.synchronize()
        op_err[name] = rel_errs(dec, ref)["max"]
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        for _ in range(args.kernel_reps):
            ctx.apply_bfp(name, outq, q0[0])
        e1.record(ctx.stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.kernel_reps
        gbs = 8 * n ** 3 / (ms / 1e3) / 1e9
        op_ms[name] = {"ms": round(ms, 4), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    del ref, outq, dec, q0
    worst = max(op_err.values())
    return {"value": round(n ** 3 / (msb / 1e3) / 1e6, 2), "ms_per_step": round(msb, 4),
            "dtype": "bfp32 fields (31-bit mantissa, exponent per 16-point x-segment), f64 stencils, "
                     "solves and Poisson",
            "speedup_vs_fp64": round(ms_step / msb, 3),
            "roofline": roofline_of(passesb, n, gpu_msb, "_bfp", stype="bfp32"),
            "step_rel_err_vs_fp64_incl_input_rounding": step_err,
            "operator_max_rel_err_vs_fp64_incl_input_rounding": op_err,
            "operator_worst_rel_err": worst,
            "operators_within_1e-6": bool(worst <= 1e-6),
            "operators": op_ms}


BATCH_SIZES = (32, 48, 64, 96, 128, 160)


def batch_dims(seed: int = 2024, ngrids: int = 96):
    """96 grids, each dimension drawn uniformly from BATCH_SIZES (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(BATCH_SIZES, 3)) for _ in range(ngrids)]


def measure_batch(args):
    """The batching form: 96 uneven grids with 8-mode Fourier fields, tilings
    from the autotuner, the batched step captured in a HIP graph and replayed."""
    import torch
    from x3div_amd import X3DBatch
    dims = batch_dims()
    out = {}
    for prec, tag in ((8, "batch96_fp64"), (4, "batch96_fp32")):
        b = X3DBatch(dims, precision=prec)
        b.init_fields(kind=2, seed=77)
        report = b.autotune(reps=2)
        b.graph_capture(gdt=0.0)
        for _ in range(args.warmup):
            b.graph_launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.graph_launch()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        pts = b.total_points
        out[tag] = {"value": round(pts / (ms / 1e3) / 1e6, 2), "unit": "Mpts/s", "ms_per_step": round(ms, 4),
                    "grids": len(dims), "points": pts,
                    "dims_seed": 2024, "fields": "8 seeded Fourier modes per component, amplitude 1/8 each",
                    "autotune": report.strip().splitlines()[-1] if report.strip() else ""}
        b.close()
        torch.cuda.empty_cache()
    return out


def measure_nonperiodic(args, n: int = 513):
    """The non-periodic first derivatives (der?_11 and der?_22, both
    instances) on the reference's n^3 = 513^3 non-periodic grid (lines of
    2^k + 1 points, padded to 64 chunks of 9): time per application between
    HIP events, 8 B read + 8 B written per point, against the HBM peak."""
    import numpy as np
    import torch
    from x3div_amd import X3D
    out = {}
    for bc in ((1, 1), (2, 2)):
        ctx = X3D(n, n, n, 2.0, 2.0, 2.0, ncl1=(bc[0],) * 3, ncln=(bc[1],) * 3)
        u = [ctx.field() for _ in range(3)]
        ctx.init_fields(*u, kind=2, seed=7, amp=0.0)
        t = ctx.field()
        for axis in "xyz":
            for name in (f"der{axis}", f"der{axis}p"):
                fn = getattr(ctx, f"der{axis}")
                fn(t, u[0], name)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(ctx.stream)
                for _ in range(args.kernel_reps):
                    fn(t, u[0], name)
                e1.record(ctx.stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / args.kernel_reps
                gbs = 16 * n ** 3 / (ms / 1e3) / 1e9
                out[f"{name}_{bc[0]}{bc[1]}"] = {"ms": round(ms, 4), "GBps": round(gbs, 1),
                                                  "frac": round(gbs / HBM_PEAK_GBS, 4)}
        ctx.close()
        del u, t
        torch.cuda.empty_cache()
    return {"grid": [n, n, n], "operators": out}


def measure_large(args):
    """FP32 precision form at 1024x768x768 beside plain FP64 at 768^3."""
    import numpy as np
    import torch
    from x3div_amd import X3D
    out = {}
    L = 2 * np.pi
    for (nx, ny, nz), prec, tag in (((1024, 768, 768), 4, "large_fp32_1024x768x768"),
                                    ((768, 768, 768), 8, "fp64_768cubed")):
        ctx = X3D(nx, ny, nz, L, L, L)
        u = [ctx.field() for _ in range(3)]
        ctx.init_fields(*u, kind=1, seed=1234, amp=1e-3)
        if prec == 4:
            f = [t.float() for t in u] + [ctx.field32() for _ in range(7)]
            del u
            torch.cuda.empty_cache()
            ms, gpu_ms, passes = time_form(ctx, f, ctx.step_f32, ctx.step_timed_f32, args)
            suffix = "_f32"
        else:
            f = u + [ctx.field() for _ in range(7)]
            ms, gpu_ms, passes = time_form(ctx, f, ctx.step, ctx.step_timed, args)
            suffix = ""
        pts = nx * ny * nz
        rl = roofline_of(passes, 0, gpu_ms, suffix, npts=pts)
        out[tag] = {"value": round(pts / (ms / 1e3) / 1e6, 2), "unit": "Mpts/s", "ms_per_step": round(ms, 4),
                    "grid": [nx, ny, nz], "dtype": "f32 line solves and FFTs, f64 spectral division"
                    if prec == 4 else "f64", "roofline_frac": rl["frac"], "step_frac": rl["step_frac"],
                    "dominant": rl["kernel"]}
        ctx.close()
        del f
        torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the "
              "multi-GPU pencil decomposition is design-only.", file=sys.stderr)
        sys.exit(2)
    if args.profile_kernels:
        profile_kernels(args.n)
        return

    import numpy as np
    import torch

    from x3div_amd import X3D

    n = args.n
    L = 2 * np.pi
    ctx = X3D(n, n, n, L, L, L)
    f = [ctx.field() for _ in range(10)]
    ctx.init_fields(f[0], f[1], f[2], kind=1, seed=1234, amp=1e-3)
    u0 = [t.clone() for t in f[:3]]

    # plain FP64 form: the drop-in, and the headline value
    ms_step, gpu_ms_step, passes = time_form(ctx, f, ctx.step, ctx.step_timed, args)
    value = n ** 3 / (ms_step / 1e3) / 1e6
    roofline = roofline_of(passes, n, gpu_ms_step, "")

    extra = [x for x in args.forms.split(",") if x]
    forms = {"fp64": {"value": round(value, 2), "ms_per_step": round(ms_step, 4), "dtype": "f64",
                      "step_frac": roofline["step_frac"]}}
    log(f"fp64 {n}^3: {ms_step:.3f} ms/step")
    operators = measure_operators(ctx, f[0], f[3])
    log("fp64 single operators done")
    if "fp32" in extra:
        forms["fp32"] = measure_fp32(ctx, u0, args, ms_step)
        log(f"fp32 {n}^3: {forms['fp32']['ms_per_step']:.3f} ms/step")
    if "bfp32" in extra and n % 16 == 0:
        forms["bfp32"] = measure_bfp(ctx, u0, args, ms_step)
        log(f"bfp32 {n}^3: {forms['bfp32']['ms_per_step']:.3f} ms/step, worst operator error "
            f"{forms['bfp32']['operator_worst_rel_err']:.2e}")
    ctx.close()
    del f, u0
    torch.cuda.empty_cache()
    if "batch" in extra:
        forms.update(measure_batch(args))
        log("batch forms done")
    if "large" in extra:
        forms.update(measure_large(args))
    if "nonperiodic" in extra:
        forms["nonperiodic_derivatives"] = measure_nonperiodic(args)
        log("non-periodic derivatives done")
        log("large grids done")

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_n, args.cpu_n1)
    out = {
        "metric": "Mpts/s per Euler timestep at 512\u00b3 (plain FP64 vs FP32/split-BF16) + HBM GB/s",
        "value": round(value, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "gpu_ms_per_step": round(gpu_ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Taylor-Green + 1e-3 uniform noise, generated on device",
        "config": {
            "workload": f"x3div Euler timestep, {n}^3 triply periodic, plain FP64 form",
            "grid": [n, n, n],
            "lengths": "2pi x 2pi x 2pi",
            "step": "transeq + Euler + divergence + FFT Poisson + gradp + cor_vel",
            "parallelism": "1 GPU (prow=pcol=1)",
        },
        "roofline": roofline,
        "operators_fp64": operators,
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
