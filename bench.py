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
  cpu_baseline: the CPU oracle (single-threaded C restatement of the
                reference) timed on a bounded sample of the same workload.
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


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="grid points per direction")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-n", type=int, default=256, help="grid of the CPU-baseline sample step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch a plain device copy of one field (counter calibration) "
                         "and whole steps, for rocprofv3 --pmc runs")
    return ap.parse_args()


def profile_kernels(n: int, reps: int = 3):
    """Launches for hardware-counter runs: a device-to-device copy of one n^3 FP64
    field (known bytes: 8 n^3 read + 8 n^3 written, the counter calibration)
    and whole steps, whose line passes are then read per kernel."""
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
    torch.cuda.synchronize()
    ctx.close()
    print(f"profile-kernels: {n}^3, {reps} copies of {8 * n ** 3} B and {reps} steps", flush=True)


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


def cpu_baseline(n: int):
    import numpy as np

    from oracle import Oracle
    o = Oracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi)
    u = o.fields(1, 1234, 1e-3)
    t0 = time.perf_counter()
    o.step(*u, gdt=0.0)
    dt = time.perf_counter() - t0
    return {
        "value": round(n ** 3 / dt / 1e6, 4),
        "unit": "Mpts/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one full Euler step of the C oracle at {n}^3 (seeded TG + 1e-3 noise, L=2pi), "
                  f"single thread, {dt:.1f} s",
    }


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
    ux, uy, uz, dux, duy, duz, pp3, px, py, pz = f
    ctx.init_fields(ux, uy, uz, kind=1, seed=1234, amp=1e-3)
    stream = ctx.stream

    def step():
        ctx.step(ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt=0.0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    gpu_ms_step = ev0.elapsed_time(ev1) / args.steps
    value = n ** 3 / (ms_step / 1e3) / 1e6

    # per-pass times of the same step, HIP events on the library's stream; the
    # dominant kernel is the slowest line pass (the Poisson pass is several
    # hipFFT kernels and is reported beside it)
    acc = {}
    for _ in range(args.kernel_reps):
        for name, (ms, bpp) in ctx.step_timed(ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt=0.0).items():
            acc.setdefault(name, [0.0, bpp])[0] += ms
    passes = {}
    for name, (ms, bpp) in acc.items():
        ms /= args.kernel_reps
        passes[name] = {"ms": round(ms, 4), "bytes_per_point": bpp,
                        "GBps": round(bpp * n ** 3 / (ms / 1e3) / 1e9, 1)}
    dom = max((k for k in passes if k != "poisson"), key=lambda k: passes[k]["ms"])
    kms = passes[dom]["ms"]
    alg_bytes = int(passes[dom]["bytes_per_point"] * n ** 3)
    achieved = alg_bytes / (kms / 1e3) / 1e9
    step_bytes = sum(p["bytes_per_point"] for p in passes.values()) * n ** 3
    roofline = {
        "bound": "hbm",
        "kernel": f"{dom} ({KERNEL_OF_PASS[dom]})",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile(dom, n),
        "avg_launch_ms": round(kms, 4),
        "algorithmic_bytes_per_launch": alg_bytes,
        "step_algorithmic_bytes": int(step_bytes),
        "step_frac": round(step_bytes / (gpu_ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "passes": passes,
    }
    ctx.close()
    del f
    torch.cuda.empty_cache()

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_n)
    out = {
        "metric": "Mpts/s per Euler timestep (plain FP64)",
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
