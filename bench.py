#!/usr/bin/env python3
"""Headline benchmark: batched FP64 flux-only DISORT on one MI355X.

Workload: 10,000 wavenumber bands of 1 cm^-1 over
100-10,100 cm^-1 x one column, nlyr=60, nstr=16, thermal emission, seeded
synthetic optical properties (disort_amd.workloads.c2).  A step is one solve
of all 10,000 instances with inputs already resident in HBM
(disort_flux_execute on device arrays); value = instances per second.

Prints one JSON line.  The CPU baseline is the oracle (the project's FP64
restatement of DisORT++'s solve_flux_spectral_bands, same OpenMP static band
loop) timed on this host's cores on the same inputs.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DISORT instances/sec (nlyr=60,nstr=16 fp64); surrogate vs plain at ≤1e-3 err"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(batch, reps: int = 3) -> dict:
    """Oracle (CPU FP64 restatement) on all host threads: warm-up, then the
    median of `reps` timed passes over the instances of `batch`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    threads = cpu_threads()
    warm = batch.__class__  # keep a reference to the type only
    del warm
    oracle_lib.flux_batch(batch, nthreads=threads)  # warm-up
    times = sorted(oracle_lib.flux_batch(batch, nthreads=threads) for _ in range(reps))
    sec = times[len(times) // 2]
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": batch.n / sec,
        "unit": "instances/s",
        "cores": threads,
        "kind": "port",
        "sample": f"all {batch.n} C2 bands, median of {reps} passes after one warm-up, "
        f"OpenMP static band loop, {threads} threads on {model}",
    }


def bench_c4(args) -> dict:
    """Ragged batch (C4): 50,000 instances, nstr ~ U{4,8,16,32,64}, nlyr ~
    U{10..120}, beam only, Lambertian.  Times the fixed schedule (default
    kernel per stream count, buckets one after another, eager launches)
    against the autotuned one (kernel variant per bucket and serial /
    concurrent schedule chosen by timing, replayed as one hipGraph)."""
    from disort_amd import flops, workloads
    from disort_amd._capi import Context, DeviceFluxBatch

    n = args.instances
    batch = workloads.c4(n)
    ctx = Context(0)
    dev = DeviceFluxBatch(ctx, batch)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) / args.steps

    fixed = timed(dev.execute)
    fixed_out = dev.download().out.copy()
    report = dev.autotune(reps=2)
    tuned = timed(dev.execute_graph)
    out = dev.download()
    assert (out.status == 0).all(), "non-zero instance status"
    import numpy as np

    same = bool(np.array_equal(out.out, fixed_out))
    total_flops = sum(flops.instance_flops(int(N), int(L), thermal=False, beam=True)
                      for N, L in zip(batch.nstr, batch.nlyr))
    line = {
        "metric": "C4 ragged batch instances/sec (nstr in {4,8,16,32,64}, nlyr 10..120, fp64), autotuned vs fixed schedule",
        "value": n / tuned,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tuned,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox draws, C2 optical-property distributions)",
        "config": {"workload": f"C4 ragged batch: {n} instances", "instances": n,
                   "layers": int(batch.nlyr.sum()), "beam": True, "thermal": False},
        "fixed_schedule": {"value": n / fixed, "ms_per_step": 1e3 * fixed,
                           "schedule": "default kernels, serial buckets, eager launches"},
        "autotuned_schedule": {"value": n / tuned, "ms_per_step": 1e3 * tuned, "tuning": report,
                               "launch": "one hipGraph per step"},
        "speedup_vs_fixed": fixed / tuned,
        "tuned_equals_fixed_bitwise": same,
        "whole_path_fp64_tflops": total_flops / tuned / 1e12,
        "whole_path_fp64_frac": total_flops / tuned / 1e12 / PEAK_FP64_TFLOPS,
    }
    if not args.no_cpu_baseline:
        sample = workloads.c4(args.cpu_sample, first=10 ** 6)
        cb = cpu_baseline(sample, reps=1)
        cb["sample"] = (f"{args.cpu_sample} C4-distributed instances (a disjoint seeded sample), "
                        + cb["sample"].split(", ", 1)[1])
        line["cpu_baseline"] = cb
    dev.close()
    return line


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--bands", type=int, default=10000)
    ap.add_argument("--workload", choices=("c2", "c4"), default="c2",
                    help="c2: the headline spectral flux solve (default); c4: ragged autotuned batch")
    ap.add_argument("--instances", type=int, default=50000, help="C4 batch size")
    ap.add_argument("--cpu-sample", type=int, default=2000, help="C4 instances timed on the CPU")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this build's instances are independent and shard "
            "across GPUs without communication (see DESIGN.md); run with --gpus 1.",
            file=sys.stderr,
        )
        return 2
    if args.workload == "c4":
        print(json.dumps(bench_c4(args)))
        return 0

    from disort_amd import flops, workloads
    from disort_amd._capi import Context, DeviceFluxBatch

    nstr, nlyr = 16, 60
    batch = workloads.c2(args.bands, nlyr=nlyr, nstr=nstr)
    ctx = Context(0)
    dev = DeviceFluxBatch(ctx, batch)

    for _ in range(args.warmup):
        dev.execute()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.execute()
    ctx.sync()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = args.bands * args.steps / sec

    # Per-kernel device time with hipEvents on the context stream (separate passes).
    ctx.set_timing(True)
    reps = 5
    stage = [0.0, 0.0, 0.0, 0.0]
    for _ in range(reps):
        dev.execute()
        for k, ms in enumerate(ctx.stage_times()):
            stage[k] += ms / reps
    ctx.set_timing(False)
    out = dev.download()
    assert (out.status == 0).all(), "non-zero instance status"

    names = ["k_column", "k_layer", "k_boundary", "k_flux"]
    kf = flops.kernel_flops(nstr, nlyr, thermal=True, beam=True)
    dom = max(range(4), key=lambda k: stage[k])
    dom_name = names[dom]
    achieved = kf[dom_name] * args.bands / (stage[dom] * 1e-3) / 1e12  # TFLOP/s
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom_name)
    total_flops = flops.instance_flops(nstr, nlyr) * args.bands
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox draws of the C2 distributions)",
        "config": {
            "workload": "C2 batched spectral flux solve: 10000 bands x 1 column" if args.bands == 10000
            else f"C2 subset: {args.bands} bands x 1 column",
            "n_bands": args.bands,
            "nlyr": nlyr,
            "nstr": nstr,
            "thermal": True,
            "path": "plain GPU DISORT (disort_flux_execute, inputs resident in HBM)",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
        },
        "stage_ms": dict(zip(names, stage)),
        "whole_path_fp64_tflops": total_flops / (sec / args.steps) / 1e12,
        "whole_path_fp64_frac": total_flops / (sec / args.steps) / 1e12 / PEAK_FP64_TFLOPS,
        "whole_path_hbm_algorithmic_gbs": flops.instance_bytes(nstr, nlyr, nstr + 1) * args.bands
        / (sec / args.steps) / 1e9,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(batch)
    dev.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
