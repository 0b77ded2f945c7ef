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


def cpu_model() -> str:
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return model


def general_from_flux(batch):
    """The same C2 instances as a comp_only_fluxes general batch (the input of
    solve_spectral_bands): one column per band."""
    import numpy as np

    from disort_amd import general as G

    n, L, P = batch.n, int(batch.nlyr[0]), int(batch.pstride[0])
    return G.GeneralBatch(flags=int(batch.flags) | G.G_ONLY_FLUXES, nstr=int(batch.nstr[0]),
                          dtau=batch.dtau.reshape(n, L), ssa=batch.ssa.reshape(n, L),
                          pmom=batch.pmom.reshape(n, L, P), scal=batch.scal.reshape(n, -1),
                          temper=None if batch.temper is None else batch.temper.reshape(n, L + 1),
                          phi_user=np.zeros(0))


def cpu_baseline(batch, reps: int = 5, general: bool = True) -> tuple[dict, object]:
    """CPU restatement of DisORT++'s two spectral drivers on all host threads,
    on the same inputs as the GPU: the flux driver (solve_flux_spectral_bands:
    DisortFluxSolver<16> per band) and the general driver (solve_spectral_bands
    with comp_only_fluxes: DisortSolver per band), each an OpenMP static band
    loop (src/python_bindings.cpp:93-104, :177-188).  One warm-up, then the
    median of `reps` passes, timed inside C++ (the loop) and around the call.
    Returns (baseline dict, flux-driver output [8][levels])."""
    import ctypes as C

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    threads = cpu_threads()
    flux_loop, flux_call = [], []
    for k in range(reps + 1):
        t0 = time.perf_counter()
        sec = oracle_lib.flux_batch(batch, nthreads=threads)
        if k:
            flux_loop.append(sec)
            flux_call.append(time.perf_counter() - t0)
    out = batch.out.copy()
    med = lambda v: float(np.median(v))  # noqa: E731
    model = cpu_model()
    base = {
        "value": batch.n / med(flux_loop),
        "unit": "instances/s",
        "cores": threads,
        "kind": "port",
        "sample": f"all {batch.n} instances; CPU restatement of DisORT++'s solve_flux_spectral_bands "
                  f"(flux driver), OpenMP static loop, {threads} threads on {model}; median of {reps} "
                  f"passes after one warm-up, C++ loop time",
        "solve_flux_spectral_bands": {"instances_per_s": batch.n / med(flux_loop),
                                      "instances_per_s_python_call": batch.n / med(flux_call),
                                      "loop_s": flux_loop},
        "nproc": os.cpu_count(),
        "cpu_model": model,
    }
    if not general:
        return base, out
    gb = general_from_flux(batch)
    lib = oracle_lib.lib()
    lib.oracle_general_batch.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    lib.oracle_general_batch.restype = C.c_int64
    st = gb.c_struct()
    gen_loop, gen_call = [], []
    for k in range(reps + 1):
        secs = C.c_double()
        t0 = time.perf_counter()
        failed = lib.oracle_general_batch(C.byref(st), 0, gb.n, threads, C.byref(secs))
        if failed:
            raise RuntimeError(f"{failed} general oracle columns failed")
        if k:
            gen_loop.append(secs.value)
            gen_call.append(time.perf_counter() - t0)
    # The two drivers solve the same problems: their fluxes agree.
    gflux = np.concatenate([gb.flux[:, i, :] for i in range(gb.n)], axis=1)
    drivers_diff = float(np.max(np.abs(gflux - out) / (np.abs(out) + 1e-3)))
    base["solve_spectral_bands"] = {"instances_per_s": gb.n / med(gen_loop),
                                    "instances_per_s_python_call": gb.n / med(gen_call),
                                    "loop_s": gen_loop, "comp_only_fluxes": True}
    base["drivers_max_rel_diff"] = drivers_diff
    return base, out


def output_errors(gpu, ref, batch) -> dict:
    """Per output field: max |gpu - cpu| / (|cpu| + 1e-3) over every band and
    level, i.e. <= 1e-9 exactly when |gpu - cpu| <= 1e-9 |cpu| + 1e-12, with
    the band and level where it occurs."""
    import numpy as np

    from disort_amd.batch import OUT_FIELDS

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    L1 = int(batch.nlyr[0]) + 1
    res = {}
    for f, name in enumerate(OUT_FIELDS):
        e = np.abs(gpu[f] - ref[f]) / (np.abs(ref[f]) + 1e-3)
        j = int(np.argmax(e))
        res[name] = {"max": float(e[j]), "band": j // L1, "level": j % L1,
                     "gpu": float(gpu[f, j]), "cpu": float(ref[f, j]),
                     "n_over_1e-9": int((e > 1e-9).sum())}
        if e[j] > 1e-9:
            # Which side is off: both against the same solve in x87 extended
            # precision (64-bit significand) of that band.
            x = float(oracle_lib.flux_one_extended(batch, j // L1)[f, j % L1])
            res[name]["extended_precision"] = x
            res[name]["gpu_vs_extended"] = abs(gpu[f, j] - x) / (abs(x) + 1e-3)
            res[name]["cpu_vs_extended"] = abs(ref[f, j] - x) / (abs(x) + 1e-3)
    return res


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
        cb, _ = cpu_baseline(sample, reps=1, general=False)
        cb["sample"] = (f"{args.cpu_sample} C4-distributed instances (a disjoint seeded sample), "
                        + cb["sample"].split("; ", 1)[1])
        line["cpu_baseline"] = cb
    dev.close()
    return line


def bench_c5(args) -> dict:
    """Learned surrogate (C5): train MLP members (609 features -> 512 x 4 ->
    183, BF16 MFMA; two by default) that predict the residual of the solver's fluxes over
    a two-stream baseline, on solver-generated columns; calibrate the gate,
    then serve fresh columns and compare with the plain FP64 GPU solver on
    the same columns.  Prints the gated throughput at the 1e-3 net-flux
    bound (relative to the incident beam) and a tolerance sweep, with the
    local-net-flux metric beside it."""
    import numpy as np

    from disort_amd import surrogate as S
    from disort_amd._capi import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    W = S.WIDTHS_TS
    x, t, n, norm = S.build_dataset(ctx, args.train_columns, seed=1, w0=W[0], t_floor=S.T_FLOOR)
    gen_s = time.perf_counter() - t0
    members, train_s, hist = [], [], []
    # Loss weights: diffuse net flux 1, diffuse sum 0.25 (the gate's metric is
    # the net flux; the direct beam is the closed form).
    wts = np.concatenate([np.ones(S.NLEV), np.full(S.NLEV, 0.25)])
    restarts = 0
    for k in range(args.members):
        m = S.MLP(ctx, W, max_batch=65536)  # training batches of 8192; serving in 65,536-column chunks
        lr = 1.2e-3
        t0 = time.perf_counter()
        while True:
            m.set_params(m.init_params(100 + k, S.N_FEAT_TS, S.N_OUT))
            h = S.train(ctx, m, x, t, n, epochs=args.epochs, batch=8192, lr=lr, seed=k, weights=wts,
                        log=lambda s_: print(s_, file=sys.stderr, flush=True))
            # A run whose loss ends far above where it started from has
            # diverged (healthy runs end near 4e-3): train again at half the rate.
            if h[-1] < 0.1 or restarts >= 2:
                break
            restarts += 1
            lr *= 0.5
            print(f"member {k} diverged (loss {h[-1]:.3e}); retraining at lr {lr:.1e}", file=sys.stderr, flush=True)
        ctx.sync()
        train_s.append(time.perf_counter() - t0)
        hist.append(h[-1])
        members.append(m)
    x.free()
    t.free()
    mlp_flops = 2.0 * (S.N_FEAT_TS * 512 + 3 * 512 * 512 + 512 * 183)  # per column, forward, unpadded
    train_tflops = 3.0 * mlp_flops * n * args.epochs / (sum(train_s) / len(train_s)) / 1e12

    cal = S.DeviceColumns(ctx, S.c5_columns(32768, seed=77))
    cal.solve()
    cal_ref = S.host_fluxes(cal.dev.download())
    gate = S.GatedSurrogate(ctx, members, norm, alpha=0.01, tol=S.TOL, metric="beam")
    cal_info = gate.calibrate(cal, cal_ref)
    sweep_gates = []
    for tol, metric in ((2e-3, "beam"), (3e-3, "beam"), (1e-2, "beam"), (1e-3, "local")):
        g = S.GatedSurrogate(ctx, members, norm, alpha=0.01, tol=tol, metric=metric)
        info = g.calibrate(cal, cal_ref)
        sweep_gates.append((tol, g, info))
    cal.close()

    # Serving chunks of 512k columns: the plain re-solve of a chunk's routed
    # columns has a latency floor of about 2 ms (one boundary-sweep lifetime),
    # which larger chunks amortise.
    chunk = 524288
    n_serve = args.serve_columns
    plain_s = gated_s = mlp_s = 0.0
    errs_local, errs_beam, acc_all = [], [], 0
    sweep_s = [0.0] * len(sweep_gates)
    sweep_acc = [0] * len(sweep_gates)
    sweep_viol = [0] * len(sweep_gates)
    n_plain = n_sampled = 0
    # Warm-up (untimed): one serve call per gate on a separate chunk, so that
    # the sub-batch plans a serving process keeps are built before timing.
    wdc = S.DeviceColumns(ctx, S.c5_columns(chunk, seed=4999))
    for g_ in [gate] + [g for _, g, _ in sweep_gates]:
        fw, iw = g_.serve(wdc, sample_frac=args.sample_frac, seed=4999)
        fw.free()
        iw["flags"].free()
    wdc.close()
    done = 0
    k = 0
    while done < n_serve:
        m_ = min(chunk, n_serve - done)
        hb = S.c5_columns(m_, seed=5000 + k)
        dc = S.DeviceColumns(ctx, hb)
        dc.solve()
        ctx.sync()
        _, dt = S.timed(ctx, dc.solve)
        plain_s += dt
        ref = S.host_fluxes(dc.dev.download())
        u = hb.scal.reshape(m_, -1)[:, 1] * hb.scal.reshape(m_, -1)[:, 0]
        (fl, sp, od), dt = S.timed(ctx, lambda: gate.predict(dc))
        mlp_s += dt
        for a in (fl, sp, od):
            a.free()
        (flux, info), dt = S.timed(ctx, lambda: gate.serve(dc, sample_frac=args.sample_frac, seed=k))
        gated_s += dt
        f = flux.get(np.float64, (m_, 3, S.NLEV))
        summ = S.flags_summary(info["flags"].get(np.int32, m_))
        flux.free()
        info["flags"].free()
        ea = S.net_flux_error(f, ref, u)
        eb = S.net_flux_error_beam(f, ref, u)
        acc = summ["accepted_idx"]
        errs_local.append(ea[acc])
        errs_beam.append(eb[acc])
        acc_all += acc.size
        n_plain += info["n_plain"]
        n_sampled += summ["n_sampled"]
        for j, (tol, g, _) in enumerate(sweep_gates):
            (flux2, inf2), dt = S.timed(ctx, lambda: g.serve(dc, sample_frac=args.sample_frac, seed=k))
            sweep_s[j] += dt
            f2 = flux2.get(np.float64, (m_, 3, S.NLEV))
            acc2 = S.flags_summary(inf2["flags"].get(np.int32, m_))["accepted_idx"]
            flux2.free()
            inf2["flags"].free()
            e2 = S.column_error(f2, ref, u, g.metric)
            sweep_acc[j] += acc2.size
            sweep_viol[j] += int((e2[acc2] > tol).sum())
        dc.close()
        done += m_
        k += 1
        print(f"served {done}/{n_serve}", file=sys.stderr, flush=True)

    def q(e):
        e = np.concatenate(e) if e else np.zeros(0)
        return {p: float(np.quantile(e, p / 100)) for p in (50, 90, 99)} if e.size else None

    plain_rate = n_serve / plain_s
    gated_rate = n_serve / gated_s
    mlp_rate = n_serve / mlp_s
    # The network's GEMMs alone (disort_mlp_infer launches only k_gemm): both
    # members on resident 65,536-row batches, back to back.
    rows = 65536
    xb = S.DevArray(ctx, rows * W[0] * 2)
    yb = S.DevArray(ctx, rows * W[-1] * 4)
    ctx.h2d(xb.ptr, np.zeros(rows * W[0], np.uint16))
    for m in members:
        m.infer(xb.ptr, rows, yb.ptr)
    ctx.sync()
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        for m in members:
            m.infer(xb.ptr, rows, yb.ptr)
    ctx.sync()
    gemm_s = (time.perf_counter() - t0) / (reps * len(members))
    pad_flops = 2.0 * rows * sum(a * b for a, b in zip(W[:-1], W[1:]))
    xb.free()
    yb.free()
    line = {
        "metric": "gated surrogate vs plain GPU DISORT columns/sec at <=1e-3 net-flux error, relative to the "
                  "incident beam (nlyr=60, nstr=16)",
        "value": gated_rate,
        "unit": "columns/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 (MLP, fp32 accumulate) + f64 (plain fallback)",
        "data": "synthetic (seeded Philox draws; C5 column distributions)",
        "config": {"workload": f"C5 surrogate: train {n} columns x {args.epochs} epochs x {args.members} members, "
                               f"serve {n_serve} fresh columns",
                   "mlp": f"{S.N_FEAT_TS} (242 column + 183 level + 62 thermal-source + 122 two-stream features) "
                          "-> 512x4 (GELU) -> 183, predicting the residual over the two-stream baseline",
                   "gate": f"OOD range + {args.members}-member spread + sampled re-solve (online error monitor), "
                           "routed on the device",
                   "error_metric": "max_levels |dF_net| / (mu0 F0)", "tol": S.TOL,
                   "sample_frac": args.sample_frac},
        "plain_gpu_columns_per_s": plain_rate,
        "gated_vs_plain": gated_rate / plain_rate,
        "accepted_frac": acc_all / n_serve,
        "plain_routed_frac": n_plain / n_serve,
        "sampled_frac": n_sampled / n_serve,
        "accepted_err_quantiles_local": q(errs_local),
        "accepted_err_quantiles_beam": q(errs_beam),
        "calibration": cal_info,
        "monitor": gate.monitor.stats(),
        "tolerance_sweep": [
            {"tol": tol, "metric": g.metric, "calibration_accept_frac": inf["accept_frac"],
             "served_accept_frac": sweep_acc[j] / n_serve,
             "accepted_violation_frac": sweep_viol[j] / max(sweep_acc[j], 1),
             "columns_per_s": n_serve / sweep_s[j], "vs_plain": (n_serve / sweep_s[j]) / plain_rate}
            for j, (tol, g, inf) in enumerate(sweep_gates)],
        "mlp_inference_columns_per_s": mlp_rate,
        "roofline": {"bound": "mfma", "kernel": "k_gemm (disort_mlp_infer, 65,536 rows, padded widths "
                     + "x".join(str(w) for w in W) + ")",
                     "achieved": pad_flops / gemm_s / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": pad_flops / gemm_s / 1e12 / 2500.0, "traffic": None},
        "predict_path_tflops": 2 * mlp_flops * mlp_rate / 1e12,
        "training": {"columns": n, "epochs": args.epochs, "seconds_per_member": train_s,
                     "final_loss": hist, "tflops": train_tflops, "dataset_seconds": gen_s,
                     "diverged_restarts": restarts},
    }
    gate.close()
    for _, g, _ in sweep_gates:
        g.close()
    for m in members:
        m.close()
    return line


def bench_jac(args) -> dict:
    """Learned temperature Jacobians (the surrogate extension): two members
    (671 features -> 512 x 4 -> 124, BF16 MFMA) learn the residual of the responses
    of the top-of-atmosphere upward flux and the surface downward flux to
    every level's Planck value from solver-generated Jacobians; the gate
    (OOD + spread + sampled re-solve, online monitor) routes failing columns
    to the plain GPU Jacobian (disort_jac_execute).  Prints the gated
    throughput at the 2e-3 bound (error relative to pi dB/dT at each
    temperature) against the plain GPU Jacobian on the same columns, a
    tolerance sweep, and the CPU oracle's Jacobian on a bounded sample."""
    import numpy as np

    from disort_amd import jac_surrogate as J
    from disort_amd import surrogate as S
    from disort_amd._capi import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    x, t, n, norm = J.build_jac_dataset(ctx, args.jac_train_columns, seed=1, chunk=16384)
    ctx.sync()
    gen_s = time.perf_counter() - t0
    print(f"dataset {n} columns in {gen_s:.1f} s", file=sys.stderr, flush=True)
    members, train_s, hist, restarts = [], [], [], 0
    for k in range(2):
        t0 = time.perf_counter()
        m, h = J.train_member(ctx, x, t, n, args.jac_epochs, seed=300 + k, batch=2048, lr=1e-3, max_batch=65536)
        ctx.sync()
        train_s.append(time.perf_counter() - t0)
        hist.append(h[-1])
        members.append(m)
        print(f"member {k}: {train_s[-1]:.1f} s, final loss {h[-1]:.3e}", file=sys.stderr, flush=True)
    # A member whose loss ends far above the other's has had a training
    # excursion: train it again at half the rate.
    worse = int(hist[1] > hist[0])
    if hist[worse] > 20.0 * hist[1 - worse]:
        restarts = 1
        members[worse].close()
        t0 = time.perf_counter()
        m, h = J.train_member(ctx, x, t, n, args.jac_epochs, seed=300 + worse, batch=2048, lr=5e-4, max_batch=65536)
        ctx.sync()
        train_s[worse] += time.perf_counter() - t0
        hist[worse] = h[-1]
        members[worse] = m
        print(f"member {worse} retrained at lr 5e-4: final loss {h[-1]:.3e}", file=sys.stderr, flush=True)
    x.free()
    t.free()
    W = J.JWIDTHS
    mlp_flops = 2.0 * (J.N_EFEAT_TS * 512 + 3 * 512 * 512 + 512 * J.N_JOUT)
    train_tflops = 3.0 * mlp_flops * n * args.jac_epochs / (sum(train_s) / 2) / 1e12

    cal = J.JacColumns(ctx, S.c5_columns(16384, seed=501))
    cal.plain()
    cal_ref = cal.plain_rows_host()
    gate = J.JacGatedSurrogate(ctx, members, norm, alpha=0.01, tol=J.JTOL)
    cal_info, cal_err = gate.calibrate(cal, cal_ref)
    sweep = []
    for tol in (1e-2, 5e-3, 2e-3, 1e-3):
        g = J.JacGatedSurrogate(ctx, members, norm, alpha=0.01, tol=tol)
        sweep.append((tol, g, g.calibrate(cal, cal_ref)[0]))
    cal.close()

    chunk = 65536
    wdc = J.JacColumns(ctx, S.c5_columns(chunk, seed=4999))
    for g_ in [gate] + [g for _, g, _ in sweep]:
        r_, i_ = g_.serve(wdc, sample_frac=args.jac_sample_frac, seed=4999)
        r_.free()
        i_["flags"].free()
    wdc.close()
    n_serve = args.jac_serve_columns
    plain_s = gated_s = mlp_s = 0.0
    errs, acc_all, n_plain, n_sampled = [], 0, 0, 0
    sweep_s = [0.0] * len(sweep)
    sweep_acc = [0] * len(sweep)
    sweep_viol = [0] * len(sweep)
    done = k = 0
    while done < n_serve:
        m_ = min(chunk, n_serve - done)
        hb = S.c5_columns(m_, seed=5000 + k)
        dc = J.JacColumns(ctx, hb)
        dc.plain()
        ctx.sync()
        _, dt = S.timed(ctx, dc.plain)
        plain_s += dt
        ref = dc.plain_rows_host()
        db = J.dbdt_dofs(hb)
        (r_, sp_, od_), dt = S.timed(ctx, lambda: gate.predict(dc))
        mlp_s += dt
        for a in (r_, sp_, od_):
            a.free()
        (rows, info), dt = S.timed(ctx, lambda: gate.serve(dc, sample_frac=args.jac_sample_frac, seed=k))
        gated_s += dt
        got = rows.get(np.float64, (m_, 2, J.D))
        fl = info["flags"].get(np.int32, m_)
        rows.free()
        info["flags"].free()
        e = J.jacobian_error(got, ref, db)
        acc = np.nonzero((fl & 1) == 0)[0]
        errs.append(e[acc])
        acc_all += acc.size
        n_plain += info["n_plain"]
        n_sampled += int(((fl & 2) != 0).sum())
        for j, (tol, g, _) in enumerate(sweep):
            (r2, i2), dt = S.timed(ctx, lambda: g.serve(dc, sample_frac=args.jac_sample_frac, seed=k))
            sweep_s[j] += dt
            got2 = r2.get(np.float64, (m_, 2, J.D))
            acc2 = np.nonzero((i2["flags"].get(np.int32, m_) & 1) == 0)[0]
            r2.free()
            i2["flags"].free()
            sweep_acc[j] += acc2.size
            sweep_viol[j] += int((J.jacobian_error(got2, ref, db)[acc2] > tol).sum())
        dc.close()
        done += m_
        k += 1
        print(f"served {done}/{n_serve}", file=sys.stderr, flush=True)
    e_all = np.concatenate(errs) if errs else np.zeros(0)
    plain_rate, gated_rate, mlp_rate = n_serve / plain_s, n_serve / gated_s, n_serve / mlp_s

    rows_ = 65536
    xb = S.DevArray(ctx, rows_ * W[0] * 2)
    yb = S.DevArray(ctx, rows_ * W[-1] * 4)
    ctx.h2d(xb.ptr, np.zeros(rows_ * W[0], np.uint16))
    for m in members:
        m.infer(xb.ptr, rows_, yb.ptr)
    ctx.sync()
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        for m in members:
            m.infer(xb.ptr, rows_, yb.ptr)
    ctx.sync()
    gemm_s = (time.perf_counter() - t0) / (reps * len(members))
    pad_flops = 2.0 * rows_ * sum(a * b for a, b in zip(W[:-1], W[1:]))
    xb.free()
    yb.free()

    line = {
        "metric": "gated Jacobian surrogate vs plain GPU temperature Jacobian, columns/sec at <=2e-3 error "
                  "(nlyr=60, nstr=16)",
        "value": gated_rate,
        "unit": "columns/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 (MLP, fp32 accumulate) + f64 (plain fallback)",
        "data": "synthetic (seeded Philox draws; C5 column distributions)",
        "config": {"workload": f"Jacobian surrogate: train {n} columns x {args.jac_epochs} epochs x 2 members, "
                               f"serve {n_serve} fresh columns",
                   "outputs": "dF/dT_k of flux_up at the top and flux_down at the surface, k = 61 levels + surface",
                   "mlp": f"{J.N_EFEAT} emission features + {J.N_JOUT} two-stream responses -> 512x4 (GELU) -> "
                          f"{J.N_JOUT}, predicting the residual over the two-stream responses",
                   "gate": "OOD range + 2-member spread + sampled re-solve (online error monitor), routed on the device",
                   "error_metric": "max_k |dF/dT_k - ref| / (pi dB/dT(T_k)) over both rows", "tol": J.JTOL,
                   "sample_frac": args.jac_sample_frac},
        "plain_gpu_columns_per_s": plain_rate,
        "plain_gpu_solves_per_s": plain_rate * J.D,
        "gated_vs_plain": gated_rate / plain_rate,
        "accepted_frac": acc_all / n_serve,
        "plain_routed_frac": n_plain / n_serve,
        "sampled_frac": n_sampled / n_serve,
        "accepted_err_quantiles": ({p: float(np.quantile(e_all, p / 100)) for p in (50, 90, 99)}
                                   if e_all.size else None),
        "accepted_violation_frac": float((e_all > J.JTOL).mean()) if e_all.size else None,
        "calibration": cal_info,
        "calibration_err_quantiles_all": {p: float(np.quantile(cal_err, p / 100)) for p in (50, 90, 99, 99.9)},
        "monitor": gate.monitor.stats(),
        "tolerance_sweep": [
            {"tol": tol, "calibration_accept_frac": inf["accept_frac"], "served_accept_frac": sweep_acc[j] / n_serve,
             "accepted_violation_frac": sweep_viol[j] / max(sweep_acc[j], 1),
             "columns_per_s": n_serve / sweep_s[j], "vs_plain": (n_serve / sweep_s[j]) / plain_rate}
            for j, (tol, _, inf) in enumerate(sweep)],
        "mlp_inference_columns_per_s": mlp_rate,
        "roofline": {"bound": "mfma", "kernel": "k_gemm (disort_mlp_infer, 65,536 rows, padded widths "
                     + "x".join(str(w) for w in W) + ")",
                     "achieved": pad_flops / gemm_s / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": pad_flops / gemm_s / 1e12 / 2500.0, "traffic": None},
        "training": {"columns": n, "epochs": args.jac_epochs, "seconds_per_member": train_s, "final_loss": hist,
                     "tflops": train_tflops, "dataset_seconds": gen_s, "retrained_members": restarts},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = jac_cpu_baseline(ctx, args.jac_cpu_sample)
    gate.close()
    for _, g, _ in sweep:
        g.close()
    for m in members:
        m.close()
    return line


def jac_cpu_baseline(ctx, sample: int) -> dict:
    """The oracle's temperature Jacobian (the analytic linearisation of
    DisortFluxSolver::computeThermalJacobian, src/FluxSolver.cpp:1294-1497)
    on `sample` fresh columns, one column per task on all host threads, and
    the plain GPU Jacobian of the same columns against it."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from disort_amd import jac_surrogate as J
    from disort_amd import surrogate as S

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    hb = S.c5_columns(sample, seed=9001)
    threads = cpu_threads()
    oracle_lib.flux_jacobian_one(hb, 0)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        ref = list(ex.map(lambda i: oracle_lib.flux_jacobian_one(hb, i)[1], range(sample)))
    sec = time.perf_counter() - t0
    ref = np.stack(ref)
    cols = J.JacColumns(ctx, hb)
    cols.plain()
    ctx.sync()
    got = cols.jac.get(np.float64, (sample, 6, S.NLEV, J.D))
    cols.close()
    # The parity bar of the Jacobian tests: 1e-9 relative with an absolute
    # floor of 1e-12 of the column's largest derivative (ratio <= 1 passes).
    bar = 1e-9 * np.abs(ref) + 1e-12 * np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    ratio = np.abs(got - ref) / bar
    # Entries outside it are arbitrated as in the tests: against the same
    # linearisation in x87 extended precision x, within the bar of x or no
    # further from x than 4 times the FP64 oracle's largest error on that
    # degree of freedom (thin deep layers make the algorithm ill-conditioned).
    bad_cols = np.nonzero((ratio > 1.0).any(axis=(1, 2, 3)))[0]
    still = 0
    for i in bad_cols:
        x = oracle_lib.flux_jacobian_one_extended(hb, int(i))[1]
        col_err = np.abs(ref[i] - x).max(axis=(0, 1), keepdims=True)
        fl = 1e-12 * np.abs(ref[i]).max()
        ok = np.abs(got[i] - x) <= np.maximum(1e-9 * np.abs(x) + fl, 4.0 * col_err)
        still += int(((ratio[i] > 1.0) & ~ok).sum())
    big = np.abs(ref) > 1e-6 * np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    big3 = np.abs(ref) > 1e-3 * np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    return {"value": sample / sec, "unit": "columns/s", "cores": threads, "kind": "port",
            "sample": f"{sample} fresh columns; the oracle's Jacobian (CPU restatement of "
                      f"computeThermalJacobian), one column per task on {threads} threads on {cpu_model()}",
            "plain_gpu_vs_oracle": {"bar": "|gpu - oracle| <= 1e-9 |oracle| + 1e-12 max|oracle of the column|",
                                    "frac_entries_within": float((ratio <= 1.0).mean()),
                                    "max_ratio_to_bar": float(ratio.max()),
                                    "columns_arbitrated_extended": int(bad_cols.size),
                                    "entries_failing_after_arbitration": still,
                                    "max_rel_of_entries_above_1e-6_of_column_max": float(
                                        (np.abs(got - ref)[big] / np.abs(ref)[big]).max()),
                                    "max_rel_of_entries_above_1e-3_of_column_max": float(
                                        (np.abs(got - ref)[big3] / np.abs(ref)[big3]).max()),
                                    "max_abs_over_column_max": float(
                                        (np.abs(got - ref) / np.abs(ref).max(axis=(1, 2, 3), keepdims=True)).max())}}


def bench_mlp(args) -> dict:
    """The surrogate network's inference GEMMs alone (disort_mlp_infer: five
    k_gemm launches per call, widths 640 x 512 x 512 x 512 x 512 x 192) on
    resident 65,536-row BF16 batches: the measurement behind the C5 line's
    roofline entry.  Random weights; the arithmetic does not depend on them."""
    import numpy as np

    from disort_amd import surrogate as S
    from disort_amd._capi import Context

    ctx = Context(0)
    rows = 65536
    m = S.MLP(ctx, S.WIDTHS_TS, max_batch=rows)
    m.set_params(m.init_params(1, S.N_FEAT_TS, S.N_OUT))
    x = S.DevArray(ctx, rows * S.WIDTHS_TS[0] * 2)
    y = S.DevArray(ctx, rows * S.WIDTHS_TS[-1] * 4)
    r = np.random.Generator(np.random.Philox(3)).normal(0, 1, rows * S.WIDTHS_TS[0]).astype(np.float32)
    ctx.h2d(x.ptr, (r.view(np.uint32) >> 16).astype(np.uint16))
    for _ in range(args.warmup):
        m.infer(x.ptr, rows, y.ptr)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.infer(x.ptr, rows, y.ptr)
    ctx.sync()
    sec = (time.perf_counter() - t0) / args.steps
    flops = 2.0 * rows * sum(a * b for a, b in zip(S.WIDTHS_TS[:-1], S.WIDTHS_TS[1:]))
    line = {"metric": "surrogate MLP inference rows/sec (BF16 MFMA, FP32 accumulate)", "value": rows / sec,
            "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
            "higher_is_better": True, "dtype": "bf16", "data": "synthetic (random weights and inputs)",
            "config": {"workload": "disort_mlp_infer, 65,536 rows", "widths": list(S.WIDTHS_TS)},
            "roofline": {"bound": "mfma", "kernel": "k_gemm (5 launches per row batch)", "achieved": flops / sec / 1e12,
                         "peak": 2500.0, "unit": "TFLOP/s", "frac": flops / sec / 1e12 / 2500.0, "traffic": None}}
    m.close()
    return line


def bench_c3(args) -> dict:
    """C3: the general solver (azimuthal modes, user intensities, RPV, pseudo-
    spherical beam, Nakajima-Tanaka) on 4,096 columns resident in HBM."""
    import ctypes as C

    import numpy as np

    from disort_amd import flops, general as G, workloads
    from disort_amd._capi import Context

    n = args.columns
    gb = workloads.c3(n)
    ctx = Context(0)
    dev = G.DeviceGeneralBatch(ctx, gb)
    for _ in range(args.warmup):
        dev.execute()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.execute()
    ctx.sync()
    sec = (time.perf_counter() - t0) / args.steps
    ctx.set_timing(True)
    dev.execute()
    stage = ctx.stage_times()
    ctx.set_timing(False)
    out = dev.download()
    assert (out.status == 0).all(), "non-zero column status"
    names = ["setup", "k_gsurf_b", "k_glayer", "k_gboundary", "k_gflux", "k_gucoef", "k_guser", "k_gaccum+corrections"]
    stage_ms = dict(zip(names, stage))
    # Layers above the absorption cut-off per column (the boundary problem's size).
    abst = np.cumsum((1.0 - gb.ssa) * gb.dtau, axis=1)
    ncut = np.minimum((abst < 10.0).sum(axis=1) + 1, gb.nlyr)
    model = {"k_glayer": 0.0, "k_gboundary": 0.0, "k_guser": 0.0}
    for nc_ in np.unique(ncut):
        cnt = int((ncut == nc_).sum())
        for k, v in flops.general_column_flops(gb.nstr, gb.nlyr, gb.numu, gb.ntau, int(nc_)).items():
            model[k] += cnt * v
    fp = {k: model[k] / (stage_ms[k] * 1e-3) / 1e12 for k in model}
    dom = max(model, key=lambda k: stage_ms[k])
    line = {
        "metric": "DISORT general-solver columns/sec (C3: nlyr=80, nstr=32, 32 modes, 16x8 user angles, fp64)",
        "value": n / sec,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox draws, disort_amd.workloads.c3)",
        "config": {"workload": f"C3 general solver: {n} columns x nlyr=80 x nstr=32, 32 Fourier modes, "
                               "16 polar x 8 azimuth user angles at 81 levels, RPV, pseudo-spherical, NT",
                   "columns": n, "nlyr": gb.nlyr, "nstr": gb.nstr, "numu": gb.numu, "nphi": gb.nphi,
                   "ntau": gb.ntau, "mean_ncut": float(ncut.mean()),
                   "path": "disort_general_solve on device arrays (DISORT_MEM_DEVICE)"},
        "roofline": {"bound": "fp64", "kernel": dom, "achieved": fp[dom], "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": fp[dom] / PEAK_FP64_TFLOPS, "traffic": None},
        "kernel_fp64_tflops": fp,
        "stage_ms": stage_ms,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib

        l = oracle_lib.lib()
        l.oracle_general_batch.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        l.oracle_general_batch.restype = C.c_int64
        threads = cpu_threads()
        sample = min(n, args.cpu_columns)
        ref = workloads.c3(sample)
        st = ref.c_struct()
        secs = C.c_double()
        l.oracle_general_batch(C.byref(st), 0, min(sample, threads), threads, C.byref(secs))  # warm-up
        failed = l.oracle_general_batch(C.byref(st), 0, sample, threads, C.byref(secs))
        k = min(sample, 8)
        err = float(np.max(np.abs(out.inten[:k] - ref.inten[:k]) / (np.abs(ref.inten[:k]) + 1e-12 * np.abs(ref.inten[:k]).max())))
        line["cpu_baseline"] = {"value": (sample - failed) / secs.value, "unit": "columns/s", "cores": threads,
                                "kind": "port", "sample": f"first {sample} C3 columns, oracle general solver, "
                                f"OpenMP static loop over columns ({secs.value:.2f} s)",
                                "max_rel_diff_vs_gpu_first_8_columns": err}
    dev.close()
    return line


def copy_probe(mib: int) -> None:
    """One elementwise FP64 copy of `mib` MiB on cuda:0: its FETCH_SIZE and
    WRITE_SIZE counters calibrate the counters of the kernels that follow."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"copy probe: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)
    del src, dst


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--bands", type=int, default=10000)
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5", "jac", "mlp"), default="c2",
                    help="c2: the headline spectral flux solve (default); c3: the general solver with "
                         "intensities; c4: ragged autotuned batch; c5: learned surrogate behind the error gate; "
                         "jac: learned temperature Jacobians behind the same gate; "
                         "mlp: the surrogate network's inference GEMMs alone")
    ap.add_argument("--columns", type=int, default=4096, help="C3 columns")
    ap.add_argument("--cpu-columns", type=int, default=64, help="C3 columns timed on the CPU")
    ap.add_argument("--train-columns", type=int, default=8388608, help="C5 training columns")
    ap.add_argument("--epochs", type=int, default=120, help="C5 training epochs per member")
    ap.add_argument("--serve-columns", type=int, default=1048576, help="C5 fresh columns served")
    ap.add_argument("--members", type=int, default=2, help="C5 ensemble members (the gate's spread is theirs)")
    ap.add_argument("--sample-frac", type=float, default=0.005,
                    help="C5 sampled re-solve fraction (0.5%% of a 524,288-column call: 2,600 monitored columns)")
    ap.add_argument("--jac-train-columns", type=int, default=2097152, help="Jacobian surrogate training columns")
    ap.add_argument("--jac-epochs", type=int, default=150, help="Jacobian surrogate epochs per member")
    ap.add_argument("--jac-serve-columns", type=int, default=262144, help="Jacobian surrogate columns served")
    ap.add_argument("--jac-cpu-sample", type=int, default=2048, help="columns of the oracle Jacobian timed")
    ap.add_argument("--jac-sample-frac", type=float, default=0.002,
                    help="Jacobian gate's sampled re-solve fraction (a plain Jacobian costs nlyr+2 solves; "
                         "0.2%% of a 65,536-column call still gives the monitor 131 columns)")
    ap.add_argument("--instances", type=int, default=50000, help="C4 batch size")
    ap.add_argument("--cpu-sample", type=int, default=2000, help="C4 instances timed on the CPU")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one FP64 copy of this many MiB (reads and writes that many bytes), "
                         "the calibration kernel of HBM counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: DISORT instances are independent and shard "
            "across GPUs without communication; run with --gpus 1.",
            file=sys.stderr,
        )
        return 2
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    if args.workload == "c3":
        print(json.dumps(bench_c3(args)))
        return 0
    if args.workload == "c4":
        print(json.dumps(bench_c4(args)))
        return 0
    if args.workload == "mlp":
        print(json.dumps(bench_mlp(args)))
        return 0
    if args.workload == "c5":
        print(json.dumps(bench_c5(args)))
        return 0
    if args.workload == "jac":
        print(json.dumps(bench_jac(args)))
        return 0

    from disort_amd import flops, workloads
    from disort_amd._capi import Context, DeviceFluxBatch

    nstr, nlyr = 16, 60
    batch = workloads.c2(args.bands, nlyr=nlyr, nstr=nstr)
    ctx = Context(0)
    dev = DeviceFluxBatch(ctx, batch)
    # The plan's autotuner picks the kernel variant and the bucket schedule
    # (streams x chunks) on this batch before the timed region.
    tuning = dev.autotune(reps=3)

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

    # Per-kernel device time with hipEvents on the context stream (separate
    # passes), each kernel as one launch over the whole batch: the plan is
    # set back to one bucket on the context stream for these passes.
    dev.set_streams(1)
    dev.set_chunks(nstr, 1)
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
            "schedule": {"streams": tuning.get("streams"), "chunks": tuning.get("chunks"),
                         "variant": tuning.get("dominant_variant", tuning["buckets"][0]["chosen"]),
                         "serial_ms": tuning.get("serial_ms"),
                         "tuned_ms": tuning.get("best_ms")},
        },
        "roofline": {
            "bound": "fp64",
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
        gpu_out = out.out.copy()
        line["cpu_baseline"], ref = cpu_baseline(batch)
        line["err_vs_cpu_oracle"] = output_errors(gpu_out, ref, batch)
    dev.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
