#!/usr/bin/env python3
"""Headline benchmark: particle-accelerations/s at N = 1M Plummer, theta = 0.5.

One step = one pass of the force path over the resident input: build the
tree from positions already in HBM (bbox, Morton, radix sort, topology,
upward pass) and compute every particle's self-acceleration.  Timed with HIP
events on the stream the library launches on; the dominant kernel's average
duration comes from the library's own event profiler over the same timed
steps.  The CPU baseline is the C restatement in oracle/ (the reference's
algorithm, OpenMP over targets) timed on this host on a bounded sample.

Prints one JSON line (see --help for the knobs; defaults run in ~1-2 min).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particle-accels/s at 1M Plummer θ=0.5; FP32 vs plain, rms rel err ≤1e-3"
FLOPS_PER_INTERACTION = 20  # 3 sub, 3 mul + 3 add (r2, eps2), sqrt, div, 2 mul r^-3, 1 mul m, 3 mul + 3 add
FLOPS_PER_QUAD_CELL = 54    # cell monopole + quadrupole term as written in k_group_eval (FMA = 2)
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--theta", type=float, default=0.5)
    p.add_argument("--softening", type=float, default=0.01)
    p.add_argument("--leaf", type=int, default=32)
    p.add_argument("--form", choices=["fp32", "fp64"], default="fp32",
                   help="form whose throughput is `value` (the other is reported beside it)")
    p.add_argument("--accuracy-targets", type=int, default=4096)
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-plain", action="store_true", help="skip timing the plain form beside the FP32 form")
    p.add_argument("--group-size", type=int, default=0, help="FP32 form targets per group (0: autotuned)")
    p.add_argument("--stack-cap", type=int, default=0, help="FP32 form LDS stack entries (0: autotuned)")
    p.add_argument("--multipole", type=int, default=2, help="FP32 form cells: 1 monopole, 2 quadrupole")
    p.add_argument("--no-kahan", action="store_true")
    return p.parse_args()


def cpu_baseline(x, m, theta, eps, leaf, budget_s):
    """Oracle build + accelerations on a bounded random sample of targets,
    scaled to particle-accelerations/s for the whole job."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    n = len(x)
    t0 = time.perf_counter()
    tree = O.build(x, m, leaf)
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(123)
    order = rng.permutation(n)
    done, t_acc = 0, 0.0
    chunk = 2000
    while done < n and t_build + t_acc < budget_s:
        sel = order[done:done + chunk]
        t0 = time.perf_counter()
        O.accelerations(tree, x, m, x[sel], softening=eps, theta=theta, nthreads=threads)
        t_acc += time.perf_counter() - t0
        done += len(sel)
        chunk = min(chunk * 2, 200_000)
    per_target = t_acc / done
    t_total = t_build + per_target * n
    cpu = "unknown"
    try:
        cpu = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        pass
    return {
        "value": n / t_total,
        "unit": "particle-accels/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle build of all {n} particles ({t_build:.2f} s, 1 thread) + accelerations of {done} "
                  f"random targets ({t_acc:.2f} s, {threads} OpenMP threads), extrapolated to {n} targets; "
                  f"host CPU {cpu}",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the target-sharded multi-GPU design is "
              "described in DESIGN.md and not measured here", file=sys.stderr)
        sys.exit(2)
    import torch

    from bhut_amd import _native, workloads

    dev = torch.device("cuda:0")
    ctx = _native.ctx()
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)

    N, theta, eps, G, leaf = args.n, args.theta, args.softening, 1.0, args.leaf
    log(f"bench: generating Plummer N={N} seed=1")
    x, m = workloads.plummer(N, seed=1)
    pos_t = torch.from_numpy(x).to(dev)
    mass_t = torch.from_numpy(m).to(dev)
    out_t = torch.empty((N, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    tuning = None
    if args.form == "fp32" and (args.group_size == 0 or args.stack_cap == 0):
        from bhut_amd import autotune
        log("bench: autotuning the FP32 form (group size, stack depth) on this input")
        probe = _native.DeviceTree.build_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, leaf, context=ctx)
        best, table = autotune.tune_f32(probe, theta, eps, G, multipole=args.multipole,
                                        kahan=not args.no_kahan,
                                        group_sizes=(args.group_size,) if args.group_size else (64, 128),
                                        stack_caps=(args.stack_cap,) if args.stack_cap else (1024, 2048))
        del probe
        tuning = {"chosen": best, "table": table}
        log(f"bench: autotune -> {best}")
        f32_opts = _native.F32Opts(best["group_size"], 0 if args.no_kahan else 1, best["stack_cap"], args.multipole)
    else:
        f32_opts = _native.F32Opts(args.group_size, 0 if args.no_kahan else 1, args.stack_cap, args.multipole)

    def step(form):
        t = _native.DeviceTree.build_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, leaf, context=ctx)
        if form == "fp64":
            t.accel_self_dev(out_t.data_ptr(), theta, eps, G)
        else:
            t.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, f32_opts)
        return t

    def timed(form):
        for _ in range(args.warmup):
            step(form)
        torch.cuda.synchronize()
        ctx.profile(True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            step(form)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        prof = ctx.profile_summary()
        ctx.profile(False)
        return e0.elapsed_time(e1) / args.steps, wall / args.steps, prof

    forms = [args.form] + ([] if args.no_plain or args.form == "fp64" else ["fp64"])
    res = {}
    for form in forms:
        log(f"bench: timing form {form}: {args.warmup} warmup + {args.steps} steps")
        ms, wall_s, prof = timed(form)
        res[form] = dict(ms=ms, wall_ms=wall_s * 1e3, prof=prof, acc=out_t.cpu().numpy())
        log(f"bench: {form} {ms:.3f} ms/step (wall {wall_s * 1e3:.3f}); kernels: {prof}")

    # interaction counts (exact device counters / stats), outside the timed region
    t = step("fp64")
    cnt = t.counters(None, theta).sum(axis=0)
    inter64 = int(cnt[2] + cnt[4])
    stats32 = None
    if "fp32" in res:
        out2 = torch.empty_like(out_t)
        stats32 = t.accel_f32_self_dev(out2.data_ptr(), theta, eps, G, f32_opts)
        torch.cuda.synchronize()

    # accuracy against FP64 direct summation on random targets
    rng = np.random.default_rng(7)
    sel = np.sort(rng.choice(N, min(args.accuracy_targets, N), replace=False))
    ref = _native.direct(x, m, x[sel], softening=eps, G=G)
    nr = np.linalg.norm(ref, axis=1)
    acc_err = {}
    for form in res:
        e = np.linalg.norm(res[form]["acc"][sel] - ref, axis=1) / nr
        acc_err[form] = (float(np.sqrt(np.mean(e ** 2))), float(e.max()))
        log(f"bench: {form} rms rel err {acc_err[form][0]:.3e} max {acc_err[form][1]:.3e} "
            f"({len(sel)} targets vs FP64 direct sum)")

    # monopole-only variant of the FP32 form, for the accuracy table only
    if "fp32" in res:
        out3 = torch.empty_like(out_t)
        t.accel_f32_self_dev(out3.data_ptr(), theta, eps, G, _native.F32Opts(0, 1, 0, 1))
        torch.cuda.synchronize()
        e = np.linalg.norm(out3.cpu().numpy()[sel] - ref, axis=1) / nr
        acc_err["fp32_monopole_only"] = (float(np.sqrt(np.mean(e ** 2))), float(e.max()))
        e = np.linalg.norm(res["fp32"]["acc"][sel] - res["fp64"]["acc"][sel], axis=1) / \
            np.linalg.norm(res["fp64"]["acc"][sel], axis=1) if "fp64" in res else None
        if e is not None:
            acc_err["fp32_vs_plain"] = (float(np.sqrt(np.mean(e ** 2))), float(e.max()))

    head = args.form
    ms = res[head]["ms"]
    if head == "fp64":
        kname, peak = "k_walk_f64", PEAK_FP64_TFLOPS
        inter = inter64
        flops = FLOPS_PER_INTERACTION * inter64
    else:
        kname, peak = "k_group_eval_f32", PEAK_FP32_TFLOPS
        inter = int(stats32[0] + stats32[1])
        flops = FLOPS_PER_INTERACTION * int(stats32[1]) + FLOPS_PER_QUAD_CELL * int(stats32[0])
    launches, kms = res[head]["prof"].get(kname, (0, 0.0))
    kavg = kms / max(launches, 1)
    achieved = flops / (kavg * 1e-3) / 1e12 if kavg > 0 else None
    # HBM bytes per launch come from the committed FETCH_SIZE/WRITE_SIZE
    # passes (rocprofv3) of this same workload, when it matches
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "plummer1m_group_eval_traffic.json")
    if head == "fp32" and N == 1_000_000 and leaf == 32 and f32_opts.group_size in (0, 64) and os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": N / (ms * 1e-3),
        "unit": "particle-accels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if head == "fp32" else "fp64",
        "data": "synthetic",
        "config": {"workload": "plummer_self_accelerations", "n": N, "theta": theta, "softening": eps,
                   "leaf_size": leaf, "G": G, "seed": 1, "step": "tree build + self accelerations",
                   "form": head},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": (achieved / peak) if achieved else None, "traffic": traffic,
                     "kernel": kname, "kernel_avg_ms": kavg,
                     "flops_per_launch": flops, "interactions_per_launch": inter},
        "accuracy": {f: {"rms_rel_err": acc_err[f][0], "max_rel_err": acc_err[f][1]} for f in acc_err},
        "interactions_per_s": inter / (ms * 1e-3),
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in res[head]["prof"].items()},
        "gpu": torch.cuda.get_device_name(0),
        "f32_options": None if head != "fp32" else {
            "group_size": f32_opts.group_size, "stack_cap": f32_opts.stack_cap, "kahan": f32_opts.kahan,
            "multipole": f32_opts.multipole, "autotune": tuning},
    }
    if head == "fp32" and "fp64" in res:
        line["plain_fp64"] = {"value": N / (res["fp64"]["ms"] * 1e-3), "ms_per_step": res["fp64"]["ms"],
                              "interactions_per_launch": inter64,
                              "kernels_ms_per_step": {k: v[1] / args.steps for k, v in res["fp64"]["prof"].items()}}
        line["speedup_fp32_vs_plain"] = res["fp64"]["ms"] / ms
    if not args.no_cpu:
        log(f"bench: CPU baseline (oracle port), budget {args.cpu_seconds:.0f} s")
        line["cpu_baseline"] = cpu_baseline(x, m, theta, eps, leaf, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
