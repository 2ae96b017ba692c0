#!/usr/bin/env python3
"""Headline benchmark: particle-accelerations/s at N = 1M Plummer, theta = 0.5.

One step = one pass of the force path over the resident input: build the
tree from positions already in HBM (bbox, Morton, radix sort, topology,
upward pass) and compute every particle's self-acceleration.  Timed with HIP
events on the stream the library launches on; the dominant kernel's average
duration comes from the library's own event profiler over the same timed
steps.  The CPU baseline is the C restatement in oracle/ (the reference's
algorithm, OpenMP over targets) timed on this host on a bounded sample.

--workload selects the other measured configurations (each prints its own
JSON line; the default run prints only the headline line):
  plummer    1M Plummer self accelerations (+ a 250k-target external query)
  ensemble   512 independent Plummer systems, one forest vs 512 single calls
  diskhalo   16M exponential disk + Hernquist halo, theta in {0.3, 0.5, 0.7},
             leaf / group size / stack depth autotuned per theta
  leapfrog   4M Plummer, 100 KDK steps, refit every step, rebuild every 10,
             steps captured in HIP graphs; FP32 vs FP64 forces
  stream     16M host-resident targets against a 1M Plummer tree, streamed in
             chunks on two streams vs one call vs the device-resident walk
  copy       three 256 MiB device-to-device copies: the known-size kernel that
             calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
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
COMPUTE_UNIT = "vector ALU (VALU); no MFMA is used"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=["plummer", "ensemble", "diskhalo", "leapfrog", "stream", "copy"],
                   default="plummer")
    p.add_argument("--n", type=int, default=0, help="particles (0: the workload's own size)")
    p.add_argument("--theta", type=float, default=0.5)
    p.add_argument("--softening", type=float, default=0.01)
    p.add_argument("--leaf", type=int, default=32)
    p.add_argument("--form", choices=["fp32", "fp64"], default="fp32",
                   help="form whose throughput is `value` (the other is reported beside it)")
    p.add_argument("--accuracy-targets", type=int, default=4096)
    p.add_argument("--query-targets", type=int, default=250_000, help="external query size (plummer; 0: skip)")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-plain", action="store_true", help="skip timing the plain form beside the FP32 form")
    p.add_argument("--group-size", type=int, default=0, help="FP32 form targets per group (0: autotuned)")
    p.add_argument("--stack-cap", type=int, default=0, help="FP32 form LDS stack entries (0: autotuned)")
    p.add_argument("--multipole", type=int, default=2, help="FP32 form cells: 1 monopole, 2 quadrupole")
    p.add_argument("--no-kahan", action="store_true")
    p.add_argument("--systems", type=int, default=512, help="ensemble: number of systems")
    p.add_argument("--lf-steps", type=int, default=100, help="leapfrog: steps")
    p.add_argument("--no-graph", action="store_true", help="leapfrog: eager steps instead of graphs")
    p.add_argument("--list-budget-gib", type=float, default=0.0,
                   help="FP32 form: bytes of interaction lists held at once, GiB (0: the library default)")
    return p.parse_args()


# ---------------------------------------------------------------- helpers
class Env:
    def __init__(self):
        import torch

        from bhut_amd import _native
        self.torch = torch
        self.nat = _native
        self.dev = torch.device("cuda:0")
        self.ctx = _native.ctx()
        # one explicit (default-priority) stream for torch and the library, so
        # the HIP events bracketing a step are ordered with its kernels (the
        # library cannot be pointed at the legacy null stream)
        self.stream = torch.cuda.Stream(self.dev)
        torch.cuda.set_stream(self.stream)
        self.ctx.set_stream(self.stream.cuda_stream)

    def upload(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)

    def sync(self):
        self.torch.cuda.synchronize()

    def timed(self, fn, steps, warmup):
        """(ms per step from HIP events, wall ms per step, kernel profile)."""
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        self.ctx.profile(True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(self.stream)
        for _ in range(steps):
            fn()
        e1.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        prof = self.ctx.profile_summary()
        self.ctx.profile(False)
        return e0.elapsed_time(e1) / steps, wall * 1e3 / steps, prof


def rel_err(got, ref):
    nr = np.linalg.norm(ref, axis=1)
    e = np.linalg.norm(got - ref, axis=1) / nr
    return {"rms_rel_err": float(np.sqrt(np.mean(e ** 2))), "max_rel_err": float(e.max()), "targets": int(len(e))}


def f32_options(args, best=None):
    from bhut_amd._native import F32Opts
    gs = best["group_size"] if best else args.group_size
    sc = best["stack_cap"] if best else args.stack_cap
    return F32Opts(gs, 0 if args.no_kahan else 1, sc, args.multipole)


def roofline_f32(prof, steps, stats):
    """FLOP rate of the evaluation kernel over one step's groups (one launch
    per step unless the lists were split into batches)."""
    launches, kms = prof.get("k_group_eval_f32", (0, 0.0))
    kstep = kms / max(steps, 1)
    flops = FLOPS_PER_INTERACTION * int(stats[1]) + FLOPS_PER_QUAD_CELL * int(stats[0])
    achieved = flops / (kstep * 1e-3) / 1e12 if kstep > 0 else None
    # compute-bound on the vector ALUs (nothing here is a dense contraction
    # for MFMA): priced against the FP32 vector figure
    return {"bound": "valu", "compute_unit": COMPUTE_UNIT, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS if achieved else None, "traffic": None,
            "kernel": "k_group_eval_f32", "kernel_avg_ms": kms / max(launches, 1),
            "launches_per_step": launches / max(steps, 1), "kernel_ms_per_step": kstep,
            "flops_per_step": flops, "cell_interactions_per_step": int(stats[0]),
            "pair_interactions_per_step": int(stats[1]), "interactions_per_launch": int(stats[0] + stats[1])}


def pair_loop_ceiling():
    """Pairs/s of the evaluation's leaf-pair loop alone on this device
    (bhut_amd/pair_ceiling, built with the library), or None."""
    exe = os.path.join(ROOT, "bhut_amd", "pair_ceiling")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # reported, not fatal: the bench line stands without it
        print(f"bench: pair-loop ceiling not measured ({e})", file=sys.stderr)
        return None


def host_cpus():
    """(threads to use, description): the CPUs this process may run on
    (affinity), capped by the cgroup CPU quota and by OMP_NUM_THREADS when
    either is set (the host a GPU is leased with grants a share of its
    cores), with the CPU model."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except Exception:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    omp = int(omp) if omp and omp.isdigit() else None
    use = min(v for v in (aff, quota, omp) if v)
    model = "unknown"
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        pass
    return use, {"nproc": aff, "cgroup_cpu_quota": quota, "omp_num_threads": omp, "threads_used": use,
                 "model": model}


def cpu_baseline(x, m, theta, eps, leaf, budget_s):
    """Oracle build + self accelerations on a bounded random sample of
    targets, each timed in 3 runs after a warm-up (median kept), scaled to
    particle-accelerations/s for the whole job."""
    from oracle import oracle as O
    threads, host = host_cpus()
    n = len(x)
    rng = np.random.default_rng(123)
    order = rng.permutation(n)

    def median_of(fn, runs=3):
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts

    tree = O.build(x, m, leaf)  # warm-up
    t_build, builds = median_of(lambda: O.build(x, m, leaf))
    # sample size: a probe of 2000 targets sets it so the three timed runs
    # take about three quarters of the budget
    probe = np.sort(order[:2000])
    t0 = time.perf_counter()
    O.accelerations_subset(tree, x, m, probe, softening=eps, theta=theta, nthreads=threads)
    per = (time.perf_counter() - t0) / len(probe)
    k = int(min(n, max(2000, 0.25 * budget_s / per)))
    sel = np.sort(order[:k])
    t_acc, accs = median_of(lambda: O.accelerations_subset(tree, x, m, sel, softening=eps, theta=theta,
                                                           nthreads=threads))
    t_total = t_build + t_acc / k * n
    # the serial loop of the reference (one target after another on one
    # core): a tenth of the budget over the three runs
    k1 = int(min(n, max(200, 0.03 * budget_s / (per * threads))))
    sel1 = np.sort(order[:k1])
    t_acc1, _ = median_of(lambda: O.accelerations_subset(tree, x, m, sel1, softening=eps, theta=theta,
                                                         nthreads=1))
    t_total1 = t_build + t_acc1 / k1 * n
    return {
        "value": n / t_total,
        "unit": "particle-accels/s",
        "cores": threads,
        "kind": "port",
        "host": host,
        "sample": f"oracle build of all {n} particles (median of 3: {t_build:.2f} s, 1 thread) + self "
                  f"accelerations of {k} random particles (median of 3: {t_acc:.2f} s, runs "
                  f"{', '.join(f'{v:.2f}' for v in accs)} s, {threads} OpenMP threads), extrapolated to {n} "
                  f"targets; host CPU {host['model']}, {host['nproc']} CPUs in the affinity mask, "
                  f"cgroup quota {host['cgroup_cpu_quota']}, OMP_NUM_THREADS {host['omp_num_threads']}",
        "single_thread": {"value": n / t_total1, "unit": "particle-accels/s", "cores": 1,
                          "sample": f"the same build + {k1} random particles (median of 3: {t_acc1:.2f} s, "
                                    f"1 thread), extrapolated to {n} targets"},
    }


def base_line(args, value, ms, dtype, config):
    return {"metric": METRIC, "value": value, "unit": "particle-accels/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


# ---------------------------------------------------------------- 1M Plummer
def workload_plummer(args, env):
    from bhut_amd import workloads
    nat, torch = env.nat, env.torch
    N = args.n or 1_000_000
    theta, eps, G, leaf = args.theta, args.softening, 1.0, args.leaf
    log(f"bench: generating Plummer N={N} seed=1")
    x, m = workloads.plummer(N, seed=1)
    pos_t, mass_t = env.upload(x), env.upload(m)
    out_t = torch.empty((N, 3), dtype=torch.float64, device=env.dev)
    env.sync()

    tuning = None
    if args.form == "fp32" and (args.group_size == 0 or args.stack_cap == 0):
        from bhut_amd import autotune
        log("bench: autotuning the FP32 form (group size, stack depth) on this input")
        probe = nat.DeviceTree.build_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, leaf, context=env.ctx)
        best, table = autotune.tune_f32(probe, theta, eps, G, multipole=args.multipole, kahan=not args.no_kahan,
                                        group_sizes=(args.group_size,) if args.group_size else (64, 128),
                                        stack_caps=(args.stack_cap,) if args.stack_cap else (512, 1024),
                                        out_ptr=out_t.data_ptr())
        del probe
        tuning = {"chosen": best, "table": table}
        log(f"bench: autotune -> {best}")
        opts = f32_options(args, best)
    else:
        opts = f32_options(args)

    def step(form):
        t = nat.DeviceTree.build_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, leaf, context=env.ctx)
        if form == "fp64":
            t.accel_self_dev(out_t.data_ptr(), theta, eps, G)
        else:
            t.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, opts)
        return t

    forms = [args.form] + ([] if args.no_plain or args.form == "fp64" else ["fp64"])
    res = {}
    for form in forms:
        log(f"bench: timing form {form}: {args.warmup} warmup + {args.steps} steps")
        ms, wall, prof = env.timed(lambda: step(form), args.steps, args.warmup)
        res[form] = dict(ms=ms, wall_ms=wall, prof=prof, acc=out_t.cpu().numpy())
        log(f"bench: {form} {ms:.3f} ms/step (wall {wall:.3f}); kernels: {prof}")

    # interaction counts (exact device counters / stats), outside the timed region
    t = step("fp64")
    cnt = t.counters(None, theta).sum(axis=0)
    inter64 = int(cnt[2] + cnt[4])
    stats32 = None
    if "fp32" in res:
        out2 = torch.empty_like(out_t)
        stats32 = t.accel_f32_self_dev(out2.data_ptr(), theta, eps, G, opts)
        env.sync()

    # accuracy against FP64 direct summation on random targets
    rng = np.random.default_rng(7)
    sel = np.sort(rng.choice(N, min(args.accuracy_targets, N), replace=False))
    ref = nat.direct(x, m, x[sel], softening=eps, G=G)
    acc_err = {form: rel_err(res[form]["acc"][sel], ref) for form in res}
    for form, e in acc_err.items():
        log(f"bench: {form} rms rel err {e['rms_rel_err']:.3e} max {e['max_rel_err']:.3e} vs FP64 direct sum")
    if "fp32" in res:
        out3 = torch.empty_like(out_t)
        t.accel_f32_self_dev(out3.data_ptr(), theta, eps, G, nat.F32Opts(0, 1, 0, 1))
        env.sync()
        acc_err["fp32_monopole_only"] = rel_err(out3.cpu().numpy()[sel], ref)
        if "fp64" in res:
            acc_err["fp32_vs_plain"] = rel_err(res["fp32"]["acc"][sel], res["fp64"]["acc"][sel])

    head = args.form
    ms = res[head]["ms"]
    if head == "fp64":
        launches, kms = res[head]["prof"].get("k_walk_f64", (0, 0.0))
        kavg = kms / max(launches, 1)
        flops = FLOPS_PER_INTERACTION * inter64
        achieved = flops / (kavg * 1e-3) / 1e12 if kavg > 0 else None
        roof = {"bound": "valu", "compute_unit": COMPUTE_UNIT, "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS if achieved else None, "traffic": None, "kernel": "k_walk_f64",
                "kernel_avg_ms": kavg, "flops_per_launch": flops, "interactions_per_launch": inter64}
        inter = inter64
    else:
        roof = roofline_f32(res[head]["prof"], args.steps, stats32)
        inter = roof["interactions_per_launch"]
        # HBM bytes per launch come from the committed FETCH_SIZE/WRITE_SIZE
        # passes (rocprofv3) of this same workload, when it matches
        tpath = os.path.join(ROOT, "profiles", "plummer1m_group_eval_traffic.json")
        if N == 1_000_000 and leaf == 32 and opts.group_size in (0, 64) and os.path.exists(tpath):
            roof["traffic"] = json.load(open(tpath)).get("hbm_bytes_per_launch")
            roof["traffic_source"] = ("read from profiles/plummer1m_group_eval_traffic.json: rocprofv3 --pmc "
                                      "FETCH_SIZE and WRITE_SIZE passes of this workload, each scaled by the "
                                      "counter's reading for a copy of known size; not measured in this run")
    line = base_line(args, N / (ms * 1e-3), ms, "fp32" if head == "fp32" else "fp64",
                     {"workload": "plummer_self_accelerations", "n": N, "theta": theta, "softening": eps,
                      "leaf_size": leaf, "G": G, "seed": 1, "step": "tree build + self accelerations", "form": head})
    if head == "fp32" and roof.get("achieved"):
        # the same loop without lists, windows or cells: what this
        # instruction mix can reach on the device (equal masses here)
        ceil = pair_loop_ceiling()
        if ceil:
            tf = ceil["uniform"] * FLOPS_PER_INTERACTION / 1e12
            roof["loop_ceiling"] = {"pairs_per_s": ceil["uniform"], "tflops": tf, "frac": roof["achieved"] / tf,
                                    "mass_pairs_per_s": ceil["mass"], "no_rsq_pairs_per_s": ceil["no_rsq"]}
            if "fma_lane_per_s" in ceil:
                # independent FP32 FMAs alone (2 flops each): the vector ALU's
                # issue ceiling on this device without packed instructions
                roof["valu_fma_ceiling_tflops"] = 2 * ceil["fma_lane_per_s"] / 1e12
            if "rsq_lane_per_s" in ceil and "fma_lane_per_s" in ceil:
                # v_rsq_f32 issue cost in FMA issue slots (one per leaf pair)
                roof["rsq_cost_in_fma_slots"] = ceil["fma_lane_per_s"] / ceil["rsq_lane_per_s"]
            print(f"bench: pair-loop ceiling {tf:.1f} TFLOP/s; evaluation at {roof['achieved'] / tf:.1%} of it",
                  file=sys.stderr)
    line["roofline"] = roof
    line["accuracy"] = acc_err
    line["interactions_per_s"] = inter / (ms * 1e-3)
    line["kernels_ms_per_step"] = {k: v[1] / args.steps for k, v in res[head]["prof"].items()}
    line["gpu"] = torch.cuda.get_device_name(0)
    line["f32_options"] = None if head != "fp32" else {
        "group_size": opts.group_size, "stack_cap": opts.stack_cap, "kahan": opts.kahan,
        "multipole": opts.multipole, "autotune": tuning}
    if head == "fp32" and "fp64" in res:
        launches, kms = res["fp64"]["prof"].get("k_walk_f64", (0, 0.0))
        kavg = kms / max(launches, 1)
        ach64 = FLOPS_PER_INTERACTION * inter64 / (kavg * 1e-3) / 1e12 if kavg > 0 else None
        line["plain_fp64"] = {"value": N / (res["fp64"]["ms"] * 1e-3), "ms_per_step": res["fp64"]["ms"],
                              "interactions_per_launch": inter64,
                              "roofline": {"bound": "valu", "kernel": "k_walk_f64", "kernel_avg_ms": kavg,
                                           "achieved": ach64, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                           "frac": ach64 / PEAK_FP64_TFLOPS if ach64 else None},
                              "kernels_ms_per_step": {k: v[1] / args.steps for k, v in res["fp64"]["prof"].items()}}
        line["speedup_fp32_vs_plain"] = res["fp64"]["ms"] / ms

    if args.query_targets:
        # external query on the resident tree (build excluded): the M targets
        # uniform in [-5, 5]^3, both forms, accuracy against FP64 direct sum
        M = args.query_targets
        rng = np.random.default_rng(11)
        q = rng.uniform(-5.0, 5.0, (M, 3))
        q_t = env.upload(q)
        qo = torch.empty((M, 3), dtype=torch.float64, device=env.dev)
        env.sync()
        qres = {}
        for form in (["fp32"] if head == "fp32" else []) + ([] if args.no_plain and head == "fp32" else ["fp64"]):
            if form == "fp32":
                fn = lambda: t.accel_f32_targets_dev(q_t.data_ptr(), M, qo.data_ptr(), theta, eps, G, opts)
            else:
                fn = lambda: t.accel_targets_dev(q_t.data_ptr(), M, qo.data_ptr(), theta, eps, G)
            qms, _, _ = env.timed(fn, max(1, args.steps // 2), 1)
            qsel = np.sort(rng.choice(M, min(args.accuracy_targets, M), replace=False))
            qref = nat.direct(x, m, q[qsel], softening=eps, G=G)
            qres[form] = {"targets_per_s": M / (qms * 1e-3), "ms": qms,
                          "accuracy": rel_err(qo.cpu().numpy()[qsel], qref)}
            log(f"bench: query {form} {qms:.3f} ms for {M} targets; {qres[form]['accuracy']}")
        line["query"] = {"targets": M, "distribution": "uniform [-5,5]^3, seed 11", "forms": qres}

    if head == "fp32":
        # the drop-in call on host NumPy arrays: positions and masses copied
        # in and accelerations copied out over PCIe around the same step
        import bhut_amd
        kw = dict(theta=theta, softening=eps, G=G, leaf_size=leaf, precision="fp32")
        bhut_amd.accelerations(x, m, **kw)
        reps, t0 = 3, time.perf_counter()
        for _ in range(reps):
            bhut_amd.accelerations(x, m, **kw)
        hms = (time.perf_counter() - t0) / reps * 1e3
        line["host_api"] = {"call": "bhut_amd.accelerations(x, m, precision='fp32') on host arrays",
                            "ms": hms, "particle_accels_per_s": N / (hms * 1e-3)}
        log(f"bench: host-array call {hms:.2f} ms ({N / (hms * 1e-3):.3e} particle-accels/s incl. PCIe)")

    if not args.no_cpu:
        log(f"bench: CPU baseline (oracle port), budget {args.cpu_seconds:.0f} s")
        line["cpu_baseline"] = cpu_baseline(x, m, theta, eps, leaf, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    return line


# ---------------------------------------------------------------- ensemble
def workload_ensemble(args, env):
    from bhut_amd import workloads
    nat, torch = env.nat, env.torch
    theta, eps, G, leaf = args.theta, args.softening, 1.0, args.leaf
    log(f"bench: generating {args.systems} Plummer systems (seeds 1000+i)")
    systems = workloads.ensemble(args.systems)
    counts = np.array([len(s[0]) for s in systems], np.int64)
    off = np.concatenate([[0], np.cumsum(counts)])
    x = np.concatenate([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    N = len(x)
    log(f"bench: ensemble of {len(counts)} systems, {N} particles (min {counts.min()}, max {counts.max()})")
    pos_t, mass_t = env.upload(x), env.upload(m)
    out_t = torch.empty((N, 3), dtype=torch.float64, device=env.dev)
    env.sync()
    opts = f32_options(args, {"group_size": args.group_size or 64, "stack_cap": args.stack_cap or 1024})

    def forest_step(form):
        f = nat.DeviceTree.build_forest_dev(pos_t.data_ptr(), mass_t.data_ptr(), counts, 3, leaf, context=env.ctx)
        if form == "fp64":
            f.accel_self_dev(out_t.data_ptr(), theta, eps, G)
        else:
            f.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, opts)
        return f

    def sequential(form):
        for s in range(len(counts)):
            a, b = int(off[s]), int(off[s + 1])
            t = nat.DeviceTree.build_dev(pos_t[a:b].data_ptr(), mass_t[a:b].data_ptr(), b - a, 3, leaf,
                                         context=env.ctx)
            if form == "fp64":
                t.accel_self_dev(out_t[a:b].data_ptr(), theta, eps, G)
            else:
                t.accel_f32_self_dev(out_t[a:b].data_ptr(), theta, eps, G, opts)

    res = {}
    plan = [("forest_fp32", lambda: forest_step("fp32"), args.steps, args.warmup),
            ("sequential_fp32", lambda: sequential("fp32"), 1, 1),
            ("forest_fp64", lambda: forest_step("fp64"), 1, 1),
            ("sequential_fp64", lambda: sequential("fp64"), 1, 1)]
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(counts), min(16, len(counts)), replace=False))
    sels = [np.sort(rng.choice(int(counts[s]), min(512, int(counts[s])), replace=False)) for s in pick]
    refs = [nat.direct(systems[s][0], systems[s][1], systems[s][0][sel], softening=eps, G=G)
            for s, sel in zip(pick, sels)]
    for name, fn, steps, warm in plan:
        log(f"bench: timing {name}: {warm} warmup + {steps} steps")
        ms, wall, prof = env.timed(fn, steps, warm)
        acc = out_t.cpu().numpy()
        got = np.concatenate([acc[off[s]:off[s + 1]][sel] for s, sel in zip(pick, sels)])
        res[name] = {"ms": ms, "wall_ms": wall, "particle_accels_per_s": N / (ms * 1e-3),
                     "accuracy": rel_err(got, np.concatenate(refs)),
                     "kernels_ms_per_step": {k: v[1] / steps for k, v in prof.items()}}
        log(f"bench: {name} {ms:.1f} ms/step; {res[name]['accuracy']}")
        if name == "forest_fp32":
            prof32 = prof
    f = forest_step("fp32")
    st = f.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, opts)
    env.sync()
    head = res["forest_fp32"]
    line = base_line(args, head["particle_accels_per_s"], head["ms"], "fp32",
                     {"workload": "plummer_ensemble_forest", "systems": int(len(counts)), "n_total": int(N),
                      "n_min": int(counts.min()), "n_max": int(counts.max()), "theta": theta, "softening": eps,
                      "leaf_size": leaf, "G": G, "seeds": "1000+i",
                      "step": "forest build (segmented sort) + FP32 self accelerations, one launch per kernel"})
    line["metric"] = "particle-accels/s, 512-system Plummer ensemble θ=0.5 (batched forest)"
    line["roofline"] = roofline_f32(prof32, args.steps, st)
    line["forms"] = res
    line["speedup_forest_fp32_vs_sequential_plain"] = res["sequential_fp64"]["ms"] / head["ms"]
    line["speedup_forest_fp32_vs_sequential_fp32"] = res["sequential_fp32"]["ms"] / head["ms"]
    line["accuracy_sample"] = f"{len(pick)} random systems x up to 512 random targets vs FP64 direct sum"
    line["cpu_baseline"] = None
    return line


# ---------------------------------------------------------------- disk + halo
def workload_diskhalo(args, env):
    from bhut_amd import autotune, workloads
    nat, torch = env.nat, env.torch
    eps, G = args.softening, 1.0
    nd = (args.n or 16_000_000) // 2
    log(f"bench: generating disk + halo, {nd} + {nd} particles")
    x, m = workloads.disk_halo(nd, nd)
    N = len(x)
    pos_t, mass_t = env.upload(x), env.upload(m)
    out_t = torch.empty((N, 3), dtype=torch.float64, device=env.dev)
    env.sync()
    rng = np.random.default_rng(9)
    sel = np.sort(rng.choice(N, 10_000, replace=False))
    log("bench: FP64 direct sum on the 10,000-target subset")
    ref = nat.direct(x, m, x[sel], softening=eps, G=G)
    rows = []
    for theta in (0.3, 0.5, 0.7):
        log(f"bench: theta={theta}: autotuning leaf / group size / stack depth")
        best, table = autotune.tune_leaf_f32_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, out_t.data_ptr(),
                                                 theta, eps, G, stack_caps=(1024,), multipole=args.multipole,
                                                 context=env.ctx)
        opts = f32_options(args, best)
        leaf = best["leaf_size"]

        def step():
            t = nat.DeviceTree.build_dev(pos_t.data_ptr(), mass_t.data_ptr(), N, 3, leaf, context=env.ctx)
            t.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, opts)
            return t

        ms, wall, prof = env.timed(step, max(1, args.steps // 3), 1)
        t = step()
        st = t.accel_f32_self_dev(out_t.data_ptr(), theta, eps, G, opts)
        env.sync()
        acc = rel_err(out_t.cpu().numpy()[sel], ref)
        row = {"theta": theta, "chosen": best, "ms": ms, "particle_accels_per_s": N / (ms * 1e-3),
               "interactions_per_s": float(st[0] + st[1]) / (ms * 1e-3), "accuracy": acc,
               "roofline": roofline_f32(prof, max(1, args.steps // 3), st), "autotune": table,
               "kernels_ms_per_step": {k: v[1] / max(1, args.steps // 3) for k, v in prof.items()}}
        log(f"bench: theta={theta} {best} {ms:.1f} ms/step; {acc}")
        rows.append(row)
        del t
    head = rows[1]
    line = base_line(args, head["particle_accels_per_s"], head["ms"], "fp32",
                     {"workload": "disk_halo_self_accelerations", "n": int(N), "n_disk": nd, "n_halo": nd,
                      "softening": eps, "G": G, "seed": 3, "theta": [0.3, 0.5, 0.7],
                      "step": "tree build + FP32 self accelerations (autotuned per theta)"})
    line["metric"] = "particle-accels/s, 16M disk+halo (value at θ=0.5)"
    line["roofline"] = head["roofline"]
    line["thetas"] = rows
    line["accuracy_sample"] = "10,000 random particles vs FP64 direct sum over all 16M"
    line["cpu_baseline"] = None
    return line


# ---------------------------------------------------------------- leapfrog
def workload_leapfrog(args, env):
    from bhut_amd import workloads
    from bhut_amd.integrate import Leapfrog
    torch = env.torch
    N = args.n or 4_000_000
    theta, eps, G, leaf = args.theta, args.softening, 1.0, args.leaf
    dt, steps, rebuild = 1e-3, args.lf_steps, 10
    log(f"bench: generating Plummer N={N} with Gaussian velocities sigma=0.3")
    x, m = workloads.plummer(N, seed=1)
    v = workloads.plummer_velocities(N, seed=2, sigma=0.3)
    res = {}
    e0 = None
    finals = {}
    for prec in ("fp32", "fp64"):
        lf = Leapfrog(x, v, m, theta=theta, softening=eps, G=G, leaf_size=leaf, precision=prec)
        if e0 is None:
            log("bench: initial energy (FP64 direct potential)")
            t0 = time.perf_counter()
            e0 = lf.energy(chunk=1 << 20, progress=lambda k: log(f"bench:   potential {k}/{N} targets, "
                                                                   f"{time.perf_counter() - t0:.1f} s"))
            log(f"bench: E0 = {e0:.12e} ({time.perf_counter() - t0:.1f} s)")
        env.sync()
        e_a = torch.cuda.Event(enable_timing=True)
        e_b = torch.cuda.Event(enable_timing=True)
        done = 0
        total_ms = 0.0
        while done < steps:
            k = min(10, steps - done)
            e_a.record(env.stream)
            lf.run(k, dt, rebuild_every=rebuild, use_graph=not args.no_graph)
            e_b.record(env.stream)
            torch.cuda.synchronize()
            total_ms += e_a.elapsed_time(e_b)
            done += k
            log(f"bench: {prec} {done}/{steps} steps, {total_ms / done:.2f} ms/step")
        log(f"bench: {prec} final energy")
        t0 = time.perf_counter()
        e1 = lf.energy(chunk=1 << 20, progress=lambda k: log(f"bench:   potential {k}/{N} targets, "
                                                               f"{time.perf_counter() - t0:.1f} s"))
        finals[prec] = lf.pos.cpu().numpy()
        res[prec] = {"ms_per_step": total_ms / steps, "particle_steps_per_s": N * steps / (total_ms * 1e-3),
                     "energy_final": e1, "rel_energy_drift": (e1 - e0) / abs(e0), "sim": lf.info()}
        log(f"bench: {prec} {res[prec]}")
        del lf
    dev = np.linalg.norm(finals["fp32"] - finals["fp64"], axis=1)
    disp = np.linalg.norm(finals["fp64"] - x, axis=1)
    head = res["fp32"]
    line = base_line(args, N / (head["ms_per_step"] * 1e-3), head["ms_per_step"], "fp32",
                     {"workload": "plummer_leapfrog", "n": N, "steps": steps, "dt": dt, "rebuild_every": rebuild,
                      "theta": theta, "softening": eps, "leaf_size": leaf, "G": G, "seed": 1,
                      "velocities": "isotropic Gaussian sigma 0.3, seed 2",
                      "step": "KDK: kick-drift, refit (or rebuild), forces, kick; HIP graph per step kind"})
    line["steps"] = steps
    line["warmup"] = 0
    line["metric"] = "particle-steps/s, 4M Plummer KDK leapfrog (refit each step, rebuild every 10)"
    line["unit"] = "particle-steps/s"
    line["forms"] = res
    line["energy_initial"] = e0
    line["final_position_rms_deviation_fp32_vs_fp64"] = float(np.sqrt(np.mean(dev ** 2)))
    line["final_position_rms_displacement"] = float(np.sqrt(np.mean(disp ** 2)))
    line["roofline"] = None
    line["cpu_baseline"] = None
    return line


# ---------------------------------------------------------------- streamed targets
def workload_stream(args, env):
    """1M Plummer tree; 16M host-resident targets uniform in [-5, 5]^3
    through the device in chunks on two streams, against one call on all
    targets (host arrays) and against the device-resident walk alone."""
    from bhut_amd import workloads
    nat, torch = env.nat, env.torch
    N = 1_000_000
    M = args.n or 16_000_000
    theta, eps, G, leaf = args.theta, args.softening, 1.0, args.leaf
    x, m = workloads.plummer(N, seed=1)
    rng = np.random.default_rng(12)
    q = rng.uniform(-5.0, 5.0, (M, 3))
    t = nat.DeviceTree.build(x, m, leaf, context=env.ctx)
    q_t = env.upload(q)
    o_t = torch.empty((M, 3), dtype=torch.float64, device=env.dev)
    env.sync()
    opts = nat.F32Opts(64, 1, 1024, 2)
    res = {}
    for form in ("fp32", "fp64"):
        reps = 3 if form == "fp32" else 1
        def dev_call():
            if form == "fp32":
                t.accel_f32_targets_dev(q_t.data_ptr(), M, o_t.data_ptr(), theta, eps, G, opts)
            else:
                t.accel_targets_dev(q_t.data_ptr(), M, o_t.data_ptr(), theta, eps, G)
        dms, _, _ = env.timed(dev_call, reps, 1)
        one = t.accel_f32(q, theta, eps, G, opts) if form == "fp32" else t.accel_targets(q, theta, eps, G)
        t0 = time.perf_counter()
        for _ in range(reps):
            one = t.accel_f32(q, theta, eps, G, opts) if form == "fp32" else t.accel_targets(q, theta, eps, G)
        one_ms = (time.perf_counter() - t0) * 1e3 / reps
        row = {"device_resident_ms": dms, "one_call_host_ms": one_ms, "chunks": {}}
        for chunk in (1 << 20, 1 << 22):
            got = t.accel_targets_streamed(q, theta, eps, G, precision=form, chunk=chunk, opts=opts)
            t0 = time.perf_counter()
            for _ in range(reps):
                got = t.accel_targets_streamed(q, theta, eps, G, precision=form, chunk=chunk, opts=opts)
            sms = (time.perf_counter() - t0) * 1e3 / reps
            same = bool(np.array_equal(got.view(np.uint64), one.view(np.uint64)))
            row["chunks"][str(chunk)] = {"ms": sms, "targets_per_s": M / (sms * 1e-3),
                                         "bit_identical_to_one_call": same,
                                         "device_bytes_of_targets": 2 * chunk * 48}
            log(f"bench: stream {form} chunk {chunk}: {sms:.1f} ms (one call {one_ms:.1f} ms, device-resident "
                f"{dms:.1f} ms); identical {same}")
        res[form] = row
    head = res["fp32"]["chunks"][str(1 << 20)]
    line = base_line(args, M / (head["ms"] * 1e-3), head["ms"], "fp32",
                     {"workload": "streamed_targets", "n_sources": N, "targets": M,
                      "targets_distribution": "uniform [-5,5]^3, seed 12", "theta": theta, "softening": eps,
                      "leaf_size": leaf, "G": G, "chunk": 1 << 20,
                      "step": "host targets -> device in chunks on two streams -> host accelerations"})
    line["metric"] = "target-accels/s, 16M host targets streamed against a 1M Plummer tree"
    line["unit"] = "target-accels/s"
    line["forms"] = res
    line["roofline"] = None
    line["cpu_baseline"] = None
    return line


def workload_copy(args, env):
    torch = env.torch
    nbytes = 256 << 20
    a = torch.full((nbytes // 4,), 1.0, dtype=torch.float32, device=env.dev)
    b = torch.empty_like(a)
    env.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    env.sync()
    ms = (time.perf_counter() - t0) * 1e3 / 3
    return {"metric": "device-to-device copy GB/s (counter calibration)", "value": 2 * nbytes / (ms * 1e-3) / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 0, "ms_per_step": ms, "higher_is_better": True,
            "config": {"workload": "copy", "bytes": nbytes}, "roofline": None, "cpu_baseline": None}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is measured on one GPU's "
              "share; sharding targets over several GPUs is a design, not measured here", file=sys.stderr)
        sys.exit(2)
    env = Env()
    if args.list_budget_gib > 0:
        env.ctx.set_list_budget(int(args.list_budget_gib * 2 ** 30))
    fn = {"plummer": workload_plummer, "ensemble": workload_ensemble, "diskhalo": workload_diskhalo,
          "leapfrog": workload_leapfrog, "stream": workload_stream, "copy": workload_copy}[args.workload]
    line = fn(args, env)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
