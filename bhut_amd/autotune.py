"""Autotuner for the FP32 grouped form.

Times the group walk and evaluation kernels (HIP events recorded by the
library on its stream) for each candidate configuration on the caller's own
tree and input, and returns the fastest.  Each stack candidate runs on a
fresh clone of the tree: the library grows a walk's LDS stack when it
overflows and later walks on the same tree start at the grown size, so the
table reports the stack each candidate actually ran with
(``stack_used``).  Leaf size changes the tree itself, so `tune_leaf_f32`
rebuilds per candidate and also times the build.
"""

from __future__ import annotations

from typing import Iterable

import numpy as np

from . import _native

_F32_KERNELS = ("k_group_walk_count", "k_group_walk_fill", "k_group_eval_f32", "k_item_reduce", "f32_prepare")


def _timed(ctx, fn, reps: int) -> float:
    fn()  # warm-up (also builds FP32 source copies and moments)
    ctx.profile(True)
    for _ in range(reps):
        fn()
    prof = ctx.profile_summary()
    ctx.profile(False)
    return sum(ms for name, (_, ms) in prof.items() if name in _F32_KERNELS or name.startswith("build_")) / reps


def tune_f32(tree: "_native.DeviceTree", theta: float, softening: float, G: float = 1.0, *, targets=None,
             group_sizes: Iterable[int] = (64, 128), stack_caps: Iterable[int] = (512, 1024), multipole: int = 2,
             kahan: bool = True, reps: int = 2, out_ptr: int | None = None):
    """Best (group_size, stack_cap) for this tree: returns (best, table).
    With out_ptr (device buffer of n x dim float64) self accelerations stay
    on the device."""
    ctx = tree.context
    table = []
    for gs in group_sizes:
        for sc in stack_caps:
            opts = _native.F32Opts(int(gs), int(kahan), int(sc), int(multipole))
            cand = tree.clone()  # starts with no grown stack of its own
            if out_ptr is not None and targets is None:
                fn = lambda: cand.accel_f32_self_dev(out_ptr, theta, softening, G, opts)
            else:
                fn = lambda: cand.accel_f32(targets, theta, softening, G, opts)
            try:
                ms = _timed(ctx, fn, reps)
            except RuntimeError as e:
                if "stack" not in str(e):
                    raise
                table.append({"group_size": gs, "stack_cap": sc, "ms": None, "note": "stack overflow"})
                continue
            grown = cand.f32_stack()[0 if targets is None else 1]
            table.append({"group_size": gs, "stack_cap": sc, "stack_used": max(int(sc), grown), "ms": ms})
            del cand
    ok = [r for r in table if r["ms"] is not None]
    if not ok:
        raise RuntimeError("no FP32 configuration fits; raise the stack candidates")
    best = min(ok, key=lambda r: r["ms"])
    return {"group_size": best["group_size"], "stack_cap": best["stack_used"]}, table


def tune_leaf_f32(positions, masses, theta: float, softening: float, G: float = 1.0, *,
                  leaf_sizes: Iterable[int] = (16, 32, 64), group_sizes: Iterable[int] = (64, 128),
                  stack_caps: Iterable[int] = (1024,), multipole: int = 2, reps: int = 1, context=None):
    """Best (leaf_size, group_size, stack_cap) including the tree build."""
    ctx = context or _native.ctx()
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    m = np.ascontiguousarray(masses, dtype=np.float64)
    table = []
    for leaf in leaf_sizes:
        ctx.profile(True)
        tree = _native.DeviceTree.build(pos, m, int(leaf), context=ctx)
        prof = ctx.profile_summary()
        ctx.profile(False)
        build_ms = sum(ms for name, (_, ms) in prof.items() if name.startswith("build_"))
        best, sub = tune_f32(tree, theta, softening, G, group_sizes=group_sizes, stack_caps=stack_caps,
                             multipole=multipole, reps=reps)
        for r in sub:
            table.append(dict(r, leaf_size=leaf, build_ms=build_ms,
                              total_ms=None if r["ms"] is None else r["ms"] + build_ms))
        del tree
    ok = [r for r in table if r["total_ms"] is not None]
    best = min(ok, key=lambda r: r["total_ms"])
    return {"leaf_size": best["leaf_size"], "group_size": best["group_size"], "stack_cap": best["stack_cap"]}, table


def tune_leaf_f32_dev(pos_ptr: int, mass_ptr: int, n: int, dim: int, out_ptr: int, theta: float,
                      softening: float, G: float = 1.0, *, leaf_sizes: Iterable[int] = (16, 32, 64),
                      group_sizes: Iterable[int] = (64, 128), stack_caps: Iterable[int] = (1024,),
                      multipole: int = 2, reps: int = 1, context=None):
    """tune_leaf_f32 on inputs already in device memory."""
    ctx = context or _native.ctx()
    table = []
    for leaf in leaf_sizes:
        tree = _native.DeviceTree.build_dev(pos_ptr, mass_ptr, n, dim, int(leaf), context=ctx)
        ctx.profile(True)
        tree = _native.DeviceTree.build_dev(pos_ptr, mass_ptr, n, dim, int(leaf), context=ctx)
        prof = ctx.profile_summary()
        ctx.profile(False)
        build_ms = sum(ms for name, (_, ms) in prof.items() if name.startswith("build_"))
        best, sub = tune_f32(tree, theta, softening, G, group_sizes=group_sizes, stack_caps=stack_caps,
                             multipole=multipole, reps=reps, out_ptr=out_ptr)
        for r in sub:
            table.append(dict(r, leaf_size=leaf, build_ms=build_ms,
                              total_ms=None if r["ms"] is None else r["ms"] + build_ms))
        del tree
    ok = [r for r in table if r["total_ms"] is not None]
    best = min(ok, key=lambda r: r["total_ms"])
    return {"leaf_size": best["leaf_size"], "group_size": best["group_size"], "stack_cap": best["stack_cap"]}, table
