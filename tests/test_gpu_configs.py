"""The measured configurations at their own sizes.

* 16M disk + halo (8M exponential disk + 8M Hernquist halo, halo:disk mass
  5:1): for every theta in {0.3, 0.5, 0.7} and leaf size in {16, 32, 64} the
  permutation equals the oracle's stable Morton argsort, the topology the
  closed form, the node arrays the oracle's tree bit for bit; the FP32 form
  against the FP64 direct sum over all 16M sources on a 10 000-particle
  random subset (rms and max relative error printed, bounds in the test);
  the plain walk against the oracle on a subset at 1e-12 and its lists.
* 512 independent Plummer spheres (N log-uniform in [2 000, 200 000]) in one
  forest: 16 sampled systems (the smallest and the largest among them) bit
  for bit against lone trees and the oracle (tree, lists, FP64 and FP32
  accelerations).
* 4M Plummer leapfrog (sigma_v 0.3, dt 1e-3): 10 steps with a rebuild every
  5, graph replays against eager steps bit for bit; 100 FP32 steps with a
  rebuild every 10, relative energy drift bounded (FP64 direct-sum
  potential).
* 1M Plummer with a 250 000-target query uniform in [-5, 5]^3: plain walk
  against the oracle at 1e-12 on a sample (lists too), FP32 form against the
  FP64 direct sum.
"""

import time

import numpy as np
import pytest

from bhut_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TREE_FIELDS = ["center", "half_size", "mass", "com", "first_child", "child_count", "start", "count", "is_leaf"]


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b, name=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, name
    if a.dtype == np.float64:
        np.testing.assert_array_equal(a.view(np.uint64), np.asarray(b, np.float64).view(np.uint64), err_msg=name)
    else:
        np.testing.assert_array_equal(a, b, err_msg=name)


def acc_close(got, ref, tol=1e-12):
    nr = np.linalg.norm(ref, axis=1)
    err = np.linalg.norm(got - ref, axis=1)
    bad = err > np.maximum(tol * nr, tol * nr.max())
    assert not bad.any(), f"{bad.sum()} targets off; worst rel {np.max(err / np.maximum(nr, 1e-300)):.3e}"


def rel_err(got, ref):
    e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


# rms and max relative error bounds of the FP32 form against FP64 direct
# summation, per theta (the measured values sit well inside them)
F32_BOUNDS = {0.3: (2e-4, 2e-3), 0.5: (1e-3, 1e-2), 0.7: (4e-3, 4e-2)}


def test_config_diskhalo_16m(native):
    t0 = time.time()
    x, m = workloads.disk_halo()
    n = len(x)
    assert n == 16_000_000
    say(f"diskhalo16m: generated in {time.time() - t0:.1f} s")
    perm_ref = O.argsort_stable(O.morton_codes(x))
    rng = np.random.default_rng(33)
    sub = np.sort(rng.choice(n, 10_000, replace=False))
    ref = native.direct(x, m, x[sub], softening=0.01)
    say(f"diskhalo16m: oracle order and direct sum on 10k targets done at {time.time() - t0:.1f} s")
    for leaf in (16, 32, 64):
        dt = native.DeviceTree.build(x, m, leaf)
        e = dt.export()
        np.testing.assert_array_equal(e["perm"], perm_ref)
        top = O.topology(n, leaf)
        for f in ("start", "count", "first_child", "child_count"):
            np.testing.assert_array_equal(e[f], getattr(top, f), err_msg=f"leaf {leaf} {f}")
        assert len(e["mass"]) == 2 * int(e["is_leaf"].sum()) - 1
        ot = O.build(x, m, leaf)
        for f in TREE_FIELDS:
            bits_equal(e[f], getattr(ot, f), f"leaf {leaf} {f}")
        del e
        for theta in (0.3, 0.5, 0.7):
            acc = dt.accel_f32(None, theta, 0.01, 1.0)
            rms, mx = rel_err(acc[sub], ref)
            say(f"diskhalo16m: leaf {leaf} theta {theta}: FP32 rms {rms:.3e} max {mx:.3e}")
            assert rms <= F32_BOUNDS[theta][0] and mx <= F32_BOUNDS[theta][1], (leaf, theta, rms, mx)
        # plain walk: 1e-12 against the oracle's self walk on a subset, lists exact
        acc = dt.accel_self(0.5, 0.01, 1.0)
        acc_close(acc[sub[:500]], O.accelerations_subset(ot, x, m, sub[:500], softening=0.01, theta=0.5))
        ptr, ids = dt.interaction_lists(x[sub[:8]], 0.5)
        for k in range(8):
            np.testing.assert_array_equal(ids[ptr[k]:ptr[k + 1]], O.target_list(ot, x[sub[k]], 0.5))
        say(f"diskhalo16m: leaf {leaf} done at {time.time() - t0:.1f} s")
        del dt, ot


def test_config_ensemble_512(native):
    from bhut_amd.batched import Forest
    t0 = time.time()
    systems = workloads.ensemble()
    counts = np.array([len(s[0]) for s in systems])
    assert len(systems) == 512 and counts.min() >= 2000 and counts.max() <= 200_000
    say(f"ensemble512: {counts.sum()} particles generated in {time.time() - t0:.1f} s")
    f = Forest.from_systems(systems, leaf_size=32)
    a64 = f.split(f.accelerations(theta=0.5, softening=0.01, precision="fp64"))
    a32 = f.split(f.accelerations(theta=0.5, softening=0.01, precision="fp32"))
    po, no = f.device_tree.systems()
    e = f.device_tree.export()
    say(f"ensemble512: forest built and walked at {time.time() - t0:.1f} s")
    rng = np.random.default_rng(512)
    sample = sorted(set(rng.choice(512, 14, replace=False)) | {int(np.argmin(counts)), int(np.argmax(counts))})
    for s in sample:
        x, m = systems[s]
        lone = native.DeviceTree.build(x, m, 32)
        le = lone.export()
        a, b = no[s], no[s + 1]
        for fld in ("center", "half_size", "mass", "com", "count", "is_leaf", "child_count"):
            bits_equal(e[fld][a:b], le[fld], f"system {s} {fld}")
        fc = e["first_child"][a:b]
        np.testing.assert_array_equal(np.where(fc >= 0, fc - a, fc), le["first_child"])
        np.testing.assert_array_equal(e["start"][a:b] - po[s], le["start"])
        np.testing.assert_array_equal(e["perm"][po[s]:po[s + 1]] - po[s], le["perm"])
        ot = O.build(x, m, 32)
        for fld in TREE_FIELDS + ["perm"]:
            bits_equal(le[fld], getattr(ot, fld), f"system {s} {fld} vs oracle")
        bits_equal(a64[s], lone.accel_self(0.5, 0.01, 1.0), f"system {s} FP64")
        bits_equal(a32[s], lone.accel_f32(None, 0.5, 0.01, 1.0), f"system {s} FP32")
        sel = rng.choice(len(x), min(200, len(x)), replace=False)
        acc_close(a64[s][sel], O.accelerations_subset(ot, x, m, sel, softening=0.01, theta=0.5))
        # lists at the particles' own positions (the visited nodes do not depend
        # on the self index; only the leaf sums do)
        ptr, ids = lone.interaction_lists(x[sel[:4]], 0.5)
        for k, i in enumerate(sel[:4]):
            np.testing.assert_array_equal(ids[ptr[k]:ptr[k + 1]], O.target_list(ot, x[i], 0.5))
        rms, mx = rel_err(a32[s][sel], native.direct(x, m, x[sel], softening=0.01))
        assert rms <= 1e-3 and mx <= 1e-2, (s, rms, mx)
    say(f"ensemble512: {len(sample)} systems checked at {time.time() - t0:.1f} s")


def test_config_leapfrog_4m(native):
    import torch
    from bhut_amd.integrate import Leapfrog
    n = 4_000_000
    x, m = workloads.plummer(n, seed=4)
    v = workloads.plummer_velocities(n, seed=4, sigma=0.3)
    t0 = time.time()
    # graph replays against eager steps: 10 steps, rebuilds at steps 5 and 10
    for prec in ("fp32", "fp64"):
        runs = {}
        for graph in (True, False):
            lf = Leapfrog(x, v, m, theta=0.5, softening=0.01, precision=prec)
            lf.run(10, 1e-3, rebuild_every=5, use_graph=graph)
            torch.cuda.synchronize()
            runs[graph] = (lf.pos.cpu().numpy(), lf.vel.cpu().numpy(), lf.info())
            del lf
        bits_equal(runs[True][0], runs[False][0], f"{prec} positions")
        bits_equal(runs[True][1], runs[False][1], f"{prec} velocities")
        info = runs[True][2]
        assert info["steps"] == 10 and info["rebuilds"] == 2 and info["graph_launches"] >= 6, info
        say(f"leapfrog4m: {prec} graph == eager over 10 steps ({info}) at {time.time() - t0:.1f} s")
    # 100 FP32 steps, rebuild every 10: energy drift (FP64 direct-sum potential)
    lf = Leapfrog(x, v, m, theta=0.5, softening=0.01, precision="fp32")
    e0 = lf.energy(chunk=1_000_000, progress=lambda k: say(f"leapfrog4m: E0 potential {k}/{n}"))
    lf.run(100, 1e-3, rebuild_every=10)
    torch.cuda.synchronize()
    info = lf.info()
    e1 = lf.energy(chunk=1_000_000, progress=lambda k: say(f"leapfrog4m: E1 potential {k}/{n}"))
    drift = (e1 - e0) / abs(e0)
    say(f"leapfrog4m: 100 FP32 steps, relative energy drift {drift:.3e} ({info})")
    assert info["steps"] == 100 and info["rebuilds"] == 10
    assert abs(drift) < 1e-6
    xs = lf.pos.cpu().numpy()
    assert np.isfinite(xs).all() and np.sqrt(np.mean(np.sum((xs - x) ** 2, axis=1))) > 1e-3  # the system moved


def test_config_plummer_1m_query_250k(native):
    n = 1_000_000
    x, m = workloads.plummer(n, seed=1)
    rng = np.random.default_rng(11)
    q = rng.uniform(-5.0, 5.0, (250_000, 3))
    dt = native.DeviceTree.build(x, m, 32)
    ot = O.build(x, m, 32)
    acc = dt.accel_targets(q, 0.5, 0.01, 1.0)
    sel = np.sort(rng.choice(len(q), 2000, replace=False))
    acc_close(acc[sel], O.accelerations(ot, x, m, q[sel], softening=0.01, theta=0.5))
    ptr, ids = dt.interaction_lists(q[sel[:16]], 0.5)
    for k in range(16):
        np.testing.assert_array_equal(ids[ptr[k]:ptr[k + 1]], O.target_list(ot, q[sel[k]], 0.5))
    a32 = dt.accel_f32(q, 0.5, 0.01, 1.0)
    sel = np.sort(rng.choice(len(q), 4096, replace=False))
    rms, mx = rel_err(a32[sel], native.direct(x, m, q[sel], softening=0.01))
    say(f"query250k: FP32 rms {rms:.3e} max {mx:.3e}")
    assert rms <= 1e-3 and mx <= 1e-2
