"""Batched form: a forest of independent systems built and walked at once.

Each system of the forest must be exactly the tree build_tree gives it alone
(bhut/tree/build.py:19-107): every node field and the permutation, after
removing the system's node / particle offsets, bit-equal to the oracle's tree
of that system alone.  Self accelerations of the forest equal, bit for bit,
the plain walk of each system alone (the walk of a system never leaves it),
and the FP32 form groups each system's targets exactly as it does alone, so
its results are bit-equal too.
"""

import numpy as np
import pytest

from oracle import oracle as O
from bhut_amd.workloads import plummer

pytestmark = pytest.mark.gpu

NODE_FIELDS = ["center", "half_size", "mass", "com", "child_count", "count", "is_leaf"]


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def bits_equal(a, b, name=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, name
    if a.dtype == np.float64:
        np.testing.assert_array_equal(a.view(np.uint64), np.asarray(b, np.float64).view(np.uint64), err_msg=name)
    else:
        np.testing.assert_array_equal(a, b, err_msg=name)


def make_systems(sizes, dim=3, seed=0):
    out = []
    for i, n in enumerate(sizes):
        rng = np.random.default_rng(seed + i)
        if i % 3 == 0 and dim == 3:
            x, m = plummer(n, seed=seed + i, a=rng.uniform(0.5, 2.0))
        else:
            x = rng.standard_normal((n, dim)) * rng.uniform(0.1, 10.0) + rng.uniform(-50, 50, dim)
            m = rng.uniform(0.1, 2.0, n)
        if i % 4 == 1 and n > 4:
            x[n // 2:n // 2 + 3] = x[0]  # duplicate positions inside one system
        out.append((x, m))
    return out


def forest_slices(e, po, no, s):
    a, b = no[s], no[s + 1]
    sl = {f: e[f][a:b] for f in NODE_FIELDS}
    fc = e["first_child"][a:b]
    sl["first_child"] = np.where(fc >= 0, fc - a, fc)
    sl["start"] = e["start"][a:b] - po[s]
    sl["perm"] = e["perm"][po[s]:po[s + 1]] - po[s]
    return sl


# sizes below, at and just above the leaf size, odd sizes, a single particle
SIZES = [1, 2, 31, 32, 33, 64, 65, 500, 1023, 4097, 7, 2000]


@pytest.mark.parametrize("leaf", [8, 32])
def test_forest_topology_bit_exact(native, leaf):
    systems = make_systems(SIZES, seed=leaf)
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, [len(x) for x, _ in systems], leaf)
    po, no = f.systems()
    np.testing.assert_array_equal(po, np.concatenate([[0], np.cumsum(SIZES)]))
    e = f.export()
    for s, (x, m) in enumerate(systems):
        ot = O.build(x, m, leaf)
        sl = forest_slices(e, po, no, s)
        assert no[s + 1] - no[s] == len(ot.mass), f"system {s} node count"
        for k, v in sl.items():
            bits_equal(v, getattr(ot, k), f"system {s} {k}")


def test_forest_accel_fp64_equals_single(native):
    systems = make_systems(SIZES + [30_000, 3], seed=7)
    counts = [len(x) for x, _ in systems]
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, counts, 32)
    acc = f.accel_self(0.5, 0.01, 1.0)
    off = np.concatenate([[0], np.cumsum(counts)])
    for s, (x, m) in enumerate(systems):
        one = native.DeviceTree.build(x, m, 32).accel_self(0.5, 0.01, 1.0)
        bits_equal(acc[off[s]:off[s + 1]], one, f"system {s}")
        if len(x) <= 4097:
            ref = O.accelerations(O.build(x, m, 32), x, m, softening=0.01, theta=0.5)
            err = np.linalg.norm(acc[off[s]:off[s + 1]] - ref, axis=1)
            nr = np.linalg.norm(ref, axis=1)
            assert np.all(err <= 1e-12 * np.maximum(nr, nr.max() if len(nr) else 0.0) + 0.0), f"system {s}"


def test_forest_counters_and_lists(native):
    systems = make_systems([300, 40, 2048, 1], seed=3)
    counts = [len(x) for x, _ in systems]
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, counts, 16)
    po, no = f.systems()
    cnt = f.counters(None, 0.6)
    ptr, ids = f.interaction_lists(None, 0.6)
    for s, (x, m) in enumerate(systems):
        ot = O.build(x, m, 16)
        np.testing.assert_array_equal(cnt[po[s]:po[s + 1]], O.counters(ot, x, self_mode=True, theta=0.6))
        for i in range(0, len(x), max(1, len(x) // 17)):
            gi = po[s] + i
            got = ids[ptr[gi]:ptr[gi + 1]] - no[s]
            np.testing.assert_array_equal(got, O.target_list(ot, x[i], 0.6), err_msg=f"system {s} target {i}")


def test_forest_refit_bit_exact(native):
    systems = make_systems([5000, 77, 1200], seed=11)
    counts = [len(x) for x, _ in systems]
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, counts, 32)
    rng = np.random.default_rng(5)
    pos2 = pos + 0.01 * rng.standard_normal(pos.shape)
    mas2 = mas * rng.uniform(0.9, 1.1, len(mas))
    f.refit(pos2, mas2)
    po, no = f.systems()
    e = f.export()
    for s, (x, m) in enumerate(systems):
        ot = O.refit(O.build(x, m, 32), pos2[po[s]:po[s + 1]], mas2[po[s]:po[s + 1]])
        sl = forest_slices(e, po, no, s)
        for k in ("mass", "com", "center", "half_size"):
            bits_equal(sl[k], getattr(ot, k), f"system {s} {k}")


def test_forest_fp32_equals_single(native):
    systems = make_systems([20_000, 100, 64, 129, 3000], seed=21)
    counts = [len(x) for x, _ in systems]
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, counts, 32)
    off = np.concatenate([[0], np.cumsum(counts)])
    for gs in (64, 128):
        opts = native.F32Opts(gs, 1, 0, 2)
        acc, st = f.accel_f32(None, 0.5, 0.01, 1.0, opts, stats=True)
        groups = 0
        for s, (x, m) in enumerate(systems):
            one, st1 = native.DeviceTree.build(x, m, 32).accel_f32(None, 0.5, 0.01, 1.0, opts, stats=True)
            bits_equal(acc[off[s]:off[s + 1]], one, f"gs {gs} system {s}")
            groups += st1[2]
        assert st[2] == groups


def test_forest_two_dimensional(native):
    systems = make_systems([1000, 33, 4000], dim=2, seed=4)
    counts = [len(x) for x, _ in systems]
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, counts, 16)
    po, no = f.systems()
    e = f.export()
    acc = f.accel_self(0.5, 0.0, 1.0)
    for s, (x, m) in enumerate(systems):
        ot = O.build(x, m, 16)
        for k, v in forest_slices(e, po, no, s).items():
            bits_equal(v, getattr(ot, k), f"system {s} {k}")
        bits_equal(acc[po[s]:po[s + 1]], native.DeviceTree.build(x, m, 16).accel_self(0.5, 0.0, 1.0))


def test_forest_rejects_single_tree_operations(native):
    systems = make_systems([100, 200], seed=1)
    pos = np.concatenate([x for x, _ in systems])
    mas = np.concatenate([m for _, m in systems])
    f = native.DeviceTree.build_forest(pos, mas, [100, 200], 32)
    with pytest.raises(ValueError, match="forest"):
        f.accel_targets(np.zeros((4, 3)), 0.5, 0.0, 1.0)
    with pytest.raises(ValueError, match="forest"):
        f.accel_f32(np.zeros((4, 3)), 0.5, 0.0, 1.0)
    with pytest.raises(ValueError, match="forest"):
        f.should_refit(pos)
    with pytest.raises(ValueError):
        native.DeviceTree.build_forest(pos, mas, [100, 0, 200], 32)


def test_batched_api(native):
    import bhut_amd
    systems = make_systems([700, 50, 2500], seed=9)
    accs = bhut_amd.accelerations_batched(systems, theta=0.5, softening=0.01)
    for (x, m), a in zip(systems, accs):
        bits_equal(a, bhut_amd.accelerations(x, m, theta=0.5, softening=0.01))
    fr = bhut_amd.Forest.from_systems(systems, leaf_size=32)
    acc32 = fr.split(fr.accelerations(theta=0.5, softening=0.01, precision="fp32"))
    for (x, m), a in zip(systems, acc32):
        bits_equal(a, bhut_amd.accelerations(x, m, theta=0.5, softening=0.01, precision="fp32"))
