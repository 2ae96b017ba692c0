"""Entry points the other GPU tests do not call: traversal statistics and
the refit-or-rebuild rule through the tree functions, the device-pointer
forms of refit, potential and direct sum against their host forms, the
leapfrog driver's energies, the leaf-size autotuner, forest bookkeeping and
context trimming."""

import numpy as np
import pytest

from bhut_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_statistics_and_refit_rule(native):
    """compute_tree_statistics (bh.py:604-700) equal to the oracle's
    restatement of the reference's values, at the tree's positions, at
    displaced positions of the same count (the walk runs at the passed
    positions) and on a host-only TreeData; traversal_counters in self mode
    equal to the oracle's exact counts; should_refit_vs_rebuild
    (refit.py:195-251) agrees with the oracle on a small drift, a far
    particle and a changed N."""
    from bhut_amd.traverse.bh import compute_tree_statistics, traversal_counters
    from bhut_amd.tree.build import build_tree
    from bhut_amd.tree.node import TreeData
    from bhut_amd.tree.refit import should_refit_vs_rebuild
    x, m = workloads.plummer(5000, seed=21)
    tree = build_tree(x, m, dim=3, leaf_size=16)
    ot = O.build(x, m, 16)
    host = TreeData(**tree.arrays(), dim=3, leaf_size=16)
    xd = x + 0.05 * np.random.default_rng(21).standard_normal(x.shape)
    for tr in (tree, host):
        for p in (x, xd, x[:700]):
            st = compute_tree_statistics(tr, p, 0.5)
            exp = O.statistics(ot, p, 0.5)
            for k, v in exp.items():
                assert st[k] == v, k
            np.testing.assert_array_equal(st["per_target"], O.counters(ot, p, self_mode=False, theta=0.5))
    np.testing.assert_array_equal(traversal_counters(tree, x, 0.5, self_mode=True),
                                  O.counters(ot, x, self_mode=True, theta=0.5))
    np.testing.assert_array_equal(traversal_counters(host, x, 0.5, self_mode=True),
                                  O.counters(ot, x, self_mode=True, theta=0.5))
    with pytest.raises(ValueError):
        compute_tree_statistics(tree, x[:, :2], 0.5)
    x2 = x + 1e-4
    assert should_refit_vs_rebuild(tree, x2) and O.should_refit(ot, x2)
    x3 = x.copy()
    x3[0] = 1e4
    assert not should_refit_vs_rebuild(tree, x3) and not O.should_refit(ot, x3)
    assert not should_refit_vs_rebuild(tree, x[:100])


def test_device_pointer_forms_equal_host_forms(native):
    """refit_dev, potential_dev and direct_dev on torch tensors give the
    bits of refit, potential and direct on the same host arrays."""
    import torch
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(23)
    x, m = workloads.plummer(20_000, seed=23)
    x2 = x + 1e-3 * rng.standard_normal(x.shape)
    c = native.ctx()
    a = native.DeviceTree.build(x, m, 32)
    a.refit(x2, m)
    b = native.DeviceTree.build(x, m, 32)
    x2t = torch.tensor(x2, device=dev)
    mt = torch.tensor(m, device=dev)
    torch.cuda.synchronize(dev)
    b.refit_dev(x2t.data_ptr(), mt.data_ptr())
    c.synchronize()
    ea, eb = a.export(), b.export()
    for f in ("mass", "com", "center", "half_size"):
        np.testing.assert_array_equal(bits(ea[f]), bits(eb[f]), err_msg=f)
    ph = a.potential(None, 0.5, 0.01, 1.0)
    pd = torch.empty(len(x), dtype=torch.float64, device=dev)
    a.potential_dev(None, len(x), pd.data_ptr(), 0.5, 0.01, 1.0)
    c.synchronize()
    np.testing.assert_array_equal(bits(pd.cpu().numpy()), bits(ph))
    n = 3000
    acc_h, pot_h = native.direct(x2[:n], m[:n], softening=0.01, potential=True)
    acc_d = torch.empty((n, 3), dtype=torch.float64, device=dev)
    pot_d = torch.empty(n, dtype=torch.float64, device=dev)
    xs = x2t[:n].contiguous()
    torch.cuda.synchronize(dev)
    native.direct_dev(xs.data_ptr(), mt.data_ptr(), n, xs.data_ptr(), n, 3, True, 0.01, 1.0, acc_d.data_ptr(),
                      pot_d.data_ptr())
    c.synchronize()
    np.testing.assert_array_equal(bits(acc_d.cpu().numpy()), bits(acc_h))
    np.testing.assert_array_equal(bits(pot_d.cpu().numpy()), bits(pot_h))


def test_leapfrog_energies(native):
    """Kinetic energy and the FP64 direct-sum potential energy of the
    leapfrog driver (whole and in chunks) against NumPy sums."""
    from bhut_amd.integrate import Leapfrog
    n = 2000
    x, m = workloads.plummer(n, seed=24)
    v = workloads.plummer_velocities(n, seed=25)
    lf = Leapfrog(x, v, m, softening=0.01)
    ke = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    assert abs(lf.kinetic_energy() - ke) <= 1e-12 * ke
    d = x[:, None, :] - x[None, :, :]
    inv = 1.0 / np.sqrt(np.sum(d * d, axis=2) + 0.01 ** 2)
    np.fill_diagonal(inv, 0.0)
    w = -0.5 * np.sum(m[:, None] * m[None, :] * inv)
    for chunk in (0, 700):
        pe = lf.potential_energy(chunk=chunk)
        assert abs(pe - w) <= 1e-11 * abs(w), (chunk, pe, w)


def test_autotuner_and_forest_bookkeeping(native):
    """tune_leaf_f32 times every (leaf, group) pair and returns the fastest;
    a forest reports its systems; a trimmed context keeps working."""
    from bhut_amd.autotune import tune_leaf_f32
    from bhut_amd.batched import Forest
    x, m = workloads.plummer(30_000, seed=26)
    best, table = tune_leaf_f32(x, m, 0.5, 0.01, leaf_sizes=(16, 32), group_sizes=(64, 128))
    assert len(table) == 4
    ok = [r for r in table if r["total_ms"] is not None]
    assert min(ok, key=lambda r: r["total_ms"])["leaf_size"] == best["leaf_size"]
    assert best["leaf_size"] in (16, 32) and best["group_size"] in (64, 128)
    systems = [workloads.plummer(k, seed=30 + k) for k in (100, 2000, 5)]
    f = Forest.from_systems(systems, leaf_size=16)
    assert f.num_systems == 3
    native.ctx().trim()
    a = native.DeviceTree.build(x, m, 32).accel_self(0.5, 0.01, 1.0)
    assert np.all(np.isfinite(a))


def test_misshaped_inputs_raise_before_any_device_call(native):
    """Targets, masses and positions whose shapes do not match the tree
    raise ValueError on the host (the C side reads rows * dim values, so a
    wrong shape must never reach it), as the reference's broadcasting fails
    with ValueError."""
    import bhut_amd
    from bhut_amd.batched import Forest
    from bhut_amd.traverse.bh import barnes_hut_accelerations_targets
    from bhut_amd.tree.build import build_tree
    from bhut_amd.tree.node import TreeData
    x, m = workloads.plummer(3000, seed=31)
    tree = build_tree(x, m, dim=3, leaf_size=16)
    dt = tree.device_tree
    bad_targets = (np.zeros((10, 2)), np.zeros(30), np.zeros((10, 4)), np.zeros((2, 5, 3)))
    for tg in bad_targets:
        for call in (lambda: dt.accel_targets(tg, 0.5, 0.01, 1.0), lambda: dt.accel_f32(tg, 0.5, 0.01, 1.0),
                     lambda: dt.potential(tg, 0.5, 0.01, 1.0), lambda: dt.counters(tg, 0.5),
                     lambda: dt.interaction_lists(tg, 0.5),
                     lambda: barnes_hut_accelerations_targets(tree, x, m, tg, 0.01, 0.5)):
            with pytest.raises(ValueError):
                call()
    with pytest.raises(ValueError):
        dt.refit(x[:100], m[:100])
    with pytest.raises(ValueError):
        dt.refit(x, m[:100])
    with pytest.raises(ValueError):
        dt.refit(x[:, :2], m)
    host = TreeData(**tree.arrays(), dim=3, leaf_size=16)
    with pytest.raises(ValueError):
        native.DeviceTree.from_arrays(host, x[:100], m[:100])
    with pytest.raises(ValueError):
        native.DeviceTree.from_arrays(host, x, m[:10])
    f = Forest(x, m, [1000, 2000], leaf_size=16)
    with pytest.raises(ValueError):
        f.refit(x, m[:5])
    t = bhut_amd.Tree(x, m, leaf_size=16).build()
    with pytest.raises(ValueError):
        t.accelerations(np.zeros((5, 2)))
    # the tree still works after the refused calls
    np.testing.assert_array_equal(dt.accel_targets(x[:7], 0.5, 0.01, 1.0), dt.accel_targets(x[:7], 0.5, 0.01, 1.0))


def test_three_child_tree_plain_walk_and_fp32_refusal(native):
    """An imported TreeData whose root has three children (bhut's walk pushes
    child_count children, bh.py:529-535): the plain walk visits all three and
    matches the oracle at theta 0 and 1; the FP32 form, whose group walk
    pushes exactly two children, refuses the tree with ValueError."""
    from bhut_amd.traverse.bh import barnes_hut_accelerations
    from bhut_amd.tree.node import TreeData
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.5, 0.0], [3.0, -1.0, 2.0]])
    m = np.array([1.0, 2.0, 0.5])
    lo, hi = x.min(axis=0), x.max(axis=0)
    center = np.vstack([(lo + hi) / 2, x])
    half = np.array([np.max(hi - lo) / 2, 0.0, 0.0, 0.0])
    mass = np.array([m.sum(), *m])
    com = np.vstack([(m[:, None] * x).sum(axis=0) / m.sum(), x])
    td = TreeData(center=center, half_size=half, mass=mass, com=com, first_child=np.array([1, -1, -1, -1]),
                  child_count=np.array([3, 0, 0, 0]), start=np.array([0, 0, 1, 2]), count=np.array([3, 1, 1, 1]),
                  is_leaf=np.array([False, True, True, True]), perm=np.arange(3), dim=3, leaf_size=1)
    ot = O.OracleTree(center=center, half_size=half, mass=mass, com=com, first_child=np.array([1, -1, -1, -1]),
                      child_count=np.array([3, 0, 0, 0]), start=np.array([0, 0, 1, 2]),
                      count=np.array([3, 1, 1, 1]), is_leaf=np.array([False, True, True, True]),
                      perm=np.arange(3), dim=3, leaf_size=1)
    for theta in (0.0, 1.0, 10.0):
        got = barnes_hut_accelerations(td, x, m, 0.01, theta)
        np.testing.assert_allclose(got, O.accelerations(ot, x, m, softening=0.01, theta=theta), rtol=1e-12)
    with pytest.raises(ValueError, match="binary"):
        barnes_hut_accelerations(td, x, m, 0.01, 0.5, precision="fp32")
