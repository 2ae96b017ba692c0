"""The FP32 grouped form against FP64 direct summation and the plain form.

Tolerances are stated per test: the FP32 form is not bit-faithful; at
theta = 0 (every node opened, bh.py:600-601) it is an FP32 direct sum whose
relative error is a few 1e-7, and at theta > 0 its conservative group test
makes it at least as accurate as the plain FP64 walk (rms relative error
against FP64 direct summation).
"""

import numpy as np
import pytest

from bhut_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def rel_err(a, ref):
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


def test_theta_zero_is_fp32_direct_sum(native):
    """Every pair in FP32 on coordinates relative to the group centre: rms
    relative error ~3e-7; the maximum (~2e-5) is at targets whose total
    acceleration nearly cancels (|a| ~ 0.016 against ~1 in the core)."""
    x, m = workloads.plummer(20_000, seed=3)
    dt = native.DeviceTree.build(x, m, 32)
    a = dt.accel_f32(None, 0.0, 0.01, 1.0)
    d = native.direct(x, m, softening=0.01)
    rms, mx = rel_err(a, d)
    assert rms < 2e-6 and mx < 5e-5, (rms, mx)


@pytest.mark.parametrize("kind", ["random", "two_masses"])
def test_unequal_masses_theta_zero(native, kind):
    """Source tiles whose masses differ multiply each pair by its mass; tiles
    of one common mass apply it to the tile's sum.  Random masses take the
    first path everywhere, a disk (mass 1/6) + halo (5/6) mix takes both;
    either way an FP32 direct sum: random masses bounded as for equal ones;
    the disk's in-plane forces cancel more (looser bounds, still orders of
    magnitude below a wrong mass factor)."""
    if kind == "random":
        x, _ = workloads.plummer(20_000, seed=4)
        m = np.random.default_rng(4).uniform(0.5, 1.5, 20_000) / 20_000
        bound = (2e-6, 5e-5)
    else:
        x, m = workloads.disk_halo(10_000, 10_000, seed=5)
        bound = (1e-5, 1e-3)
    dt = native.DeviceTree.build(x, m, 32)
    a = dt.accel_f32(None, 0.0, 0.01, 1.0)
    d = native.direct(x, m, softening=0.01)
    rms, mx = rel_err(a, d)
    assert rms < bound[0] and mx < bound[1], (kind, rms, mx)


@pytest.mark.parametrize("gs,kahan,multipole", [(64, 1, 2), (128, 1, 2), (64, 0, 2), (64, 1, 1)])
def test_theta_half_accuracy(native, gs, kahan, multipole):
    """rms relative error against FP64 direct summation: with quadrupole
    cells below 1e-3 at theta = 0.5; monopole-only no worse than the plain
    per-target walk (the group test only ever opens more)."""
    n = 100_000
    x, m = workloads.plummer(n, seed=1)
    dt = native.DeviceTree.build(x, m, 32)
    opts = native.F32Opts(gs, kahan, 0, multipole)
    a32, st = dt.accel_f32(None, 0.5, 0.01, 1.0, opts, stats=True)
    a64 = dt.accel_self(0.5, 0.01, 1.0)
    sel = np.random.default_rng(0).choice(n, 2000, replace=False)
    d = native.direct(x, m, x[sel], softening=0.01)
    rms32, _ = rel_err(a32[sel], d)
    rms64, _ = rel_err(a64[sel], d)
    assert rms32 <= rms64 * 1.05, (rms32, rms64)
    if multipole == 2:
        assert rms32 < 1e-3, rms32
    assert st[0] > 0 and st[1] > 0 and st[2] == -(-n // gs)


def test_theta_half_accuracy_two_masses(native):
    """The disk (mass 1/6) + halo (5/6) mix at theta = 0.5, where tiles of
    one mass and mixed tiles alternate: rms relative error against FP64
    direct summation below 1e-3 and no worse than the plain per-target walk."""
    x, m = workloads.disk_halo(100_000, 100_000, seed=8)
    n = len(x)
    dt = native.DeviceTree.build(x, m, 16)
    a32 = dt.accel_f32(None, 0.5, 0.01, 1.0)
    a64 = dt.accel_self(0.5, 0.01, 1.0)
    sel = np.random.default_rng(8).choice(n, 2000, replace=False)
    d = native.direct(x, m, x[sel], softening=0.01)
    rms32, _ = rel_err(a32[sel], d)
    rms64, _ = rel_err(a64[sel], d)
    assert rms32 < 1e-3 and rms32 <= rms64 * 1.05, (rms32, rms64)


def test_external_targets(native):
    x, m = workloads.plummer(50_000, seed=2)
    dt = native.DeviceTree.build(x, m, 16)
    tg = np.random.default_rng(2).uniform(-5, 5, (20_000, 3))
    a0 = dt.accel_f32(tg, 0.0, 0.01, 1.0)
    d = native.direct(x, m, tg, softening=0.01)
    rms, mx = rel_err(a0, d)
    assert rms < 2e-6 and mx < 2e-5
    a = dt.accel_f32(tg, 0.5, 0.01, 1.0)
    a64 = dt.accel_targets(tg, 0.5, 0.01, 1.0)
    assert rel_err(a, d)[0] <= rel_err(a64, d)[0] * 1.05


@pytest.mark.parametrize("dim,theta", [(2, 0.5), (3, 0.3), (3, 1.0)])
def test_external_targets_dims_and_thetas(native, dim, theta):
    """External targets (no self exclusion, bh.py:30-80) in 2-D and 3-D and
    at a tighter and a looser theta: an FP32 direct sum at theta = 0, and at
    theta > 0 no less accurate than the plain walk against FP64 direct
    summation (the group test opens only more than bhut's).  The 2-D
    direct-sum bound is the one of the 2-D self test."""
    rng = np.random.default_rng(40 + dim)
    x = rng.standard_normal((20_000, dim))
    m = rng.uniform(0.5, 1.5, 20_000)
    tg = rng.uniform(-3, 3, (5_000, dim))
    dt = native.DeviceTree.build(x, m, 16)
    d = native.direct(x, m, tg, softening=0.01)
    rms0, _ = rel_err(dt.accel_f32(tg, 0.0, 0.01, 1.0), d)
    assert rms0 < (2e-6 if dim == 3 else 1e-5), rms0  # in-plane sums cancel more (2-D)
    rms32, _ = rel_err(dt.accel_f32(tg, theta, 0.01, 1.0), d)
    rms64, _ = rel_err(dt.accel_targets(tg, theta, 0.01, 1.0), d)
    assert rms32 <= rms64 * 1.05, (rms32, rms64)


@pytest.mark.parametrize("masses", ["random", "equal"])
def test_two_dimensional_and_zero_softening(native, masses):
    """2-D, softened and unsoftened (the pair with r = 0 adds nothing), with
    per-pair masses and with the equal-mass tiles."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8000, 2))
    m = rng.uniform(0.5, 1.5, 8000) if masses == "random" else np.full(8000, 1.0 / 8000)
    dt = native.DeviceTree.build(x, m, 16)
    for eps in (0.01, 0.0):
        a = dt.accel_f32(None, 0.0, eps, 1.0)
        d = native.direct(x, m, softening=eps)
        rms, _ = rel_err(a, d)
        assert rms < 1e-5, (eps, rms)


@pytest.mark.parametrize("n", [1, 2, 63, 65, 130])
def test_small_and_partial_groups(native, n):
    """N below, at and just above one group of 64 targets (a partly filled
    last group, whose spare lanes hold copies of its first target): an FP32
    direct sum at theta = 0, at theta = 0.5 no worse than the plain walk; a
    lone particle has zero acceleration."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, 3))
    m = rng.uniform(0.5, 1.5, n)
    dt = native.DeviceTree.build(x, m, 8)
    for theta in (0.0, 0.5):
        a = dt.accel_f32(None, theta, 0.01, 1.0)
        if n == 1:
            assert np.all(a == 0)
            continue
        d = native.direct(x, m, softening=0.01)
        rms, _ = rel_err(a, d)
        rms64, _ = rel_err(dt.accel_self(theta, 0.01, 1.0), d)
        assert rms <= max(rms64 * 1.05, 2e-6), (theta, rms, rms64)


def test_coincident_particles_unsoftened(native):
    """Two particles at one position with no softening: the pair has
    r = 0 and adds nothing (the plain form's pure-Python leaf fallback,
    bh.py:307-311, gives the same), in both the equal-mass and the per-pair
    mass loops."""
    x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [2.0, 2.0, 3.0], [0.0, 0.5, 0.0]])
    for m in (np.array([1.0, 2.0, 3.0, 4.0]), np.full(4, 0.25)):
        dt = native.DeviceTree.build(x, m, 32)
        a = dt.accel_f32(None, 0.5, 0.0, 1.0)
        ref = dt.accel_self(0.5, 0.0, 1.0)
        assert np.all(np.isfinite(a))
        np.testing.assert_allclose(a, ref, rtol=1e-5, atol=1e-7)


def test_far_external_targets(native):
    """Targets around 1e6 times farther out than the sources, spread in all
    directions (their bounding box encloses the whole system, so the exact
    test decides): the root's moments serve them; FP32 relative error
    against FP64 direct summation stays at the FP32 level."""
    x, m = workloads.plummer(20_000, seed=14)
    dt = native.DeviceTree.build(x, m, 32)
    tg = np.random.default_rng(14).standard_normal((300, 3)) * 1e6
    a = dt.accel_f32(tg, 0.5, 0.01, 1.0)
    d = native.direct(x, m, tg, softening=0.01)
    rms, mx = rel_err(a, d)
    assert rms < 1e-5 and mx < 1e-4, (rms, mx)


def test_refit_then_fp32(native):
    x, m = workloads.plummer(30_000, seed=6)
    dt = native.DeviceTree.build(x, m, 32)
    dt.accel_f32(None, 0.5, 0.01, 1.0)
    x2 = x + 1e-3 * np.random.default_rng(6).standard_normal(x.shape)
    dt.refit(x2, m)
    a = dt.accel_f32(None, 0.0, 0.01, 1.0)
    rms, _ = rel_err(a, native.direct(x2, m, softening=0.01))
    assert rms < 2e-6


def test_small_stack_is_grown(native):
    """A tiny LDS stack bound, forced from the host: the walk kernel's own
    bound check stops the group, flags the overflow, and the count pass is
    rerun with a larger stack; the lists and results do not depend on it."""
    x, m = workloads.plummer(50_000, seed=7)
    dt = native.DeviceTree.build(x, m, 8)
    small = dt.accel_f32(None, 0.3, 0.01, 1.0, native.F32Opts(64, 1, 64, 0))
    ref = dt.accel_f32(None, 0.3, 0.01, 1.0, native.F32Opts(64, 1, 8192, 0))
    np.testing.assert_array_equal(small.view(np.uint64), ref.view(np.uint64))
    with pytest.raises(ValueError, match="stack_cap"):
        dt.accel_f32(None, 0.3, 0.01, 1.0, native.F32Opts(64, 1, 9000, 0))


def test_api_precision_keyword(native):
    import bhut_amd
    x, m = workloads.plummer(5000, seed=9)
    a32 = bhut_amd.accelerations(x, m, theta=0.0, softening=0.01, precision="fp32")
    a64 = bhut_amd.accelerations(x, m, theta=0.0, softening=0.01)
    assert rel_err(a32, a64)[0] < 2e-6
    t = bhut_amd.Tree(x, m).build()
    assert rel_err(t.accelerations(theta=0.0, softening=0.01, precision="fp32"), a64)[0] < 2e-6
    with pytest.raises(ValueError, match="precision"):
        t.accelerations(precision="fp16")


def test_one_pass_lists_match_counted_lists(native):
    """A context's first call on a set of groups counts the lists, then fills
    them packed; later calls on the same groups write them in one pass into
    regions sized from the previous call's list lengths -- or count and pack
    again when the regions exceed the list budget.  At theta = 0 every leaf
    is in every group's list.  The lists, and so the results, are
    bit-identical on every path, also when regions sized from shorter lists
    overflow."""
    x, m = workloads.plummer(300_000, seed=13)
    c = native.Context(0)
    t = native.DeviceTree.build(x, m, 32, context=c)
    first = t.accel_f32(None, 0.0, 0.01, 1.0)   # no previous call: counted and packed
    second = t.accel_f32(None, 0.0, 0.01, 1.0)  # regions sized by the first call: one pass
    np.testing.assert_array_equal(second.view(np.uint64), first.view(np.uint64))
    c.set_list_budget(1 << 20)                   # regions over budget: counted, in batches
    third = t.accel_f32(None, 0.0, 0.01, 1.0)
    np.testing.assert_array_equal(third.view(np.uint64), first.view(np.uint64))
    c.set_list_budget(0)
    t.accel_f32(None, 0.5, 0.01, 1.0)            # short lists
    fourth = t.accel_f32(None, 0.0, 0.01, 1.0)  # regions from the short lists overflow: counted again
    np.testing.assert_array_equal(fourth.view(np.uint64), first.view(np.uint64))
    d = native.direct(x[:2000], m[:2000], softening=0.01)
    sub = native.DeviceTree.build(x[:2000], m[:2000], 32, context=c).accel_f32(None, 0.0, 0.01, 1.0)
    assert rel_err(sub, d)[0] < 2e-6


def test_list_batches_give_identical_results(native):
    """A small list budget splits the groups into many batches (each filled,
    ordered and evaluated on its own); every group's lists and sums are the
    same, so the results are bit-identical to one batch."""
    x, m = workloads.plummer(30_000, seed=12)
    rng = np.random.default_rng(3)
    tg = rng.uniform(-3, 3, (7000, 3))
    t = native.DeviceTree.build(x, m, 32)
    c = native.ctx()
    full = t.accel_f32(None, 0.5, 0.01, 1.0)
    full_t = t.accel_f32(tg, 0.5, 0.01, 1.0)
    try:
        for budget in (1 << 20, 200_000, 1):
            c.set_list_budget(budget)
            np.testing.assert_array_equal(t.accel_f32(None, 0.5, 0.01, 1.0).view(np.uint64), full.view(np.uint64))
            np.testing.assert_array_equal(t.accel_f32(tg, 0.5, 0.01, 1.0).view(np.uint64), full_t.view(np.uint64))
    finally:
        c.set_list_budget(0)


def test_full_size_plummer(native):
    """1M Plummer at theta 0.5 (the headline configuration): rms relative
    error against FP64 direct summation below 1e-3 and no worse than the
    plain walk's; a second call (one-pass lists in regions sized by the
    first, which counted and packed them) gives the same bits; one group
    per 64 targets."""
    n = 1_000_000
    x, m = workloads.plummer(n, seed=1)
    dt = native.DeviceTree.build(x, m, 32)
    opts = native.F32Opts(64, 1, 1024, 2)
    a1, st = dt.accel_f32(None, 0.5, 0.01, 1.0, opts, stats=True)
    a2 = dt.accel_f32(None, 0.5, 0.01, 1.0, opts)
    np.testing.assert_array_equal(a1.view(np.uint64), a2.view(np.uint64))
    sel = np.random.default_rng(3).choice(n, 2000, replace=False)
    d = native.direct(x, m, x[sel], softening=0.01)
    rms32, _ = rel_err(a1[sel], d)
    rms64, _ = rel_err(dt.accel_self(0.5, 0.01, 1.0)[sel], d)
    assert rms32 < 1e-3 and rms32 <= rms64 * 1.05, (rms32, rms64)
    assert st[2] == n // 64
