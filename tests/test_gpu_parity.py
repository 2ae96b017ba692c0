"""Parity of the HIP plain form with the oracle and the golden fixtures.

Bit-exact: Morton order (perm), topology, node boxes, mass, centre of mass,
refit, per-target interaction lists and exact counters.  Accelerations: the
same summation order as bhut; the only arithmetic difference is r2s*sqrt(r2s)
in place of libm pow(r2s, 1.5) in monopoles, so per target
    ||a_gpu - a_ref|| <= 1e-12 * ||a_ref||   (with a floor of 1e-12 * max||a_ref||
for targets whose acceleration cancels to near zero).
"""

import os

import numpy as np
import pytest

from conftest import golden_files
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TREE_FIELDS = ["perm", "start", "count", "first_child", "child_count", "is_leaf", "center", "half_size", "mass",
               "com"]


def bits_equal(a, b, name=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, name
    if a.dtype == np.float64:
        np.testing.assert_array_equal(a.view(np.uint64), np.asarray(b, np.float64).view(np.uint64), err_msg=name)
    else:
        np.testing.assert_array_equal(a, b, err_msg=name)


def acc_close(got, ref, tol=1e-12):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape
    nr = np.linalg.norm(ref, axis=1)
    floor = tol * (nr.max() if len(nr) else 0.0)
    err = np.linalg.norm(got - ref, axis=1)
    bad = err > np.maximum(tol * nr, floor)
    assert not bad.any(), f"{bad.sum()} targets off; worst rel {np.max(err / np.maximum(nr, 1e-300)):.3e}"


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


@pytest.fixture(params=golden_files(), ids=lambda p: os.path.basename(p)[7:-4])
def golden(request):
    return np.load(request.param, allow_pickle=False)


def test_golden_tree_bit_exact(native, golden):
    g = golden
    for leaf in g["leaves"]:
        dt = native.DeviceTree.build(g["pos"], g["mass"], int(leaf))
        e = dt.export()
        for f in TREE_FIELDS:
            bits_equal(e[f], g[f"L{leaf}_{f}"], f"L{leaf} {f}")


def test_golden_refit_bit_exact(native, golden):
    g = golden
    for leaf in g["leaves"]:
        dt = native.DeviceTree.build(g["pos"], g["mass"], int(leaf))
        dt.refit(g["pos2"], g["mass"])
        e = dt.export()
        bits_equal(e["mass"], g[f"L{leaf}_refit_mass"], "refit mass")
        bits_equal(e["com"], g[f"L{leaf}_refit_com"], "refit com")
        o = O.build(g["pos"], g["mass"], int(leaf))
        bits_equal(e["center"], o.center, "center kept")
        bits_equal(e["half_size"], o.half_size, "half_size kept")


@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_golden_lists_and_accelerations(native, golden, eps):
    g = golden
    pos, mass, ext = g["pos"], g["mass"], g["ext"]
    for leaf in g["leaves"]:
        dt = native.DeviceTree.build(pos, mass, int(leaf))
        for theta in g["thetas"]:
            tag = f"L{leaf}_t{theta}_e{eps}"
            acc = dt.accel_self(theta, eps, 1.0)
            nt = len(g[f"{tag}_acc_numba"])
            acc_close(acc[:nt], g[f"{tag}_acc_numba"])
            ptr, ids = dt.interaction_lists(None, theta)
            gp, gi = g[f"{tag}_list_ptr"], g[f"{tag}_list_ids"]
            for i in range(nt):
                np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], gi[gp[i]:gp[i + 1]], err_msg=f"{tag} t{i}")
            acc_e = dt.accel_targets(ext, theta, eps, 1.0)
            acc_close(acc_e, g[f"{tag}_ext_acc"])
            ptr, ids = dt.interaction_lists(ext, theta)
            gp, gi = g[f"{tag}_ext_list_ptr"], g[f"{tag}_ext_list_ids"]
            for i in range(len(ext)):
                np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], gi[gp[i]:gp[i + 1]])


REF_STAT_KEYS = ("avg_nodes_visited", "avg_particles_computed", "avg_monopoles_used", "total_operations")


def ref_files():
    d = os.path.join(os.path.dirname(__file__), "golden")
    return sorted(os.path.join(d, f) for f in os.listdir(d)
                  if f.startswith("ref_") and f.endswith(".npz") and f != "ref_normal10k.npz")


@pytest.mark.parametrize("path", ref_files(), ids=lambda p: os.path.basename(p)[4:-4])
def test_reference_walk_fixtures(native, path):
    """The HIP plain form against outputs the reference's own walk wrote
    (tests/golden/make_golden_ref.py): TreeData bit for bit, per-target
    lists in pop order, pops / opening tests / monopoles as the reference
    counts them, compute_tree_statistics exactly (also at displaced
    positions of the same count), accelerations within 1e-12 of both leaf
    formulas."""
    from bhut_amd.traverse.bh import compute_tree_statistics
    from bhut_amd.tree.build import build_tree
    g = np.load(path, allow_pickle=False)
    pos, mass, ext = g["pos"], g["mass"], g["ext"]
    n, dim = pos.shape
    for leaf in g["leaves"]:
        tree = build_tree(pos, mass, dim=dim, leaf_size=int(leaf))
        dt = tree.device_tree
        e = dt.export()
        for f in TREE_FIELDS:
            bits_equal(e[f], g[f"L{leaf}_{f}"], f"L{leaf} {f}")
        for theta in g["thetas"]:
            tag = f"L{leaf}_t{theta}"
            ops = g[f"{tag}_ops"]
            nt = len(ops)
            c = dt.counters(None, theta)[:nt]
            np.testing.assert_array_equal(c[:, 0], ops[:, 0], err_msg=tag)
            np.testing.assert_array_equal(c[:, 2], ops[:, 2], err_msg=tag)
            np.testing.assert_array_equal(c[:, 1], g[f"{tag}_e0.0_tests"], err_msg=tag)
            ce = dt.counters(ext, theta)
            np.testing.assert_array_equal(ce[:, [0, 2]], g[f"{tag}_ext_ops"][:, [0, 2]], err_msg=tag)
            np.testing.assert_array_equal(ce[:, 4] - ce[:, 3], g[f"{tag}_ext_ops"][:, 1], err_msg=tag)
            for key, p in (("stats", pos), ("stats_pos2", g["pos2"])):
                st = compute_tree_statistics(tree, p, theta)
                np.testing.assert_array_equal(np.array([float(st[k]) for k in REF_STAT_KEYS]),
                                              g[f"{tag}_{key}"], err_msg=f"{tag} {key}")
            ptr, ids = dt.interaction_lists(None, theta)
            gp, gi = g[f"{tag}_e0.0_list_ptr"], g[f"{tag}_e0.0_list_ids"]
            for i in range(nt):
                np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], gi[gp[i]:gp[i + 1]], err_msg=f"{tag} t{i}")
            ptr, ids = dt.interaction_lists(ext, theta)
            gp, gi = g[f"{tag}_e0.0_ext_list_ptr"], g[f"{tag}_e0.0_ext_list_ids"]
            for i in range(len(ext)):
                np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], gi[gp[i]:gp[i + 1]], err_msg=f"{tag} e{i}")
            for eps in (0.0, 0.01):
                tag = f"L{leaf}_t{theta}_e{eps}"
                acc = dt.accel_self(theta, eps, 1.0)[:nt]
                acc_close(acc, g[f"{tag}_acc_numba"])
                acc_close(acc, g[f"{tag}_acc_python"])
                acc_e = dt.accel_targets(ext, theta, eps, 1.0)
                acc_close(acc_e, g[f"{tag}_ext_acc_numba"])
                acc_close(acc_e, g[f"{tag}_ext_acc_python"])


def test_reference_normal10k_public_api(native):
    """bhut_amd.accelerations on the CPU-runnable workload (10 000
    standard-normal particles, seed 0, unit masses, theta 0.5, softening
    0.01, leaf 32) against bhut.accelerations' own output for every
    particle, within 1e-12; sampled lists and the statistics exactly."""
    import bhut_amd
    from bhut_amd.traverse.bh import compute_tree_statistics
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_normal10k.npz"), allow_pickle=False)
    x, m = g["pos"], g["mass"]
    acc_close(bhut_amd.accelerations(x, m, theta=0.5, softening=0.01, G=1.0, leaf_size=32), g["acc"])
    t = bhut_amd.Tree(x, m, leaf_size=32).build()
    sel = g["sel"]
    acc_close(t.accelerations(theta=0.5, softening=0.01)[sel], g["sel_acc_numba"])
    ptr, ids = t._tree.device_tree.interaction_lists(None, 0.5)
    gp, gi = g["sel_list_ptr"], g["sel_list_ids"]
    for k, i in enumerate(sel):
        np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], gi[gp[k]:gp[k + 1]])
    st = compute_tree_statistics(t._tree, x, 0.5)
    np.testing.assert_array_equal(np.array([float(st[k]) for k in REF_STAT_KEYS]), g["stats"])


def plummer(n, seed, a=1.0, rmax=100.0):
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    fill = 0
    while fill < n:
        u = rng.uniform(0, 1, n - fill)
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr < rmax]
        r[fill:fill + len(rr)] = rr
        fill += len(rr)
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    return np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)


@pytest.mark.parametrize("n,leaf,theta", [(10_000, 32, 0.5), (20_000, 7, 0.3), (65_537, 64, 0.7)])
def test_medium_against_oracle(native, n, leaf, theta):
    """Config-1 style inputs: tree bit-exact, lists and counters exact,
    accelerations to 1e-12, against the oracle on the same seeded inputs."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, 3)) if n != 20_000 else plummer(n, 5)
    m = np.ones(n) if n == 10_000 else rng.uniform(0.1, 2.0, n)
    eps = 0.01
    dt = native.DeviceTree.build(x, m, leaf)
    ot = O.build(x, m, leaf)
    e = dt.export()
    for f in TREE_FIELDS:
        bits_equal(e[f], getattr(ot, f), f)
    acc = dt.accel_self(theta, eps, 1.0)
    ref = O.accelerations(ot, x, m, softening=eps, theta=theta)
    acc_close(acc, ref)
    sel = rng.choice(n, 64, replace=False)
    ptr, ids = dt.interaction_lists(None, theta)
    for i in sel:
        np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], O.target_list(ot, x[i], theta))
    cnt = dt.counters(None, theta)
    np.testing.assert_array_equal(cnt, O.counters(ot, x, self_mode=True, theta=theta))
    tg = rng.uniform(-4, 4, (3000, 3))
    acc_close(dt.accel_targets(tg, theta, eps, 1.0), O.accelerations(ot, x, m, tg, softening=eps, theta=theta))


def test_two_dimensional(native):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5000, 2))
    m = rng.uniform(0.5, 1.5, 5000)
    dt = native.DeviceTree.build(x, m, 16)
    ot = O.build(x, m, 16)
    e = dt.export()
    for f in TREE_FIELDS:
        bits_equal(e[f], getattr(ot, f), f)
    acc_close(dt.accel_self(0.5, 0.01, 1.0), O.accelerations(ot, x, m, softening=0.01, theta=0.5))


def test_refit_matches_oracle_and_rebuild_rule(native):
    import bhut_amd
    rng = np.random.default_rng(8)
    x = plummer(30_000, 8)
    m = np.full(len(x), 1.0 / len(x))
    t = bhut_amd.Tree(x, m).build()
    x2 = x + 0.001 * rng.standard_normal(x.shape)
    t.refit(x2)
    ot = O.refit(O.build(x, m, 32), x2, m)
    for f in ("mass", "com", "center", "half_size"):
        bits_equal(getattr(t._tree, f), getattr(ot, f), f)
    acc_close(t.accelerations(theta=0.5, softening=0.01), O.accelerations(ot, x2, m, softening=0.01, theta=0.5))
    # a particle far outside 1.1 root half sizes forces a rebuild (refit.py:236-249)
    x3 = x2.copy()
    x3[0] = 1e4
    assert not O.should_refit(ot, x3)
    t.refit(x3)
    ob = O.build(x3, m, 32)
    bits_equal(t._tree.perm, ob.perm, "rebuilt perm")
    bits_equal(t._tree.half_size, ob.half_size, "rebuilt half_size")


def test_refit_tree_on_clone_and_import(native):
    """refit_tree (refit.py:17-109) on a tree built here refits a clone, and
    on host-only arrays an imported tree: both carry no depth tables and go
    up by the atomic-flag climb instead of one launch per depth.  Either way
    the oracle's bits, with unequal masses (NumPy's pairwise leaf sums)."""
    from bhut_amd.tree.build import build_tree
    from bhut_amd.tree.node import TreeData
    from bhut_amd.tree.refit import refit_tree
    rng = np.random.default_rng(12)
    x = plummer(20_000, 12)
    m = rng.uniform(0.5, 1.5, len(x)) / len(x)
    x2 = x + 0.001 * rng.standard_normal(x.shape)
    ot = O.refit(O.build(x, m, 32), x2, m)
    tree = build_tree(x, m, dim=3, leaf_size=32)
    host = TreeData(**tree.arrays(), dim=3, leaf_size=32)
    assert host.device_tree is None
    for name, src in (("clone", tree), ("import", host)):
        r = refit_tree(src, x2, m)
        for f in ("mass", "com", "center", "half_size"):
            bits_equal(getattr(r, f), getattr(ot, f), f"{name} {f}")
    # the source tree is untouched by the refit of its clone
    bits_equal(tree.device_tree.export()["mass"], O.build(x, m, 32).mass, "source mass")


def test_traversal_with_other_source_arrays(native):
    """barnes_hut_accelerations(_targets) (bh.py:428-482, 30-223) walking a
    TreeData with source arrays other than the ones it was built from (the
    tree is imported with those arrays), and with host-only TreeData: the
    oracle's accelerations to 1e-12, lists exact."""
    from bhut_amd.traverse.bh import (barnes_hut_accelerations, barnes_hut_accelerations_targets,
                                      interaction_lists)
    from bhut_amd.tree.build import build_tree
    from bhut_amd.tree.node import TreeData
    rng = np.random.default_rng(13)
    x = plummer(8_000, 13)
    m = rng.uniform(0.5, 1.5, len(x))
    ot = O.build(x, m, 16)
    tree = build_tree(x, m, dim=3, leaf_size=16)
    host = TreeData(**tree.arrays(), dim=3, leaf_size=16)
    m2 = m * 2.0  # other source masses on the same tree
    tg = rng.uniform(-3, 3, (500, 3))
    for name, tr in (("device", tree), ("host", host)):
        acc_close(barnes_hut_accelerations(tr, x.copy(), m2, 0.01, 0.5),
                  O.accelerations(ot, x, m2, softening=0.01, theta=0.5))
        acc_close(barnes_hut_accelerations_targets(tr, x.copy(), m2, tg, 0.01, 0.5),
                  O.accelerations(ot, x, m2, tg, softening=0.01, theta=0.5))
        ptr, ids = interaction_lists(tr, x.copy(), m, 0.5)
        for i in (0, 17, 4000):
            np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], O.target_list(ot, x[i], 0.5), err_msg=name)


def test_edge_cases(native):
    import bhut_amd
    # single particle: zero acceleration
    a = bhut_amd.accelerations(np.array([[1.0, 2.0, 3.0]]), np.array([1.0]))
    assert a.shape == (1, 3) and np.all(a == 0)
    # coincident particles with eps = 0: Numba's ZeroDivisionError fallback
    # to the pure-Python leaf gives zero for the coincident pair
    x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [2.0, 2.0, 3.0]])
    m = np.array([1.0, 2.0, 3.0])
    ot = O.build(x, m, 32)
    acc_close(bhut_amd.accelerations(x, m), O.accelerations(ot, x, m, softening=0.0, theta=0.5))
    # leaf_size larger than N, leaf_size 1, zero and negative masses
    rng = np.random.default_rng(4)
    x = rng.standard_normal((300, 3))
    m = rng.uniform(-0.2, 1.0, 300)
    m[:20] = 0.0
    for leaf in (1, 3, 1000):
        ot = O.build(x, m, leaf)
        dt = native.DeviceTree.build(x, m, leaf)
        e = dt.export()
        for f in TREE_FIELDS:
            bits_equal(e[f], getattr(ot, f), f"leaf {leaf} {f}")
        acc_close(dt.accel_self(0.5, 0.05, 1.0), O.accelerations(ot, x, m, softening=0.05, theta=0.5))
    # zero-mass system: every node skipped (bh.py:512-513)
    dt = native.DeviceTree.build(x, np.zeros(300), 8)
    assert np.all(dt.accel_self(0.5, 0.01, 1.0) == 0)


def test_leaf_larger_than_pairwise_block(native):
    """Leaves above 128 particles take NumPy's recursive pairwise split, and
    leaves above 8192 np.sum's buffers of 8192 elements (leaf masses equal
    np.sum of the leaf's sorted masses, bit for bit)."""
    rng = np.random.default_rng(6)
    for n, leaves in ((3000, (200, 700)), (30_000, (16_000, 30_000))):
        x = rng.standard_normal((n, 3))
        m = rng.uniform(0.0, 1.0, n) * 10.0 ** rng.integers(-6, 6, n)
        for leaf in leaves:
            dt = native.DeviceTree.build(x, m, leaf)
            ot = O.build(x, m, leaf)
            e = dt.export()
            for f in ("mass", "com"):
                bits_equal(e[f], getattr(ot, f), f"leaf {leaf} {f}")
            for v in np.flatnonzero(e["is_leaf"]):
                a, c = int(e["start"][v]), int(e["count"][v])
                assert e["mass"][v] == np.sum(m[e["perm"][a:a + c]]), (leaf, v)


def test_full_size_plummer_properties(native):
    """1M Plummer (the headline configuration): size-independent checks."""
    n = 1_000_000
    x = plummer(n, 1)
    m = np.full(n, 1.0 / n)
    dt = native.DeviceTree.build(x, m, 32)
    e = dt.export()
    assert e["mass"].shape == (65535,) and int(e["is_leaf"].sum()) == 32768
    assert np.array_equal(np.sort(e["perm"]), np.arange(n))
    np.testing.assert_array_equal(e["perm"], O.argsort_stable(O.morton_codes(x)))
    ot = O.topology(n, 32)
    for f in ("start", "count", "first_child"):
        np.testing.assert_array_equal(e[f], getattr(ot, f))
    np.testing.assert_allclose(e["mass"][0], 1.0, rtol=1e-12)
    # sampled exact lists and 1e-12 accelerations against the oracle walk
    full = O.build(x, m, 32)
    for f in ("center", "half_size", "mass", "com"):
        bits_equal(e[f], getattr(full, f), f)
    rng = np.random.default_rng(1)
    sel = rng.choice(n, 32, replace=False)
    ptr, ids = dt.interaction_lists(None, 0.5)
    for i in sel:
        np.testing.assert_array_equal(ids[ptr[i]:ptr[i + 1]], O.target_list(full, x[i], 0.5))
    acc = dt.accel_self(0.5, 0.01, 1.0)
    sub = rng.choice(n, 2000, replace=False)
    # the self walk of these particles (each skipping itself) in the oracle
    acc_close(acc[sub], O.accelerations_subset(full, x, m, sub, softening=0.01, theta=0.5))
    ref = O.accelerations(full, x, m, x[sub], softening=0.01, theta=0.5)
    acc_close(native.DeviceTree.accel_targets(dt, x[sub], 0.5, 0.01, 1.0), ref)


@pytest.mark.parametrize("dim", [3, 2])
def test_direct_sum_matches_oracle(native, dim):
    """The FP64 direct sum (k_direct), the accuracy reference of every FP32
    test and of the bench line, against the oracle's direct sum within 1e-12:
    self (j != i) and external targets, softened and unsoftened with a
    coincident pair (that pair is skipped, r^2 + eps^2 == 0), potential too.
    Its chunked summation differs from the oracle's sequential one only in
    rounding."""
    rng = np.random.default_rng(50 + dim)
    n = 20_000
    x = rng.standard_normal((n, dim))
    x[7] = x[3]  # coincident particles
    m = rng.uniform(0.1, 2.0, n)
    tg = rng.uniform(-3, 3, (3000, dim))
    tg[0] = x[11]
    for eps in (0.0, 0.02):
        acc, pot = native.direct(x, m, softening=eps, potential=True)
        acc_close(acc, O.direct(x, m, softening=eps))
        np.testing.assert_allclose(pot, O.direct_potential(x, m, softening=eps), rtol=1e-12)
        acc, pot = native.direct(x, m, tg, softening=eps, potential=True)
        acc_close(acc, O.direct(x, m, tg, softening=eps))
        np.testing.assert_allclose(pot, O.direct_potential(x, m, tg, softening=eps), rtol=1e-12)
    # theta = 0 walk == direct sum (bh.py:600-601)
    dt = native.DeviceTree.build(x, m, 32)
    acc_close(dt.accel_self(0.0, 0.02, 1.0), O.direct(x, m, softening=0.02))


@pytest.mark.parametrize("dim", [3, 2])
def test_theta0_walk_bitwise_against_oracle(native, dim):
    """At theta = 0 every node is opened and every leaf summed directly
    (bh.py:600-601), so nothing but the leaf formula and the pop order enter
    the result, and the plain walk must equal the oracle (glibc's correctly
    rounded sqrt, IEEE division, the reference's summation order) bit for
    bit.  Softened (the walk's middle-range square root and reciprocal), at
    G = 1 (the mass product skipped, exact) and at other G, and unsoftened
    (sqrt() and '/'); self and external targets."""
    rng = np.random.default_rng(70 + dim)
    n = 6000
    x = rng.standard_normal((n, dim))
    m = rng.uniform(0.5, 1.5, n)
    q = rng.uniform(-3, 3, (1500, dim))
    dt = native.DeviceTree.build(x, m, 16)
    ot = O.build(x, m, 16)
    for eps, G in ((0.01, 1.0), (0.01, 6.674e-11), (1e-3, 2.5), (0.0, 1.0)):
        bits_equal(dt.accel_self(0.0, eps, G), O.accelerations(ot, x, m, softening=eps, theta=0.0, G=G),
                   f"self eps {eps} G {G}")
        bits_equal(dt.accel_targets(q, 0.0, eps, G), O.accelerations(ot, x, m, q, softening=eps, theta=0.0, G=G),
                   f"targets eps {eps} G {G}")


def test_external_targets_with_non_finite_rows(native):
    """External targets are walked in wavefronts of 64 whose lanes all compute
    every pair of a visited leaf; a target with a NaN or infinite coordinate
    sharing a wavefront with finite ones leaves their results bit for bit
    what they are without it (each lane's sums are its own), and its own row
    follows the oracle (non-finite where the oracle's is)."""
    rng = np.random.default_rng(77)
    x = rng.standard_normal((20_000, 3))
    m = rng.uniform(0.5, 1.5, 20_000)
    q = rng.standard_normal((700, 3)) * 0.7
    bad = [7, 100, 101, 650]
    q[7, 0] = np.nan
    q[100, 2] = np.inf
    q[101] = -np.inf
    q[650, 1] = np.nan
    dt = native.DeviceTree.build(x, m, 32)
    ot = O.build(x, m, 32)
    good = np.setdiff1d(np.arange(len(q)), bad)
    for eps, G in ((0.01, 1.0), (0.01, 3.0), (0.0, 1.0)):
        got = dt.accel_targets(q, 0.5, eps, G)
        alone = dt.accel_targets(q[good], 0.5, eps, G)
        bits_equal(got[good], alone, f"finite rows, eps {eps} G {G}")
        ref = O.accelerations(ot, x, m, q, softening=eps, theta=0.5, G=G)
        acc_close(got[good], ref[good])
        np.testing.assert_array_equal(np.isfinite(got[bad]), np.isfinite(ref[bad]))
