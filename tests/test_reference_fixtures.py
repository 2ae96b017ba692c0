"""The C oracle against outputs the reference itself wrote (CPU only).

tests/golden/ref_*.npz were written by tests/golden/make_golden_ref.py, which
ran maroba/bhut's own traversal (bhut/traverse/bh.py) in the build container:
per-target interaction lists in pop order, the walk's accelerations with the
pure-Python leaf path and with the Numba leaf loop run by CPython, per-target
operation counts, compute_tree_statistics, and the reference's TreeData.  These
tests pin the oracle's walk order, its counters and its statistics against
the reference directly, not against a second restatement.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIELDS = ["center", "half_size", "mass", "com", "first_child", "child_count", "start", "count",
          "is_leaf", "perm"]
STAT_KEYS = ("avg_nodes_visited", "avg_particles_computed", "avg_monopoles_used", "total_operations")


def ref_files():
    return sorted(os.path.join(GOLD, f) for f in os.listdir(GOLD)
                  if f.startswith("ref_") and f.endswith(".npz") and f != "ref_normal10k.npz")


def bits_equal(a, b, msg=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, msg
    if a.dtype == np.float64 or b.dtype == np.float64:
        np.testing.assert_array_equal(np.asarray(a, np.float64).view(np.uint64),
                                      np.asarray(b, np.float64).view(np.uint64), err_msg=msg)
    else:
        np.testing.assert_array_equal(a, b.astype(a.dtype), err_msg=msg)


@pytest.fixture(params=ref_files(), ids=lambda p: os.path.basename(p)[4:-4])
def ref(request):
    return np.load(request.param, allow_pickle=False)


def test_fixtures_present():
    assert len(ref_files()) == 10 and os.path.exists(os.path.join(GOLD, "ref_normal10k.npz"))


def test_tree_equals_reference_treedata(ref):
    for leaf in ref["leaves"]:
        t = O.build(ref["pos"], ref["mass"], int(leaf))
        for f in FIELDS:
            bits_equal(getattr(t, f), ref[f"L{leaf}_{f}"], f"L{leaf} {f}")


def _derived_counts(t, lists_ptr, lists_ids, self_idx):
    """(leaves, exact pairs) per target from the reference's lists: a leaf
    contributes its count, minus one where the target is one of its
    particles in self mode (bh.py:253-255)."""
    n = len(lists_ptr) - 1
    leaves = np.zeros(n, np.int64)
    pairs = np.zeros(n, np.int64)
    for i in range(n):
        ids = lists_ids[lists_ptr[i]:lists_ptr[i + 1]]
        lf = ids[t.is_leaf[ids]]
        leaves[i] = len(lf)
        pairs[i] = int(t.count[lf].sum())
        if self_idx is not None:
            for k in lf:
                if self_idx[i] in t.perm[t.start[k]:t.start[k] + t.count[k]]:
                    pairs[i] -= 1
    return leaves, pairs


@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_walk_lists_and_accelerations(ref, eps):
    """Pop-order lists, self and external, and both leaf formulas, bit for
    bit against the reference's own walk."""
    pos, mass, ext = ref["pos"], ref["mass"], ref["ext"]
    for leaf in ref["leaves"]:
        t = O.build(pos, mass, int(leaf))
        for theta in ref["thetas"]:
            tag = f"L{leaf}_t{theta}_e{eps}"
            nt = len(ref[f"{tag}_acc_python"])
            bits_equal(O.accelerations(t, pos, mass, softening=eps, theta=theta, leaf_formula=O.LEAF_PYTHON)[:nt],
                       ref[f"{tag}_acc_python"], tag + " python")
            bits_equal(O.accelerations(t, pos, mass, softening=eps, theta=theta, leaf_formula=O.LEAF_NUMBA)[:nt],
                       ref[f"{tag}_acc_numba"], tag + " numba")
            bits_equal(O.accelerations(t, pos, mass, ext, softening=eps, theta=theta, leaf_formula=O.LEAF_PYTHON),
                       ref[f"{tag}_ext_acc_python"], tag + " ext python")
            bits_equal(O.accelerations(t, pos, mass, ext, softening=eps, theta=theta, leaf_formula=O.LEAF_NUMBA),
                       ref[f"{tag}_ext_acc_numba"], tag + " ext numba")
            for kind, tp in (("", pos[:nt]), ("_ext", ext)):
                p, ids = ref[f"{tag}{kind}_list_ptr"], ref[f"{tag}{kind}_list_ids"]
                for i in range(len(tp)):
                    np.testing.assert_array_equal(O.target_list(t, tp[i], theta), ids[p[i]:p[i + 1]],
                                                  err_msg=f"{tag}{kind} target {i}")


def test_counters_against_reference(ref):
    """O.counters per target: pops, monopoles and opening tests as the
    reference counts them (_count_operations_single_target and the walk's
    calls of _should_open_node); leaves and exact pairs as the reference's
    lists imply."""
    pos, ext = ref["pos"], ref["ext"]
    for leaf in ref["leaves"]:
        t = O.build(pos, ref["mass"], int(leaf))
        for theta in ref["thetas"]:
            tag = f"L{leaf}_t{theta}"
            ops = ref[f"{tag}_ops"]
            nt = len(ops)
            c = O.counters(t, pos, self_mode=True, theta=theta)[:nt]
            np.testing.assert_array_equal(c[:, 0], ops[:, 0], err_msg=tag + " pops")
            np.testing.assert_array_equal(c[:, 2], ops[:, 2], err_msg=tag + " monopoles")
            np.testing.assert_array_equal(c[:, 1], ref[f"{tag}_e0.0_tests"], err_msg=tag + " tests")
            lv, pr = _derived_counts(t, ref[f"{tag}_e0.0_list_ptr"], ref[f"{tag}_e0.0_list_ids"], np.arange(nt))
            np.testing.assert_array_equal(c[:, 3], lv, err_msg=tag + " leaves")
            np.testing.assert_array_equal(c[:, 4], pr, err_msg=tag + " pairs")
            # the reference's particles_computed: count - 1 per visited leaf
            np.testing.assert_array_equal(
                O.counters(t, pos[:nt], self_mode=False, theta=theta)[:, 4] - c[:, 3], ops[:, 1], err_msg=tag)
            ce = O.counters(t, ext, self_mode=False, theta=theta)
            eops = ref[f"{tag}_ext_ops"]
            np.testing.assert_array_equal(ce[:, 0], eops[:, 0])
            np.testing.assert_array_equal(ce[:, 2], eops[:, 2])
            np.testing.assert_array_equal(ce[:, 1], ref[f"{tag}_e0.0_ext_tests"])
            lv, pr = _derived_counts(t, ref[f"{tag}_e0.0_ext_list_ptr"], ref[f"{tag}_e0.0_ext_list_ids"], None)
            np.testing.assert_array_equal(ce[:, 3], lv)
            np.testing.assert_array_equal(ce[:, 4], pr)
            np.testing.assert_array_equal(ce[:, 4] - ce[:, 3], eops[:, 1])


def test_statistics_against_reference(ref):
    """compute_tree_statistics at the tree's positions and at displaced
    positions of the same count: the reference's four values exactly."""
    for leaf in ref["leaves"]:
        t = O.build(ref["pos"], ref["mass"], int(leaf))
        for theta in ref["thetas"]:
            for key, p in (("stats", ref["pos"]), ("stats_pos2", ref["pos2"])):
                got = O.statistics(t, p, theta)
                exp = ref[f"L{leaf}_t{theta}_{key}"]
                np.testing.assert_array_equal(np.array([float(got[k]) for k in STAT_KEYS]), exp,
                                              err_msg=f"L{leaf} t{theta} {key}")


def test_counters_theta_zero_closed_form():
    """theta = 0 opens every internal node with mass (bh.py:600-601): each
    target pops every node, tests every internal node, accepts no monopole,
    visits every leaf and sums N - 1 pairs (N externally)."""
    rng = np.random.default_rng(17)
    for n, leaf in ((1, 4), (7, 1), (300, 8), (1025, 16)):
        x = rng.standard_normal((n, 3))
        m = rng.uniform(0.5, 1.5, n)
        t = O.build(x, m, leaf)
        M, nleaves = len(t.mass), int(t.is_leaf.sum())
        c = O.counters(t, x, self_mode=True, theta=0.0)
        np.testing.assert_array_equal(c, np.tile([M, M - nleaves, 0, nleaves, n - 1], (n, 1)))
        ce = O.counters(t, x + 0.5, self_mode=False, theta=0.0)
        np.testing.assert_array_equal(ce, np.tile([M, M - nleaves, 0, nleaves, n], (n, 1)))


def test_normal10k_public_api():
    """bhut.accelerations on 10 000 standard-normal particles (seed 0), unit
    masses, theta 0.5, softening 0.01, leaf 32 -- the reference's own output
    (pure-Python leaf path) -- equals the oracle bit for bit, for every
    particle; lists, Numba-formula accelerations and counts of 256 sampled
    targets too."""
    g = np.load(os.path.join(GOLD, "ref_normal10k.npz"), allow_pickle=False)
    x, m = g["pos"], g["mass"]
    t = O.build(x, m, 32)
    for f in FIELDS:
        bits_equal(getattr(t, f), g[f"L32_{f}"], f)
    bits_equal(O.accelerations(t, x, m, softening=0.01, theta=0.5, leaf_formula=O.LEAF_PYTHON), g["acc"], "acc")
    sel = g["sel"]
    bits_equal(O.accelerations(t, x, m, softening=0.01, theta=0.5)[sel], g["sel_acc_numba"], "numba")
    p, ids = g["sel_list_ptr"], g["sel_list_ids"]
    for k, i in enumerate(sel):
        np.testing.assert_array_equal(O.target_list(t, x[i], 0.5), ids[p[k]:p[k + 1]])
    c = O.counters(t, x, self_mode=True, theta=0.5)[sel]
    np.testing.assert_array_equal(c[:, 0], g["sel_ops"][:, 0])
    np.testing.assert_array_equal(c[:, 2], g["sel_ops"][:, 2])
    np.testing.assert_array_equal(c[:, 1], g["sel_tests"])
    st = O.statistics(t, x, 0.5)
    np.testing.assert_array_equal(np.array([float(st[k]) for k in STAT_KEYS]), g["stats"])


def _preorder_rank(t, right_first: bool):
    rank = np.zeros(len(t.mass), np.int64)
    stack, r = [0], 0
    while stack:
        v = stack.pop()
        rank[v] = r
        r += 1
        if t.first_child[v] >= 0:
            c = int(t.first_child[v])
            stack += [c, c + 1] if right_first else [c + 1, c]
    return rank


def test_reference_pop_order_is_right_first_preorder():
    """The reference's LIFO stack (bh.py:506-546) pushes the children in
    index order and pops the right one first: every reference list is
    increasing in right-first preorder rank, and a left-first walk (the
    plausible slip) would order some of them differently."""
    g = np.load(os.path.join(GOLD, "ref_gauss2000.npz"), allow_pickle=False)
    t = O.build(g["pos"], g["mass"], 32)
    rf, lf = _preorder_rank(t, True), _preorder_rank(t, False)
    p, ids = g["L32_t0.5_e0.01_list_ptr"], g["L32_t0.5_e0.01_list_ids"]
    left_sorted = 0
    for i in range(len(p) - 1):
        lst = ids[p[i]:p[i + 1]]
        assert np.all(np.diff(rf[lst]) > 0), i
        left_sorted += bool(np.all(np.diff(lf[lst]) > 0))
    assert left_sorted == 0


def test_subset_self_walk_equals_full_self_walk():
    """O.accelerations_subset (self mode on chosen particles, each skipping
    itself) returns the rows of the full self walk bit for bit, which the
    reference fixtures pin."""
    g = np.load(os.path.join(GOLD, "ref_gauss2000.npz"), allow_pickle=False)
    x, m = g["pos"], g["mass"]
    t = O.build(x, m, 8)
    idx = np.random.default_rng(3).choice(len(x), 150, replace=False)
    for eps in (0.0, 0.01):
        full = O.accelerations(t, x, m, softening=eps, theta=0.5)
        bits_equal(O.accelerations_subset(t, x, m, idx, softening=eps, theta=0.5), full[idx])
    bits_equal(O.accelerations_subset(t, x, m, np.arange(96), softening=0.01, theta=0.5),
               g["L8_t0.5_e0.01_acc_numba"])
