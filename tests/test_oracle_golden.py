"""The C oracle against the golden fixtures (CPU only).

The fixtures were computed with the NumPy/OpenBLAS/libm calls the reference
itself makes (tests/golden/make_golden.py), so these tests pin the oracle's
restatement of np.sum order, the fma chain of np.linalg.norm, libm pow and
the stable argsort -- bit for bit.
"""

import os

import numpy as np
import pytest

from conftest import golden_files
from oracle import oracle as O

FIELDS = ["center", "half_size", "mass", "com", "first_child", "child_count", "start", "count",
          "is_leaf", "perm"]


@pytest.fixture(params=golden_files(), ids=lambda p: os.path.basename(p)[7:-4])
def golden(request):
    return np.load(request.param, allow_pickle=False)


def test_morton_codes(golden):
    g = golden
    leaf = int(g["leaves"][0])
    np.testing.assert_array_equal(O.morton_codes(g["pos"]), g[f"L{leaf}_codes"])


def test_stable_argsort_matches_numpy(golden):
    g = golden
    leaf = int(g["leaves"][0])
    codes = g[f"L{leaf}_codes"]
    np.testing.assert_array_equal(O.argsort_stable(codes), np.argsort(codes, kind="stable"))


def test_build_bit_exact(golden):
    g = golden
    for leaf in g["leaves"]:
        t = O.build(g["pos"], g["mass"], int(leaf))
        for f in FIELDS:
            exp = g[f"L{leaf}_{f}"]
            got = getattr(t, f)
            assert got.shape == exp.shape, f
            # bitwise equality (also distinguishes -0.0 / NaN payloads)
            if got.dtype == np.float64:
                np.testing.assert_array_equal(got.view(np.uint64), exp.view(np.uint64), err_msg=f)
            else:
                np.testing.assert_array_equal(got, exp, err_msg=f)


def test_refit_bit_exact(golden):
    g = golden
    for leaf in g["leaves"]:
        t = O.build(g["pos"], g["mass"], int(leaf))
        r = O.refit(t, g["pos2"], g["mass"])
        np.testing.assert_array_equal(r.mass.view(np.uint64), g[f"L{leaf}_refit_mass"].view(np.uint64))
        np.testing.assert_array_equal(r.com.view(np.uint64), g[f"L{leaf}_refit_com"].view(np.uint64))
        # geometry untouched by refit (refit.py:84-97)
        np.testing.assert_array_equal(r.center, t.center)
        np.testing.assert_array_equal(r.half_size, t.half_size)


@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_accelerations_and_lists_bit_exact(golden, eps):
    g = golden
    pos, mass = g["pos"], g["mass"]
    for leaf in g["leaves"]:
        t = O.build(pos, mass, int(leaf))
        for theta in g["thetas"]:
            tag = f"L{leaf}_t{theta}_e{eps}"
            exp_n = g[f"{tag}_acc_numba"]
            nt = len(exp_n)
            got_n = O.accelerations(t, pos, mass, softening=eps, theta=theta, leaf_formula=O.LEAF_NUMBA)[:nt]
            got_p = O.accelerations(t, pos, mass, softening=eps, theta=theta, leaf_formula=O.LEAF_PYTHON)[:nt]
            np.testing.assert_array_equal(got_n, exp_n, err_msg=tag)
            np.testing.assert_array_equal(got_p, g[f"{tag}_acc_python"], err_msg=tag)
            ptr, ids = g[f"{tag}_list_ptr"], g[f"{tag}_list_ids"]
            for i in range(nt):
                np.testing.assert_array_equal(O.target_list(t, pos[i], theta), ids[ptr[i]:ptr[i + 1]])
            ext = g["ext"]
            got_e = O.accelerations(t, pos, mass, ext, softening=eps, theta=theta)
            np.testing.assert_array_equal(got_e, g[f"{tag}_ext_acc"], err_msg=tag)
            ptr, ids = g[f"{tag}_ext_list_ptr"], g[f"{tag}_ext_list_ids"]
            for i in range(len(ext)):
                np.testing.assert_array_equal(O.target_list(t, ext[i], theta), ids[ptr[i]:ptr[i + 1]])


def test_np_sum_restatement_against_numpy():
    """orc_np_sum against np.sum itself on many lengths (pairwise blocks,
    the 128 threshold, the recursive split, the reduction's buffers of 8192
    elements) and magnitudes."""
    rng = np.random.default_rng(5)
    for n in list(range(0, 300)) + [511, 1000, 4097, 8191, 8192, 8193, 9000, 10000, 12345, 16385, 20000, 24577,
                                    50000, 250001]:
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, size=n)
        assert O.np_sum(a) == np.sum(a), n


def test_aabb_nan_propagates_like_numpy():
    p = np.array([[1.0, 2.0, 3.0], [np.nan, 0.0, 5.0], [0.5, -1.0, 4.0]])
    lo, hi = O.aabb(p)
    np.testing.assert_array_equal(lo, np.min(p, axis=0))
    np.testing.assert_array_equal(hi, np.max(p, axis=0))


def test_theta_zero_is_direct_sum():
    """bh.py:600-601: theta=0 opens every internal node, so the walk is the
    exact direct sum (the config-1 oracle)."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((500, 3))
    m = rng.uniform(0.5, 1.5, 500)
    t = O.build(x, m, 8)
    a = O.accelerations(t, x, m, softening=0.01, theta=0.0)
    d = O.direct(x, m, softening=0.01)
    np.testing.assert_allclose(a, d, rtol=1e-12, atol=1e-12 * np.abs(d).max())


def test_two_body_closed_form():
    """|a| = G m / r^2 (reference tests/validation/test_physics_validation.py:17-48)."""
    x = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    m = np.array([3.0, 5.0])
    t = O.build(x, m, 1)
    a = O.accelerations(t, x, m, softening=0.0, theta=0.5, G=2.0)
    np.testing.assert_allclose(a[0], [2.0 * 5.0 / 4.0, 0, 0], rtol=1e-15)
    np.testing.assert_allclose(a[1], [-2.0 * 3.0 / 4.0, 0, 0], rtol=1e-15)


def test_planted_mistakes_fail_the_fixtures():
    """Plausible slips in the oracle's restatement -- leaf masses summed left
    to right instead of in NumPy's pairwise order, or the opening distance
    without the fma chain -- change bits the fixtures hold, so the bit-exact
    tests above have teeth."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_gauss2000.npz"))
    sm = g["mass"][g["L32_perm"]]
    leaves = np.nonzero(g["L32_is_leaf"])[0]
    naive_differs = 0
    for k in leaves:
        s, c = g["L32_start"][k], g["L32_count"][k]
        acc = 0.0
        for v in sm[s:s + c]:
            acc += v
        naive_differs += acc != g["L32_mass"][k]
    assert naive_differs > 0
    # plain sqrt((x*x + y*y) + z*z) flips at least one opening decision
    # somewhere, or at least one distance bit
    com, pos = g["L32_com"], g["pos"]
    v = pos[:96, None, :] - com[None, :, :]
    plain = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
    # the reference calls norm on one 1-D vector at a time (bh.py:593), which
    # goes through ddot; the axis= form would take a different path
    lib = np.array([[np.linalg.norm(v[i, j]) for j in range(v.shape[1])] for i in range(v.shape[0])])
    assert np.any(plain != lib)
