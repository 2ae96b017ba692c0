"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/bhut_hip.h declares, and the API validates its inputs
with the reference's messages before any device call (bhut/api.py:84-152,
tests/edge_cases/test_boundary_conditions.py of the reference)."""

import os
import re
import subprocess

import numpy as np
import pytest

import bhut_amd
from bhut_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "bhut_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (bh_\w+)", out))
    assert set(syms) <= exported
    # the ctypes table covers the header exactly
    assert set(_native._SIGS) == set(syms)


def declared_arities():
    """Parameter count of every function the header declares."""
    src = open(os.path.join(ROOT, "include", "bhut_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(bh_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_table_matches_header_arity():
    """Every ctypes binding takes as many arguments as its C declaration (a
    mismatch is only noticed when the binding is first called)."""
    ar = declared_arities()
    assert set(ar) == set(_native._SIGS)
    for name, (_, argtypes) in _native._SIGS.items():
        assert len(argtypes) == ar[name], (name, len(argtypes), ar[name])


def test_library_reports_no_device_cleanly():
    if _native.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        _native.Context(0)


def test_validation_messages_before_device():
    # tests/edge_cases/test_boundary_conditions.py:17-56 of the reference
    with pytest.raises(ValueError, match="2D array|dimension"):
        bhut_amd.accelerations(np.array([[1, 2]]), np.array([1]))
    with pytest.raises(ValueError, match="2D array|dimension"):
        bhut_amd.accelerations(np.array([1, 2, 3]), np.array([1]))
    with pytest.raises(ValueError, match="1D array|dimension"):
        bhut_amd.accelerations(np.array([[1, 2, 3]]), np.array([[1]]))
    with pytest.raises(ValueError, match="particles|elements|has"):
        bhut_amd.accelerations(np.random.randn(5, 3), np.random.randn(3))
    with pytest.raises(ValueError, match="non-negative"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), softening=-1.0)
    with pytest.raises(ValueError, match="theta"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), theta=-1.0)
    with pytest.raises(ValueError, match="leaf_size|positive"):
        bhut_amd.Tree(np.zeros((2, 3)), np.ones(2), leaf_size=0)
    with pytest.raises(ValueError, match="leaf_size|positive"):
        bhut_amd.Tree(np.zeros((2, 3)), np.ones(2), leaf_size=-5)
    with pytest.raises((ValueError, TypeError)):
        bhut_amd.Tree(np.zeros((2, 3)), np.ones(2), leaf_size="invalid")
    with pytest.raises(ValueError, match="criterion"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), criterion="fmm")
    with pytest.raises(ValueError, match="multipole"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), multipole="quad")
    with pytest.raises(ValueError, match="backend"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), backend="cupy")
    with pytest.raises(ImportError, match="[Dd]ask"):
        bhut_amd.accelerations(np.zeros((2, 3)), np.ones(2), backend="dask")
    with pytest.raises(ValueError, match="dim"):
        bhut_amd.Tree(np.zeros((2, 3)), np.ones(2), dim=4)


def test_empty_input_is_rejected():
    # "empty|zero" (test_boundary_conditions.py:17-24); the build raises
    # before touching a device
    from bhut_amd.tree.build import build_tree
    with pytest.raises(ValueError, match="empty|zero"):
        build_tree(np.empty((0, 3)), np.empty(0), dim=3, leaf_size=32)


def test_tree_not_built_errors():
    t = bhut_amd.Tree(np.zeros((2, 3)), np.ones(2))
    with pytest.raises(ValueError, match="must be built"):
        t.accelerations()
    with pytest.raises(ValueError, match="must be built"):
        t.refit(np.zeros((2, 3)))
    with pytest.raises(RuntimeError, match="not built"):
        t.evaluate_accelerations(0.01)


def test_treedata_host_arrays_validate():
    from bhut_amd.tree.node import TreeData
    from oracle import oracle as O
    rng = np.random.default_rng(0)
    x = rng.standard_normal((100, 3))
    o = O.build(x, np.ones(100), 8)
    td = TreeData(**{k: getattr(o, k) for k in ("center", "half_size", "mass", "com", "first_child", "child_count",
                                                 "start", "count", "is_leaf", "perm")}, dim=3, leaf_size=8)
    td.validate()
    assert td.num_nodes == o.num_nodes and td.num_particles == 100
    np.testing.assert_array_equal(td.get_children(0), [o.first_child[0], o.first_child[0] + 1])
    assert sorted(td.get_particles(0)) == list(range(100))
    td.count = td.count.copy()
    td.count[0] = 5
    with pytest.raises(ValueError, match="Root node count"):
        td.validate()
