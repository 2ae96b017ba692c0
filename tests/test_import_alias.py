"""`import bhut` resolves to bhut_amd's modules (the reference's import names,
bhut/__init__.py:8-11), without touching a device."""

import importlib
import sys

import pytest


def test_bhut_names_resolve_to_bhut_amd():
    sys.modules.pop("bhut", None)
    bhut = importlib.import_module("bhut")
    import bhut_amd
    assert bhut.Tree is bhut_amd.Tree and bhut.accelerations is bhut_amd.accelerations
    assert bhut.__all__ == ["Tree", "accelerations"]
    from bhut.traverse.bh import HAVE_NUMBA, barnes_hut_accelerations, compute_tree_statistics
    from bhut.tree.build import build_tree
    from bhut.tree.node import TreeData
    from bhut.tree.refit import refit_tree, should_refit_vs_rebuild
    import bhut_amd.traverse.bh as B
    import bhut_amd.tree.build as TB
    assert barnes_hut_accelerations is B.barnes_hut_accelerations
    assert compute_tree_statistics is B.compute_tree_statistics and HAVE_NUMBA is False
    assert build_tree is TB.build_tree and TreeData is bhut_amd.tree.node.TreeData
    assert refit_tree is bhut_amd.tree.refit.refit_tree
    assert should_refit_vs_rebuild is bhut_amd.tree.refit.should_refit_vs_rebuild
    import bhut.api
    assert bhut.api.Tree is bhut_amd.Tree


def test_validation_errors_through_the_alias():
    """Host-side validation runs before any device call (bhut/api.py:90-152)."""
    import numpy as np
    import bhut
    with pytest.raises(ValueError, match="2D array"):
        bhut.accelerations(np.zeros(3), np.ones(3))
    with pytest.raises(ValueError, match="non-negative"):
        bhut.accelerations(np.zeros((3, 3)), np.ones(3), theta=-1.0)
