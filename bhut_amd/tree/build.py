"""build_tree (bhut/tree/build.py:19-107) on the device.

bounding box -> Morton codes -> stable radix sort -> median-split topology ->
tight node boxes -> mass / centre of mass, all in libbhut_hip; the result is a
TreeData backed by the device tree.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from .node import TreeData


def build_tree(positions, masses, *, dim: int, leaf_size: int, xp: Any = None, multipole: str = "mono") -> TreeData:
    """Same signature and errors as bhut.tree.build.build_tree; ``xp`` is
    accepted for compatibility (the arrays are NumPy carriers)."""
    if multipole != "mono":
        raise ValueError(f"multipole must be 'mono', got '{multipole}'")
    positions_np = np.asarray(positions, dtype=np.float64)
    masses_np = np.asarray(masses, dtype=np.float64)
    N = positions_np.shape[0]
    if N == 0:
        raise ValueError("Cannot build tree with zero particles")
    if positions_np.ndim != 2 or positions_np.shape[1] != dim:
        raise ValueError(f"positions must be 2D array with {dim} columns, got shape {positions_np.shape}")
    if masses_np.shape != (N,):
        raise ValueError(f"positions has {N} particles but masses has {masses_np.shape[0]} elements")
    dtree = _native.DeviceTree.build(positions_np, masses_np, leaf_size)
    dtree.src_pos, dtree.src_mass = positions_np, masses_np
    return TreeData.from_device(dtree)


def device_tree_for(tree: TreeData, positions: np.ndarray, masses: np.ndarray):
    """A device tree that walks `tree` with (positions, masses) as sources
    (original order): the tree's own device copy when it was built or refit
    from these very arrays, else a fresh import of the host arrays."""
    dt = tree.device_tree
    if dt is not None and getattr(dt, "src_pos", None) is positions and getattr(dt, "src_mass", None) is masses:
        return dt
    return _native.DeviceTree.from_arrays(tree, positions, masses)
