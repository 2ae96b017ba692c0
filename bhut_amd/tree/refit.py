"""refit_tree / should_refit_vs_rebuild (bhut/tree/refit.py:17-251) on the device."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from .node import TreeData


def refit_tree(tree: TreeData, new_positions, masses, xp: Any = None) -> TreeData:
    """New TreeData with mass and centre of mass recomputed from
    ``new_positions`` (original order); centres and half sizes are kept
    (refit.py:84-109).  Same shape checks and messages as the reference."""
    new_positions = np.asarray(new_positions, dtype=np.float64)
    masses = np.asarray(masses, dtype=np.float64)
    n = tree.num_particles
    if new_positions.shape[0] != n:
        raise ValueError(f"new_positions has {new_positions.shape[0]} particles "
                         f"but tree was built with {n} particles")
    if masses.shape[0] != n:
        raise ValueError(f"masses has {masses.shape[0]} elements but tree was built with {n} particles")
    if new_positions.shape[1] != tree.dim:
        raise ValueError(f"new_positions has {new_positions.shape[1]} dimensions "
                         f"but tree has {tree.dim} dimensions")
    if tree.device_tree is not None:
        dt = tree.device_tree.clone()
    else:
        dt = _native.DeviceTree.from_arrays(tree, new_positions, masses)
    dt.refit(new_positions, masses)
    dt.src_pos, dt.src_mass = new_positions, masses
    return TreeData.from_device(dt)


def should_refit_vs_rebuild(tree: TreeData, new_positions, position_tolerance: float = 0.1,
                            boundary_check: bool = True) -> bool:
    """True when refit is recommended (refit.py:195-251): same N and dim and,
    with boundary_check, no particle beyond 1.1 root half sizes of the
    stored root centre."""
    new_positions = np.asarray(new_positions, dtype=np.float64)
    if new_positions.shape[0] != tree.num_particles or new_positions.shape[1] != tree.dim:
        return False
    if not boundary_check:
        return True
    dt = tree.device_tree
    if dt is None:
        dt = _native.DeviceTree.from_arrays(tree, new_positions, np.zeros(len(new_positions)))
    return dt.should_refit(new_positions)
