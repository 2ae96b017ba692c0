"""Barnes-Hut traversal entry points (bhut/traverse/bh.py) on the device.

``barnes_hut_accelerations`` and ``barnes_hut_accelerations_targets`` keep
the reference signatures.  The walk, the opening test and the leaf sums run
in libbhut_hip; ``precision="fp32"`` selects the grouped FP32 form instead of
the plain FP64 form.  ``HAVE_NUMBA`` is kept as a name (README.md:148 of the
reference) and is False: there is no Numba here, the device kernels use the
Numba loop's formula (bh.py:249-278).
"""

from __future__ import annotations

from typing import Any, Dict

import numpy as np

from ..tree.build import device_tree_for
from ..tree.node import TreeData

HAVE_NUMBA = False


def validate_kernel_inputs(positions, masses, softening) -> None:
    """bhut/traverse/kernels.py:171-197, same messages."""
    if len(positions.shape) != 2:
        raise ValueError(f"positions must be 2D array, got shape {positions.shape}")
    if len(masses.shape) != 1:
        raise ValueError(f"masses must be 1D array, got shape {masses.shape}")
    if positions.shape[0] != masses.shape[0]:
        raise ValueError(f"positions has {positions.shape[0]} particles but masses has {masses.shape[0]} elements")
    if not np.isscalar(softening):
        softening = np.asarray(softening)
        if len(softening.shape) != 1:
            raise ValueError(f"softening must be scalar or 1D array, got shape {softening.shape}")
        if softening.shape[0] != positions.shape[0]:
            raise ValueError(f"softening array has {softening.shape[0]} elements but need {positions.shape[0]}")


def _scalar_softening(softening) -> float:
    if np.isscalar(softening):
        return float(softening)
    s = np.asarray(softening, dtype=np.float64)
    if s.size and np.all(s == s.flat[0]):
        return float(s.flat[0])
    raise ValueError("Only scalar softening is currently supported")


def barnes_hut_accelerations(tree: TreeData, positions, masses, softening, theta: float = 0.5, G: float = 1.0,
                             *, precision: str = "fp64") -> np.ndarray:
    """Self accelerations in original particle order (bh.py:428-482)."""
    validate_kernel_inputs(positions, masses, softening)
    eps = _scalar_softening(softening)
    dt = device_tree_for(tree, positions, masses)
    if precision == "fp64":
        return dt.accel_self(theta, eps, G)
    if precision == "fp32":
        return dt.accel_f32(None, theta, eps, G)
    raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")


def barnes_hut_accelerations_targets(tree: TreeData, tree_positions, tree_masses, target_positions, softening,
                                     theta: float = 0.5, G: float = 1.0, *, precision: str = "fp64") -> np.ndarray:
    """Accelerations at external targets, no self exclusion (bh.py:30-80)."""
    validate_kernel_inputs(tree_positions, tree_masses, softening)
    eps = _scalar_softening(softening)
    tgt = np.asarray(target_positions, dtype=np.float64)
    dt = device_tree_for(tree, tree_positions, tree_masses)
    if precision == "fp64":
        return dt.accel_targets(tgt, theta, eps, G)
    if precision == "fp32":
        return dt.accel_f32(tgt, theta, eps, G)
    raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")


def interaction_lists(tree: TreeData, positions, masses, theta: float = 0.5, targets=None):
    """Per-target interaction lists of the plain walk, in pop order: returns
    (ptr, ids) with the node ids of target i at ids[ptr[i]:ptr[i+1]]."""
    dt = device_tree_for(tree, positions, masses)
    return dt.interaction_lists(None if targets is None else np.asarray(targets, np.float64), theta)


def compute_tree_statistics(tree: TreeData, positions, theta: float = 0.5, masses=None) -> Dict[str, Any]:
    """Traversal statistics with exact device counters.  Keys of the
    reference (bh.py:604-646) plus exact extras; 'avg_particles_computed'
    counts the pairs actually summed (the reference counts count-1 per leaf
    whether or not the target lies in it, bh.py:678-681)."""
    positions = np.asarray(positions, dtype=np.float64)
    if masses is None:
        dt = tree.device_tree
        if dt is None:
            raise ValueError("masses are needed for a tree without a device copy")
    else:
        dt = device_tree_for(tree, positions, np.asarray(masses, dtype=np.float64))
    c = dt.counters(None, theta) if len(positions) == dt.n else dt.counters(positions, theta)
    n = len(positions)
    tot = c.sum(axis=0)
    return {
        "avg_nodes_visited": tot[0] / n,
        "avg_particles_computed": tot[4] / n,
        "avg_monopoles_used": tot[2] / n,
        "total_operations": int(tot[4] + tot[2]),
        "avg_opening_tests": tot[1] / n,
        "avg_leaves_visited": tot[3] / n,
        "per_target": c,
    }
