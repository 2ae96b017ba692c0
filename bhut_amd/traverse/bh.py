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
                                     theta: float = 0.5, G: float = 1.0, *, precision: str = "fp64",
                                     chunk_size: int | None = None) -> np.ndarray:
    """Accelerations at external targets, no self exclusion (bh.py:30-80).
    With chunk_size the targets stream through the device in chunks on two
    streams (the Dask target chunks of api.py:28-78 and dask_.py:173-228),
    bit-identical to the unchunked call; device memory then holds two chunks
    of targets whatever their number."""
    validate_kernel_inputs(tree_positions, tree_masses, softening)
    eps = _scalar_softening(softening)
    tgt = np.asarray(target_positions, dtype=np.float64)
    if tgt.ndim != 2 or tgt.shape[1] != tree_positions.shape[1]:
        raise ValueError(f"target_positions must have shape (M, {tree_positions.shape[1]}), got {tgt.shape}")
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")
    dt = device_tree_for(tree, tree_positions, tree_masses)
    if chunk_size is not None:
        if int(chunk_size) <= 0:
            raise ValueError(f"chunk_size must be positive, got {chunk_size}")
        return dt.accel_targets_streamed(tgt, theta, eps, G, precision=precision, chunk=int(chunk_size))
    if precision == "fp64":
        return dt.accel_targets(tgt, theta, eps, G)
    return dt.accel_f32(tgt, theta, eps, G)


def interaction_lists(tree: TreeData, positions, masses, theta: float = 0.5, targets=None):
    """Per-target interaction lists of the plain walk, in pop order: returns
    (ptr, ids) with the node ids of target i at ids[ptr[i]:ptr[i+1]]."""
    dt = device_tree_for(tree, positions, masses)
    return dt.interaction_lists(None if targets is None else np.asarray(targets, np.float64), theta)


def _counting_tree(tree: TreeData):
    """A device tree whose node records are `tree`'s.  The counters read only
    node records (mass, centre of mass, size, children, particle ranges), so a
    host-only TreeData is imported with zero source arrays."""
    dt = tree.device_tree
    if dt is not None:
        return dt
    from .. import _native
    n = tree.num_particles
    return _native.DeviceTree.from_arrays(tree, np.zeros((n, tree.dim)), np.zeros(n))


def traversal_counters(tree: TreeData, positions, theta: float = 0.5, *, self_mode: bool = False) -> np.ndarray:
    """Exact per-target counts of the walk, (m, 5) int64: nodes popped,
    opening tests, monopoles accepted, leaves visited, particle pairs summed.
    With self_mode, positions must be the tree's own particles in original
    order and target i skips source i (bh.py:253-255); otherwise nothing is
    skipped (bh.py:196-221)."""
    positions = np.asarray(positions, dtype=np.float64)
    if not self_mode:
        return _counting_tree(tree).counters(positions, theta)
    dt = tree.device_tree
    if dt is None or getattr(dt, "src_pos", None) is not positions:
        from .. import _native
        dt = _native.DeviceTree.from_arrays(tree, positions, np.zeros(len(positions)))
    return dt.counters(None, theta)


def compute_tree_statistics(tree: TreeData, positions, theta: float = 0.5) -> Dict[str, Any]:
    """compute_tree_statistics (bh.py:604-646) with the reference's values.

    As in _count_operations_single_target (bh.py:648-700), the walk runs at
    each passed position (the index is not used): 'avg_nodes_visited' counts
    every pop, zero-mass nodes included; 'avg_particles_computed' adds
    count - 1 for every visited leaf (bh.py:678-681) whether or not the
    position is one of the leaf's particles; 'total_operations' is particles
    plus monopoles.  The exact counts of the same walks are added under keys
    of their own: 'avg_opening_tests', 'avg_leaves_visited',
    'total_pairs_exact' (the particles a target at that position sums, no
    self exclusion) and 'per_target' ((m, 5) int64, see traversal_counters).
    """
    positions = np.asarray(positions, dtype=np.float64)
    n = len(positions)
    if n == 0:
        raise ZeroDivisionError("division by zero")  # total / n_particles with no positions (bh.py:641)
    c = _counting_tree(tree).counters(positions, theta)
    tot = [int(v) for v in c.sum(axis=0)]
    particles = tot[4] - tot[3]  # sum over visited leaves of (count - 1)
    monopoles = tot[2]
    return {
        "avg_nodes_visited": tot[0] / n,
        "avg_particles_computed": particles / n,
        "avg_monopoles_used": monopoles / n,
        "total_operations": particles + monopoles,
        "avg_opening_tests": tot[1] / n,
        "avg_leaves_visited": tot[3] / n,
        "total_pairs_exact": tot[4],
        "per_target": c,
    }
