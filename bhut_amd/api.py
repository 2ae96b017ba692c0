"""Public API: accelerations() and Tree (bhut/api.py:254-786), device-backed.

Signatures, defaults, validation order and error messages follow the
reference so code written against bhut runs unchanged; every computation
runs in libbhut_hip on the GPU.  One optional keyword is added:
``precision`` = "fp64" (default; the plain form, matching bhut's float64
walk) or "fp32" (the grouped FP32 relative-coordinate form).
"""

from __future__ import annotations

from typing import Any, Optional, Union

import numpy as np

from . import backends
from .traverse.bh import barnes_hut_accelerations, barnes_hut_accelerations_targets
from .tree.build import build_tree as _build_tree_impl
from .tree.build import device_tree_for
from .tree.node import TreeData
from .tree.refit import refit_tree, should_refit_vs_rebuild

ArrayLike = Any


def _validate_dimensions(dim: int) -> None:
    if dim not in (2, 3):
        raise ValueError(f"dim must be 2 or 3, got {dim}")


def _validate_accelerations_inputs(positions, masses, *, dim, softening, theta, G, leaf_size):
    """api.py:84-152: float64 conversion and checks, same order and messages."""
    _validate_dimensions(dim)
    positions = np.asarray(positions, dtype=np.float64)
    masses = np.asarray(masses, dtype=np.float64)
    if len(positions.shape) != 2:
        raise ValueError(f"positions must be 2D array, got shape {positions.shape}")
    if len(masses.shape) != 1:
        raise ValueError(f"masses must be 1D array, got shape {masses.shape}")
    N, actual_dim = positions.shape
    if actual_dim != dim:
        raise ValueError(f"positions has {actual_dim} dimensions but dim={dim}")
    if masses.shape[0] != N:
        raise ValueError(f"positions has {N} particles but masses has {masses.shape[0]} elements")
    if np.isscalar(softening):
        softening = float(softening)
        if softening < 0:
            raise ValueError(f"softening must be non-negative, got {softening}")
    else:
        softening = np.asarray(softening, dtype=np.float64)
        if len(softening.shape) != 1:
            raise ValueError(f"softening array must be 1D, got shape {softening.shape}")
        if softening.shape[0] != N:
            raise ValueError(f"softening array has {softening.shape[0]} elements but need {N}")
        if np.any(softening < 0):
            raise ValueError("all softening values must be non-negative")
    if theta < 0:
        raise ValueError(f"theta must be non-negative, got {theta}")
    if leaf_size <= 0:
        raise ValueError(f"leaf_size must be positive, got {leaf_size}")
    return positions, masses, softening


def build_tree(positions, masses, *, leaf_size: int, backend: str, dim: int) -> TreeData:
    """api.py:155-191."""
    backends.get_namespace(positions, backend)
    return _build_tree_impl(positions, masses, dim=dim, leaf_size=leaf_size, multipole="mono")


def accelerations(positions: ArrayLike, masses: ArrayLike, *, theta: float = 0.5,
                  softening: Union[float, ArrayLike] = 0.0, G: float = 1.0, dim: int = 3,
                  backend: str = "auto", leaf_size: int = 32, criterion: str = "bh", multipole: str = "mono",
                  precision: str = "fp64") -> np.ndarray:
    """Gravitational accelerations by Barnes-Hut (api.py:254-364)."""
    if criterion != "bh":
        raise ValueError(f"criterion must be 'bh', got '{criterion}'")
    if multipole != "mono":
        raise ValueError(f"multipole must be 'mono', got '{multipole}'")
    backends.get_namespace(positions, backend)
    positions, masses, softening = _validate_accelerations_inputs(
        positions, masses, dim=dim, softening=softening, theta=theta, G=G, leaf_size=leaf_size)
    tree = build_tree(positions, masses, leaf_size=leaf_size, backend=backend, dim=dim)
    return evaluate_accelerations(tree, positions, masses, theta=theta, softening=softening, G=G,
                                  precision=precision)


def evaluate_accelerations(tree: TreeData, positions, masses, theta: float = 0.5,
                           softening: Union[float, ArrayLike] = 0.01, G: float = 1.0, *,
                           precision: str = "fp64") -> np.ndarray:
    """api.py:367-429: softening arrays collapse to a scalar when uniform."""
    softening_array = np.asarray(softening, dtype=np.float64)
    if softening_array.ndim > 1:
        raise ValueError("Softening must be scalar or 1D array")
    if softening_array.ndim == 1 and len(softening_array) != len(positions):
        raise ValueError("Softening array length must match number of particles")
    if np.any(softening_array < 0):
        raise ValueError("Softening must be non-negative")
    if softening_array.ndim == 1:
        if not np.allclose(softening_array, softening_array[0]):
            raise ValueError("Variable softening per particle not yet implemented")
        softening_scalar = float(softening_array[0])
    else:
        softening_scalar = float(softening_array)
    if theta < 0:
        raise ValueError("Opening angle theta must be non-negative")
    positions_np = np.asarray(positions)
    masses_np = np.asarray(masses)
    return barnes_hut_accelerations(tree, positions_np, masses_np, softening_scalar, theta, G, precision=precision)


def potential(positions: ArrayLike, masses: ArrayLike, *, theta: float = 0.5, softening: float = 0.0,
              G: float = 1.0, dim: int = 3, backend: str = "auto", leaf_size: int = 32) -> np.ndarray:
    """Softened gravitational potential phi_i = -G sum_j m_j / sqrt(r_ij^2 + eps^2)
    (self excluded) by the plain Barnes-Hut walk.  bhut's docs call
    `bhut.potential(positions, masses, theta=0.5)` (docs/getting-started.md:36)
    but the package does not define it; the arguments and validation here
    follow accelerations().  Total potential energy = 0.5 * sum(m * phi)."""
    backends.get_namespace(positions, backend)
    positions, masses, softening = _validate_accelerations_inputs(
        positions, masses, dim=dim, softening=softening, theta=theta, G=G, leaf_size=leaf_size)
    if not np.isscalar(softening):
        raise ValueError("Only scalar softening is currently supported")
    tree = build_tree(positions, masses, leaf_size=leaf_size, backend=backend, dim=dim)
    return device_tree_for(tree, positions, masses).potential(None, theta, float(softening), G)


class Tree:
    """Barnes-Hut tree object (api.py:432-786): build, refit, rebuild and
    accelerations, with the tree resident in device memory."""

    def __init__(self, positions: ArrayLike, masses: ArrayLike, *, leaf_size: int = 32, backend: str = "auto",
                 dim: int = 3) -> None:
        _validate_dimensions(dim)
        self.positions, self.masses, _ = _validate_accelerations_inputs(
            positions, masses, dim=dim, softening=0.0, theta=0.5, G=1.0, leaf_size=leaf_size)
        self.dim = dim
        self.leaf_size = leaf_size
        self.backend = backend
        self._tree: Optional[TreeData] = None
        self._is_built = False
        self._xp = backends.get_namespace(self.positions, backend)

    def build(self) -> "Tree":
        self._tree = build_tree(self.positions, self.masses, leaf_size=self.leaf_size, backend=self.backend,
                                dim=self.dim)
        self._is_built = True
        return self

    def refit(self, new_positions: ArrayLike, new_masses: Optional[ArrayLike] = None) -> "Tree":
        """api.py:513-562: refit when every particle stays within 1.1 root
        half sizes of the stored root centre, otherwise rebuild."""
        if not self._is_built or self._tree is None:
            raise ValueError("Tree must be built before refitting. Call build() first.")
        if new_masses is None:
            new_masses = self.masses
        new_positions, new_masses, _ = _validate_accelerations_inputs(
            new_positions, new_masses, dim=self.dim, softening=0.0, theta=0.5, G=1.0, leaf_size=self.leaf_size)
        if not should_refit_vs_rebuild(self._tree, new_positions):
            return self.rebuild(new_positions, new_masses)
        self._tree = refit_tree(self._tree, new_positions, new_masses, self._xp)
        self.positions = new_positions
        self.masses = new_masses
        return self

    def rebuild(self, new_positions: ArrayLike, new_masses: Optional[ArrayLike] = None) -> "Tree":
        if new_masses is None:
            new_masses = self.masses
        new_positions, new_masses, _ = _validate_accelerations_inputs(
            new_positions, new_masses, dim=self.dim, softening=0.0, theta=0.5, G=1.0, leaf_size=self.leaf_size)
        self.positions = new_positions
        self.masses = new_masses
        self._is_built = False
        self._tree = None
        return self.build()

    def accelerations(self, targets: Optional[ArrayLike] = None, *, theta: float = 0.5,
                      softening: Union[float, ArrayLike] = 0.0, G: float = 1.0,
                      precision: str = "fp64", chunk_size: Optional[int] = None) -> np.ndarray:
        """api.py:608-688.  chunk_size streams external targets through the
        device in chunks of that many targets (the counterpart of the
        reference's Dask target chunks, api.py:690-731); the result equals
        the unchunked call bit for bit."""
        if not self._is_built:
            raise ValueError("Tree must be built before evaluating accelerations")
        if targets is None:
            return self.evaluate_accelerations(softening, theta, G, precision=precision)
        targets = np.asarray(targets, dtype=np.float64)
        if len(targets.shape) != 2:
            raise ValueError(f"targets must be 2D array, got shape {targets.shape}")
        if targets.shape[1] != self.dim:
            raise ValueError(f"targets has {targets.shape[1]} dimensions but tree has {self.dim}")
        if not np.isscalar(softening):
            softening = np.asarray(softening, dtype=np.float64)
            M = targets.shape[0]
            if softening.shape != (M,):
                raise ValueError(f"softening array shape {softening.shape} does not match targets shape ({M},)")
        if theta < 0:
            raise ValueError(f"theta must be non-negative, got {theta}")
        return barnes_hut_accelerations_targets(self._tree, self.positions, self.masses, targets, softening, theta,
                                                G, precision=precision, chunk_size=chunk_size)

    def potential(self, targets: Optional[ArrayLike] = None, *, theta: float = 0.5, softening: float = 0.0,
                  G: float = 1.0) -> np.ndarray:
        """Softened potential at the particles (self excluded) or at targets."""
        if not self._is_built:
            raise ValueError("Tree must be built before evaluating accelerations")
        if not isinstance(softening, (int, float)) or softening < 0:
            raise ValueError("Softening must be a non-negative scalar")
        if theta < 0:
            raise ValueError(f"theta must be non-negative, got {theta}")
        if targets is not None:
            targets = np.asarray(targets, dtype=np.float64)
            if len(targets.shape) != 2 or targets.shape[1] != self.dim:
                raise ValueError(f"targets must be (M, {self.dim}), got shape {targets.shape}")
        return device_tree_for(self._tree, self.positions, self.masses).potential(targets, theta, float(softening),
                                                                                  G)

    def evaluate_accelerations(self, softening: Union[float, ArrayLike], theta: float = 0.5, G: float = 1.0, *,
                               precision: str = "fp64") -> np.ndarray:
        """api.py:733-786."""
        if self._tree is None:
            raise RuntimeError("Tree not built. Call build() first.")
        if not isinstance(softening, (int, float)):
            raise ValueError("Only scalar softening is currently supported")
        if softening < 0:
            raise ValueError("Softening must be non-negative")
        if theta < 0:
            raise ValueError("Opening angle theta must be non-negative")
        acc = evaluate_accelerations(self._tree, self.positions, self.masses, theta, softening, G,
                                     precision=precision)
        return np.asarray(acc)
