"""Batched form: many independent systems in one forest.

bhut has no batched entry point; a user of the reference loops
``bhut.accelerations`` (api.py:254-364) over the systems.  Here the systems
are stored back to back, sorted once on (system id, Morton code), built as
one forest (every system gets exactly the tree ``build_tree`` gives it alone,
bhut/tree/build.py:19-107) and walked by one launch in which each target only
sees its own system.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from ._native import DeviceTree, F32Opts


def _stack(systems):
    pos = [np.asarray(p, dtype=np.float64) for p, _ in systems]
    mas = [np.asarray(m, dtype=np.float64) for _, m in systems]
    if not pos:
        raise ValueError("need at least one system")
    dim = pos[0].shape[1] if pos[0].ndim == 2 else -1
    for i, (p, m) in enumerate(zip(pos, mas)):
        if p.ndim != 2 or p.shape[1] != dim:
            raise ValueError(f"system {i}: positions must be (n, {dim}), got shape {p.shape}")
        if m.shape != (p.shape[0],):
            raise ValueError(f"system {i}: positions has {p.shape[0]} particles but masses has shape {m.shape}")
        if p.shape[0] == 0:
            raise ValueError(f"system {i}: Cannot build tree with zero particles")
    counts = np.array([p.shape[0] for p in pos], dtype=np.int64)
    return np.concatenate(pos), np.concatenate(mas), counts


class Forest:
    """A forest of independent systems resident on the device.

    ``positions`` (n, dim) and ``masses`` (n,) hold the systems back to back,
    ``counts[s]`` particles each.
    """

    def __init__(self, positions, masses, counts, *, leaf_size: int = 32):
        self.positions = np.ascontiguousarray(positions, dtype=np.float64)
        self.masses = np.ascontiguousarray(masses, dtype=np.float64)
        self.counts = np.ascontiguousarray(counts, dtype=np.int64)
        if leaf_size <= 0:
            raise ValueError(f"leaf_size must be positive, got {leaf_size}")
        if self.positions.ndim != 2 or self.positions.shape[1] not in (2, 3):
            raise ValueError(f"positions must be (n, 2) or (n, 3), got shape {self.positions.shape}")
        if np.any(self.counts <= 0):
            raise ValueError("Cannot build tree with zero particles")
        self.leaf_size = int(leaf_size)
        self.device_tree = DeviceTree.build_forest(self.positions, self.masses, self.counts, self.leaf_size)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)])

    @classmethod
    def from_systems(cls, systems: Sequence, *, leaf_size: int = 32) -> "Forest":
        pos, mas, counts = _stack(systems)
        return cls(pos, mas, counts, leaf_size=leaf_size)

    @property
    def num_systems(self) -> int:
        return len(self.counts)

    def refit(self, positions, masses=None) -> "Forest":
        """New masses and centres of mass of every system's nodes (refit.py:17-109)."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        if pos.shape != self.positions.shape:
            raise ValueError(f"positions shape {pos.shape} does not match {self.positions.shape}")
        m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
        self.device_tree.refit(pos, m)
        self.positions = pos
        if m is not None:
            self.masses = m
        return self

    def accelerations(self, *, theta: float = 0.5, softening: float = 0.0, G: float = 1.0,
                      precision: str = "fp64", opts: F32Opts | None = None) -> np.ndarray:
        """(n, dim) self accelerations; system s only feels system s."""
        if theta < 0:
            raise ValueError("Opening angle theta must be non-negative")
        if softening < 0:
            raise ValueError("Softening must be non-negative")
        if precision == "fp64":
            return self.device_tree.accel_self(theta, softening, G)
        if precision == "fp32":
            return self.device_tree.accel_f32(None, theta, softening, G, opts)
        raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")

    def split(self, arr: np.ndarray) -> list:
        """Per-system views of an (n, ...) array."""
        return [arr[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]


def accelerations_batched(systems: Sequence, *, theta: float = 0.5, softening: float = 0.0, G: float = 1.0,
                          leaf_size: int = 32, precision: str = "fp64") -> list:
    """Self accelerations of each (positions, masses) system, as a list; equal
    to calling accelerations() on each system alone (api.py:254-364)."""
    f = Forest.from_systems(systems, leaf_size=leaf_size)
    return f.split(f.accelerations(theta=theta, softening=softening, G=G, precision=precision))
