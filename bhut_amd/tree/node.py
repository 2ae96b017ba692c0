"""TreeData: the Structure-of-Arrays tree of bhut/tree/node.py:15-181.

Same fields, dtypes (int64 indices, float64 geometry, bool leaf flags) and
methods as the reference.  A TreeData produced by this package is backed by a
tree resident in device memory; its host arrays are copied out on first
access.  Assigning a field detaches it from the device tree, and the next
device use imports the host arrays instead.
"""

from __future__ import annotations

from typing import Any

import numpy as np

FIELDS = ("center", "half_size", "mass", "com", "first_child", "child_count", "start", "count", "is_leaf",
          "perm")


class TreeData:
    """Structure-of-Arrays Barnes-Hut tree (bhut/tree/node.py:15-77).

    Parameters follow the reference: center (M, D), half_size (M,), mass (M,),
    com (M, D), first_child (M,) (-1 for leaves), child_count (M,), start (M,),
    count (M,), is_leaf (M,), perm (N,) (sorted position -> original index),
    dim, leaf_size.
    """

    def __init__(self, center=None, half_size=None, mass=None, com=None, first_child=None, child_count=None,
                 start=None, count=None, is_leaf=None, perm=None, dim: int = 3, leaf_size: int = 32, *,
                 _device=None):
        object.__setattr__(self, "_device", _device)
        object.__setattr__(self, "_host", None)
        object.__setattr__(self, "dim", int(dim))
        object.__setattr__(self, "leaf_size", int(leaf_size))
        given = dict(center=center, half_size=half_size, mass=mass, com=com, first_child=first_child,
                     child_count=child_count, start=start, count=count, is_leaf=is_leaf, perm=perm)
        if _device is None:
            missing = [k for k, v in given.items() if v is None]
            if missing:
                raise TypeError(f"TreeData missing fields: {', '.join(missing)}")
            object.__setattr__(self, "_host", dict(given))

    # -- device backing ---------------------------------------------------
    @classmethod
    def from_device(cls, dtree) -> "TreeData":
        return cls(dim=dtree.dim, leaf_size=dtree.leaf_size, _device=dtree)

    @property
    def device_tree(self):
        """The device tree backing these arrays, or None once detached."""
        return self._device

    def _materialize(self) -> dict:
        if self._host is None:
            object.__setattr__(self, "_host", self._device.export())
        return self._host

    def __getattr__(self, name: str) -> Any:
        if name in FIELDS:
            return self._materialize()[name]
        raise AttributeError(name)

    def __setattr__(self, name: str, value: Any) -> None:
        if name in FIELDS:
            host = self._materialize()
            host[name] = value
            object.__setattr__(self, "_device", None)  # host arrays now authoritative
        else:
            object.__setattr__(self, name, value)

    def __repr__(self) -> str:
        where = "device" if self._device is not None else "host"
        return f"TreeData(num_nodes={self.num_nodes}, num_particles={self.num_particles}, dim={self.dim}, " \
               f"leaf_size={self.leaf_size}, {where})"

    # -- reference API ------------------------------------------------------
    @property
    def num_nodes(self) -> int:
        if self._device is not None and self._host is None:
            return self._device.num_nodes
        return len(self.center)

    @property
    def num_particles(self) -> int:
        if self._device is not None and self._host is None:
            return self._device.n
        return len(self.perm)

    def validate(self) -> None:
        """Consistency checks of bhut/tree/node.py:89-137, same messages."""
        M = self.num_nodes
        N = self.num_particles
        D = self.dim
        if D not in (2, 3):
            raise ValueError(f"dim must be 2 or 3, got {D}")
        arrays_M = [("center", self.center, (M, D)), ("half_size", self.half_size, (M,)),
                    ("mass", self.mass, (M,)), ("com", self.com, (M, D)),
                    ("first_child", self.first_child, (M,)), ("child_count", self.child_count, (M,)),
                    ("start", self.start, (M,)), ("count", self.count, (M,)), ("is_leaf", self.is_leaf, (M,))]
        for name, arr, expected in arrays_M:
            if arr.shape != expected:
                raise ValueError(f"{name} has shape {arr.shape}, expected {expected}")
        if self.perm.shape != (N,):
            raise ValueError(f"perm has shape {self.perm.shape}, expected ({N},)")
        if M > 0:
            if self.count[0] != N:
                raise ValueError(f"Root node count {self.count[0]} != num_particles {N}")
            if not np.array_equal(np.sort(self.perm), np.arange(N)):
                raise ValueError("perm is not a valid permutation")
            valid = self.first_child >= 0
            if np.any(self.first_child[valid] >= M):
                raise ValueError("first_child indices exceed number of nodes")
            if np.any(self.count[self.is_leaf] > self.leaf_size):
                raise ValueError("Some leaf nodes exceed leaf_size")

    def get_children(self, node_idx: int) -> np.ndarray:
        if self.is_leaf[node_idx]:
            return np.array([], dtype=np.intp)
        first = self.first_child[node_idx]
        return np.arange(first, first + self.child_count[node_idx], dtype=np.intp)

    def get_particles(self, node_idx: int) -> np.ndarray:
        start = self.start[node_idx]
        return self.perm[np.arange(start, start + self.count[node_idx], dtype=np.intp)]

    def arrays(self) -> dict:
        """All fields as a dict of host arrays."""
        return dict(self._materialize())
