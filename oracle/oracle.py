"""ctypes front end of the CPU oracle (oracle/bhut_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` as the checker or the timed CPU
baseline.  The product package ``bhut_amd`` never imports this module.

Every function returns plain NumPy arrays in the reference's layout
(``bhut/tree/node.py:59-77``): int64 indices, float64 geometry.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

LEAF_NUMBA = 0
LEAF_PYTHON = 1


class _OrcTree(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int64),
        ("N", ctypes.c_int64),
        ("dim", ctypes.c_int64),
        ("leaf_size", ctypes.c_int64),
        ("center", ctypes.c_void_p),
        ("half_size", ctypes.c_void_p),
        ("mass", ctypes.c_void_p),
        ("com", ctypes.c_void_p),
        ("first_child", ctypes.c_void_p),
        ("child_count", ctypes.c_void_p),
        ("start", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("is_leaf", ctypes.c_void_p),
        ("perm", ctypes.c_void_p),
    ]


def build_library() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_library()
        L = ctypes.CDLL(_LIB_PATH)
        i64, dbl, vp, ci = ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_int
        L.orc_np_sum.restype = dbl
        L.orc_np_sum.argtypes = [vp, i64]
        L.orc_aabb.argtypes = [vp, i64, ci, vp, vp]
        L.orc_normalize.argtypes = [vp, i64, ci, vp, vp, vp]
        L.orc_morton.argtypes = [vp, i64, ci, vp]
        L.orc_argsort_stable.argtypes = [vp, i64, vp]
        L.orc_max_nodes.restype = i64
        L.orc_max_nodes.argtypes = [i64, i64]
        L.orc_topology.restype = i64
        L.orc_topology.argtypes = [ctypes.POINTER(_OrcTree)]
        L.orc_build.restype = i64
        L.orc_build.argtypes = [vp, vp, i64, ci, i64, ctypes.POINTER(_OrcTree)]
        L.orc_refit.argtypes = [ctypes.POINTER(_OrcTree), vp, vp]
        L.orc_should_refit.restype = ci
        L.orc_should_refit.argtypes = [ctypes.POINTER(_OrcTree), vp, i64, ci]
        L.orc_accel.argtypes = [ctypes.POINTER(_OrcTree), vp, vp, vp, i64, ci, dbl, dbl, dbl, ci, ci, vp]
        L.orc_accel_subset.argtypes = [ctypes.POINTER(_OrcTree), vp, vp, vp, i64, dbl, dbl, dbl, ci, ci, vp]
        L.orc_potential.argtypes = [ctypes.POINTER(_OrcTree), vp, vp, vp, i64, ci, dbl, dbl, dbl, ci, vp]
        L.orc_target_list.restype = i64
        L.orc_target_list.argtypes = [ctypes.POINTER(_OrcTree), vp, dbl, vp, i64]
        L.orc_counters.argtypes = [ctypes.POINTER(_OrcTree), vp, i64, ci, dbl, ci, vp]
        L.orc_direct.argtypes = [vp, vp, i64, vp, i64, ci, ci, dbl, dbl, ci, vp]
        L.orc_direct_potential.argtypes = [vp, vp, i64, vp, i64, ci, ci, dbl, dbl, ci, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class OracleTree:
    """TreeData arrays (bhut/tree/node.py:59-77) produced by the oracle."""

    center: np.ndarray
    half_size: np.ndarray
    mass: np.ndarray
    com: np.ndarray
    first_child: np.ndarray
    child_count: np.ndarray
    start: np.ndarray
    count: np.ndarray
    is_leaf: np.ndarray
    perm: np.ndarray
    dim: int
    leaf_size: int

    @property
    def num_nodes(self) -> int:
        return len(self.mass)

    def _struct(self) -> _OrcTree:
        # keep contiguous references alive on the instance
        self._keep = [np.ascontiguousarray(getattr(self, f)) for f in (
            "center", "half_size", "mass", "com", "first_child", "child_count",
            "start", "count", "is_leaf", "perm")]
        c, h, m, co, fc, cc, st, cn, il, pm = self._keep
        return _OrcTree(len(m), len(pm), self.dim, self.leaf_size, _p(c), _p(h), _p(m), _p(co),
                        _p(fc), _p(cc), _p(st), _p(cn), _p(il), _p(pm))


def _alloc(n: int, dim: int, leaf: int):
    cap = lib().orc_max_nodes(n, leaf)
    arrs = dict(
        center=np.zeros((cap, dim)), half_size=np.zeros(cap), mass=np.zeros(cap),
        com=np.zeros((cap, dim)), first_child=np.full(cap, -1, np.int64),
        child_count=np.zeros(cap, np.int64), start=np.zeros(cap, np.int64),
        count=np.zeros(cap, np.int64), is_leaf=np.zeros(cap, np.uint8),
        perm=np.zeros(max(n, 1), np.int64))
    st = _OrcTree(0, n, dim, leaf, *[_p(arrs[k]) for k in (
        "center", "half_size", "mass", "com", "first_child", "child_count",
        "start", "count", "is_leaf", "perm")])
    return arrs, st


def _finish(arrs, M: int, n: int, dim: int, leaf: int) -> OracleTree:
    return OracleTree(
        center=arrs["center"][:M].copy(), half_size=arrs["half_size"][:M].copy(),
        mass=arrs["mass"][:M].copy(), com=arrs["com"][:M].copy(),
        first_child=arrs["first_child"][:M].copy(), child_count=arrs["child_count"][:M].copy(),
        start=arrs["start"][:M].copy(), count=arrs["count"][:M].copy(),
        is_leaf=arrs["is_leaf"][:M].astype(bool), perm=arrs["perm"][:n].copy(),
        dim=dim, leaf_size=leaf)


def np_sum(a) -> float:
    a = _f64(a)
    return lib().orc_np_sum(_p(a), len(a))


def aabb(pos):
    pos = _f64(pos)
    dim = pos.shape[1]
    lo, hi = np.empty(dim), np.empty(dim)
    lib().orc_aabb(_p(pos), len(pos), dim, _p(lo), _p(hi))
    return lo, hi


def morton_codes(pos):
    """Morton codes of raw positions (bbox.py:14-96 then morton.py:48-102)."""
    pos = _f64(pos)
    n, dim = pos.shape
    lo, hi = aabb(pos)
    p01 = np.empty_like(pos)
    lib().orc_normalize(_p(pos), n, dim, _p(lo), _p(hi), _p(p01))
    codes = np.empty(n, np.uint64)
    lib().orc_morton(_p(p01), n, dim, _p(codes))
    return codes


def argsort_stable(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.empty(len(keys), np.int64)
    lib().orc_argsort_stable(_p(keys), len(keys), _p(perm))
    return perm


def topology(n: int, leaf: int, dim: int = 3) -> OracleTree:
    arrs, st = _alloc(n, dim, leaf)
    M = lib().orc_topology(ctypes.byref(st))
    if M < 0:
        raise RuntimeError("node index exceeds max_nodes")
    return _finish(arrs, M, n, dim, leaf)


def build(pos, mass, leaf: int = 32) -> OracleTree:
    pos, mass = _f64(pos), _f64(mass)
    n, dim = pos.shape
    arrs, st = _alloc(n, dim, leaf)
    M = lib().orc_build(_p(pos), _p(mass), n, dim, leaf, ctypes.byref(st))
    if M < 0:
        raise ValueError("Cannot build tree with zero particles" if M == -2 else "tree build failed")
    return _finish(arrs, M, n, dim, leaf)


def refit(tree: OracleTree, pos, mass) -> OracleTree:
    pos, mass = _f64(pos), _f64(mass)
    out = OracleTree(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in tree.__dict__.items()
                        if not k.startswith("_")})
    st = out._struct()
    lib().orc_refit(ctypes.byref(st), _p(pos), _p(mass))
    out.mass, out.com = out._keep[2], out._keep[3]
    return out


def should_refit(tree: OracleTree, pos) -> bool:
    pos = _f64(pos)
    st = tree._struct()
    return bool(lib().orc_should_refit(ctypes.byref(st), _p(pos), pos.shape[0], pos.shape[1]))


def accelerations(tree: OracleTree, src_pos, src_mass, targets=None, *, softening=0.0,
                  theta=0.5, G=1.0, leaf_formula=LEAF_NUMBA, nthreads=0):
    src_pos, src_mass = _f64(src_pos), _f64(src_mass)
    self_mode = targets is None
    tgt = src_pos if self_mode else _f64(targets)
    out = np.zeros_like(tgt)
    st = tree._struct()
    lib().orc_accel(ctypes.byref(st), _p(src_pos), _p(src_mass), _p(tgt), len(tgt), int(self_mode),
                    float(softening), float(theta), float(G), int(leaf_formula), int(nthreads), _p(out))
    return out


def accelerations_subset(tree: OracleTree, src_pos, src_mass, idx, *, softening=0.0, theta=0.5, G=1.0,
                         leaf_formula=LEAF_NUMBA, nthreads=0):
    """Self accelerations of the particles idx only (each skips itself, as
    in bh.py:475-480), (len(idx), dim)."""
    src_pos, src_mass = _f64(src_pos), _f64(src_mass)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    if len(idx) and (idx.min() < 0 or idx.max() >= len(src_pos)):
        raise IndexError("idx out of range")
    out = np.zeros((len(idx), src_pos.shape[1]))
    st = tree._struct()
    lib().orc_accel_subset(ctypes.byref(st), _p(src_pos), _p(src_mass), _p(idx), len(idx), float(softening),
                           float(theta), float(G), int(leaf_formula), int(nthreads), _p(out))
    return out


def potential(tree: OracleTree, src_pos, src_mass, targets=None, *, softening=0.0, theta=0.5, G=1.0,
              nthreads=0):
    """Softened potential over the walk (no bhut counterpart; see orc_potential)."""
    src_pos, src_mass = _f64(src_pos), _f64(src_mass)
    self_mode = targets is None
    tgt = src_pos if self_mode else _f64(targets)
    out = np.zeros(len(tgt))
    st = tree._struct()
    lib().orc_potential(ctypes.byref(st), _p(src_pos), _p(src_mass), _p(tgt), len(tgt), int(self_mode),
                        float(softening), float(theta), float(G), int(nthreads), _p(out))
    return out


def target_list(tree: OracleTree, tpos, theta: float) -> np.ndarray:
    tpos = _f64(tpos)
    st = tree._struct()
    cap = 4096
    while True:
        out = np.empty(cap, np.int64)
        n = lib().orc_target_list(ctypes.byref(st), _p(tpos), float(theta), _p(out), cap)
        if n <= cap:
            return out[:n].copy()
        cap = int(n)


def counters(tree: OracleTree, targets, *, self_mode: bool, theta: float, nthreads=0) -> np.ndarray:
    """(m, 5) int64: pops, opening tests, monopoles, leaves, pairs."""
    tgt = _f64(targets)
    out = np.zeros((len(tgt), 5), np.int64)
    st = tree._struct()
    lib().orc_counters(ctypes.byref(st), _p(tgt), len(tgt), int(self_mode), float(theta), int(nthreads),
                       _p(out))
    return out


def direct(src_pos, src_mass, targets=None, *, softening=0.0, G=1.0, nthreads=0):
    src_pos, src_mass = _f64(src_pos), _f64(src_mass)
    self_mode = targets is None
    tgt = src_pos if self_mode else _f64(targets)
    out = np.zeros_like(tgt)
    lib().orc_direct(_p(src_pos), _p(src_mass), len(src_pos), _p(tgt), len(tgt), tgt.shape[1],
                     int(self_mode), float(softening), float(G), int(nthreads), _p(out))
    return out


def direct_potential(src_pos, src_mass, targets=None, *, softening=0.0, G=1.0, nthreads=0):
    src_pos, src_mass = _f64(src_pos), _f64(src_mass)
    self_mode = targets is None
    tgt = src_pos if self_mode else _f64(targets)
    out = np.zeros(len(tgt))
    lib().orc_direct_potential(_p(src_pos), _p(src_mass), len(src_pos), _p(tgt), len(tgt), tgt.shape[1],
                               int(self_mode), float(softening), float(G), int(nthreads), _p(out))
    return out


def statistics(tree: OracleTree, positions, theta: float) -> dict:
    """compute_tree_statistics (bhut/traverse/bh.py:604-646) from the exact
    counters of external-mode walks at `positions`: the reference's
    _count_operations_single_target (bh.py:648-700) runs the opening test at
    positions[i] and never uses the index, counts every pop, and adds
    count - 1 for each visited leaf (bh.py:678-681), i.e. pairs - leaves."""
    c = counters(tree, positions, self_mode=False, theta=theta)
    n = len(c)
    tot = [int(v) for v in c.sum(axis=0)]
    particles = tot[4] - tot[3]
    return {"avg_nodes_visited": tot[0] / n, "avg_particles_computed": particles / n,
            "avg_monopoles_used": tot[2] / n, "total_operations": particles + tot[2]}
