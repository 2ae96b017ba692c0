"""Write tests/golden/ref_*.npz: outputs of the reference's own traversal.

Runs ONLY in the build container, where the reference package (maroba/bhut,
pure Python; numba and dask absent) sits at /root/reference.  The reference
never travels with the repository: only the .npz files written here do.

What it records are the outputs of the functions whose bits no library call
or closed form fixes on its own:

* the per-target interaction list of ``_barnes_hut_single_target`` and
  ``_barnes_hut_single_target_external`` (bhut/traverse/bh.py:83-146,
  485-548) in POP ORDER: the walk is run by the reference itself, with
  ``_should_open_node`` (bh.py:576-601) and the leaf functions (bh.py:
  551-573, 150-171) wrapped so that every accepted monopole (an internal
  node whose opening test returned False) and every visited leaf is
  appended in the order the walk reaches it; the wrappers return what the
  wrapped functions return;
* the walk's accelerations as the reference computes them here (pure-Python
  leaf path, bh.py:384-425, kernels.py:14-95), and with the Numba leaf loop
  (bh.py:174-280) executed by CPython: numba is absent, so its ``jit``
  decorator is a no-op (bh.py:17-27) and setting ``HAVE_NUMBA`` selects the
  undecorated loop; ``np.errstate(divide='raise')`` turns a zero distance
  into an exception as Numba's ZeroDivisionError does, so the reference's own
  ``except Exception`` fallback (bh.py:307-311, 339-343) runs;
* per-target operation counts of ``_count_operations_single_target``
  (bh.py:648-700) and the number of opening tests the walk makes, and the
  dict returned by ``compute_tree_statistics`` (bh.py:604-646), also at
  displaced positions of the same count;
* the TreeData the reference builds for each case (bhut/tree/build.py:
  19-107), which is the walk's input;
* ``bhut.accelerations`` (bhut/api.py:254-364) on the CPU-runnable workload:
  10 000 standard-normal particles (seed 0), unit masses, theta 0.5,
  softening 0.01, leaf size 32, with lists and counts for 256 targets.

Run:  python tests/golden/make_golden_ref.py [/root/reference]
"""

from __future__ import annotations

import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import cases  # noqa: E402  (the fixture systems)

TREE_FIELDS = ("center", "half_size", "mass", "com", "first_child", "child_count", "start", "count",
               "is_leaf", "perm")


class Recorder:
    """Wraps the reference's per-node callees so a walk reports its list."""

    def __init__(self, B):
        self.B = B
        self.on = False
        self.lst: list[int] = []
        self.tests = 0
        self._open = B._should_open_node
        self._leaf = B._leaf_acceleration
        self._leaf_ext = B._leaf_acceleration_external

        def should_open(tree, node_idx, target_pos, theta):
            r = self._open(tree, node_idx, target_pos, theta)
            if self.on:
                self.tests += 1
                if not r:
                    self.lst.append(int(node_idx))
            return r

        def leaf(tree, leaf_idx, *a):
            if self.on:
                self.lst.append(int(leaf_idx))
            return self._leaf(tree, leaf_idx, *a)

        def leaf_ext(tree, leaf_idx, *a):
            if self.on:
                self.lst.append(int(leaf_idx))
            return self._leaf_ext(tree, leaf_idx, *a)

        B._should_open_node = should_open
        B._leaf_acceleration = leaf
        B._leaf_acceleration_external = leaf_ext

    def start(self):
        self.on, self.lst, self.tests = True, [], 0

    def stop(self):
        self.on = False
        return self.lst, self.tests


def walk_self(B, rec, tree, pos, mass, i, eps, theta, numba):
    B.HAVE_NUMBA = numba
    rec.start()
    with np.errstate(divide="raise"):
        a = B._barnes_hut_single_target(tree, pos[i], i, pos, mass, eps, theta, 1.0)
    lst, tests = rec.stop()
    B.HAVE_NUMBA = False
    return np.asarray(a, np.float64), lst, tests


def walk_ext(B, rec, tree, pos, mass, tp, eps, theta, numba):
    B.HAVE_NUMBA = numba
    rec.start()
    with np.errstate(divide="raise"):
        a = B._barnes_hut_single_target_external(tree, pos, mass, tp, eps, theta, 1.0)
    lst, tests = rec.stop()
    B.HAVE_NUMBA = False
    return np.asarray(a, np.float64), lst, tests


def ops(B, tree, tp, i, theta):
    c = B._count_operations_single_target(tree, tp, i, theta)
    return c["nodes_visited"], c["particles_computed"], c["monopoles_used"]


def stats_arrays(d):
    keys = ("avg_nodes_visited", "avg_particles_computed", "avg_monopoles_used", "total_operations")
    return np.array([float(d[k]) for k in keys], np.float64)


def main(refroot: str) -> None:
    sys.path.insert(0, refroot)
    import bhut
    import bhut.traverse.bh as B
    from bhut.backends.numpy_ import numpy_namespace
    from bhut.tree.build import build_tree

    assert not B.HAVE_NUMBA, "this script expects the reference's pure-Python path (numba absent)"
    rec = Recorder(B)
    meta = dict(numpy=np.__version__, python=platform.python_version(), machine=platform.machine(),
                cpu=[l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][:1],
                blas=str(np.show_config(mode="dicts")["Build Dependencies"]["blas"].get("version")),
                reference=os.path.abspath(refroot), bhut_version=getattr(bhut, "__version__", "?"))
    thetas = [0.0, 0.3, 0.5, 1.0]
    for name, pos, mass, leaves in cases():
        t0 = time.time()
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        n, dim = pos.shape
        # the same external targets and displaced positions as golden_<name>.npz
        rng = np.random.default_rng(99)
        ext = rng.uniform(-3, 3, size=(min(32, 4 + n // 20), dim))
        pos2 = pos + 0.01 * rng.standard_normal(pos.shape)
        out = dict(pos=pos, mass=mass, ext=ext, pos2=pos2, thetas=np.array(thetas), leaves=np.array(leaves),
                   meta=np.array(repr(meta)))
        for leaf in leaves:
            tree = build_tree(pos, mass, dim=dim, leaf_size=leaf, xp=numpy_namespace)
            for f in TREE_FIELDS:
                out[f"L{leaf}_{f}"] = np.asarray(getattr(tree, f))
            nt = min(n, 96)
            for theta in thetas:
                out[f"L{leaf}_t{theta}_stats"] = stats_arrays(B.compute_tree_statistics(tree, pos, theta))
                out[f"L{leaf}_t{theta}_stats_pos2"] = stats_arrays(B.compute_tree_statistics(tree, pos2, theta))
                cnt = np.array([ops(B, tree, pos[i], i, theta) for i in range(nt)], np.int64).reshape(nt, 3)
                cnt_e = np.array([ops(B, tree, ext[i], -1, theta) for i in range(len(ext))],
                                 np.int64).reshape(len(ext), 3)
                out[f"L{leaf}_t{theta}_ops"] = cnt
                out[f"L{leaf}_t{theta}_ext_ops"] = cnt_e
                for eps in (0.0, 0.01):
                    tag = f"L{leaf}_t{theta}_e{eps}"
                    accp, accn = np.zeros((nt, dim)), np.zeros((nt, dim))
                    ptr, ids, tests = [0], [], []
                    for i in range(nt):
                        accp[i], lst, nts = walk_self(B, rec, tree, pos, mass, i, eps, theta, False)
                        accn[i], lst2, _ = walk_self(B, rec, tree, pos, mass, i, eps, theta, True)
                        assert lst == lst2
                        ids += lst
                        ptr.append(len(ids))
                        tests.append(nts)
                    out[f"{tag}_acc_python"], out[f"{tag}_acc_numba"] = accp, accn
                    out[f"{tag}_list_ptr"] = np.array(ptr, np.int64)
                    out[f"{tag}_list_ids"] = np.array(ids, np.int64)
                    out[f"{tag}_tests"] = np.array(tests, np.int64)
                    accp, accn = np.zeros_like(ext), np.zeros_like(ext)
                    ptr, ids, tests = [0], [], []
                    for i in range(len(ext)):
                        accp[i], lst, nts = walk_ext(B, rec, tree, pos, mass, ext[i], eps, theta, False)
                        accn[i], lst2, _ = walk_ext(B, rec, tree, pos, mass, ext[i], eps, theta, True)
                        assert lst == lst2
                        ids += lst
                        ptr.append(len(ids))
                        tests.append(nts)
                    out[f"{tag}_ext_acc_python"], out[f"{tag}_ext_acc_numba"] = accp, accn
                    out[f"{tag}_ext_list_ptr"] = np.array(ptr, np.int64)
                    out[f"{tag}_ext_list_ids"] = np.array(ids, np.int64)
                    out[f"{tag}_ext_tests"] = np.array(tests, np.int64)
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} {os.path.getsize(path)} bytes ({time.time() - t0:.1f} s)", flush=True)

    # the CPU-runnable workload through the public API
    t0 = time.time()
    rng = np.random.default_rng(0)
    x = rng.standard_normal((10_000, 3))
    m = np.ones(10_000)
    acc = bhut.accelerations(x, m, theta=0.5, softening=0.01, G=1.0, leaf_size=32)
    t_api = time.time() - t0
    tree = build_tree(x, m, dim=3, leaf_size=32, xp=numpy_namespace)
    out = dict(pos=x, mass=m, acc=np.asarray(acc, np.float64), meta=np.array(repr(meta)),
               stats=stats_arrays(B.compute_tree_statistics(tree, x, 0.5)))
    for f in TREE_FIELDS:
        out[f"L32_{f}"] = np.asarray(getattr(tree, f))
    sel = np.random.default_rng(5).choice(10_000, 256, replace=False)
    ptr, ids, tests, accn = [0], [], [], np.zeros((len(sel), 3))
    for k, i in enumerate(sel):
        accn[k], lst, nts = walk_self(B, rec, tree, x, m, int(i), 0.01, 0.5, True)
        ids += lst
        ptr.append(len(ids))
        tests.append(nts)
    out.update(sel=sel.astype(np.int64), sel_list_ptr=np.array(ptr, np.int64),
               sel_list_ids=np.array(ids, np.int64), sel_tests=np.array(tests, np.int64),
               sel_acc_numba=accn,
               sel_ops=np.array([ops(B, tree, x[i], int(i), 0.5) for i in sel], np.int64))
    path = os.path.join(HERE, "ref_normal10k.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} {os.path.getsize(path)} bytes (bhut.accelerations {t_api:.1f} s, "
          f"total {time.time() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
