"""Write the golden fixtures under tests/golden/*.npz.

The fixtures pin the C oracle's restatement of the arithmetic that bhut hands
to NumPy/OpenBLAS/libm.  This script does not import bhut: it re-derives every
quantity by calling the same library routines bhut calls, on the same slices:

* ``np.argsort(codes, kind='stable')``                    (morton.py:135)
* ``np.min`` / ``np.max`` over node slices                 (bbox.py:43-44, build.py:236-240)
* ``np.sum(m)``, ``np.sum(m[:, None] * x, axis=0) / M``    (build.py:263-266, 284-287)
* ``np.linalg.norm(target - com)``                         (bh.py:593)
* ``np.sum(dx**2)``, ``r2s ** 1.5``                        (kernels.py:59-66)

so a wrong summation order, a missing fma or a transposed index in the oracle
fails ``tests/test_oracle_golden.py``.  Morton codes are computed by explicit
bit interleaving (the definition of the code), independent of the oracle.

Run:  python tests/golden/make_golden.py      (writes golden_*.npz beside it)
"""

from __future__ import annotations

import os
import platform

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- geometry
def morton_by_definition(pos: np.ndarray) -> np.ndarray:
    pmin, pmax = np.min(pos, axis=0), np.max(pos, axis=0)
    extent = pmax - pmin
    ext = np.where(extent <= 0, 1.0, extent)
    off = np.where(extent <= 0, 0.5, 0.0)
    p01 = np.minimum((pos - pmin) / ext + off, 1.0 - 1e-12)
    q = np.minimum((p01 * (1 << 21)).astype(np.uint64), np.uint64((1 << 21) - 1))
    n, dim = pos.shape
    codes = np.zeros(n, np.uint64)
    for d in range(dim):
        for b in range(21):
            codes |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim + d)
    return codes


def topology(n: int, leaf: int):
    """Median split with LIFO numbering, written from the description of
    build.py:138-186; its output is also checked against invariants and the
    closed-form node counts in tests/test_topology.py."""
    cap = max(4 * n // leaf, 10)
    first_child = np.full(cap, -1, np.int64)
    child_count = np.zeros(cap, np.int64)
    start = np.zeros(cap, np.int64)
    count = np.zeros(cap, np.int64)
    is_leaf = np.zeros(cap, bool)
    todo = [(0, 0, n)]
    nxt = 1
    while todo:
        node, s, e = todo.pop()
        start[node], count[node] = s, e - s
        if e - s <= leaf:
            is_leaf[node] = True
            continue
        mid = s + (e - s) // 2
        first_child[node], child_count[node] = nxt, 2
        todo.append((nxt, s, mid))
        todo.append((nxt + 1, mid, e))
        nxt += 2
    M = nxt
    return first_child[:M], child_count[:M], start[:M], count[:M], is_leaf[:M]


def tree_arrays(pos: np.ndarray, mass: np.ndarray, leaf: int):
    n, dim = pos.shape
    codes = morton_by_definition(pos)
    perm = np.argsort(codes, kind="stable").astype(np.int64)
    sp, sm = pos[perm], mass[perm]
    fc, cc, st, cn, il = topology(n, leaf)
    M = len(fc)
    center = np.zeros((M, dim))
    half = np.zeros(M)
    pmin, pmax = np.min(pos, axis=0), np.max(pos, axis=0)
    center[0] = (pmin + pmax) / 2
    half[0] = np.max(pmax - pmin) / 2
    for k in range(1, M):
        sl = sp[st[k]:st[k] + cn[k]]
        lo, hi = np.min(sl, axis=0), np.max(sl, axis=0)
        center[k] = (lo + hi) / 2
        half[k] = np.max(hi - lo) / 2
    nmass = np.zeros(M)
    com = np.zeros((M, dim))
    for k in range(M - 1, -1, -1):
        if il[k]:
            m = sm[st[k]:st[k] + cn[k]]
            x = sp[st[k]:st[k] + cn[k]]
        else:
            ch = np.arange(fc[k], fc[k] + cc[k])
            m = nmass[ch]
            x = com[ch]
        tot = np.sum(m)
        nmass[k] = tot
        com[k] = np.sum(m[:, None] * x, axis=0) / tot if tot > 0 else center[k]
    return dict(codes=codes, perm=perm, center=center, half_size=half, mass=nmass, com=com,
                first_child=fc, child_count=cc, start=st, count=cn, is_leaf=il)


def refit_arrays(t: dict, pos: np.ndarray, mass: np.ndarray):
    """refit.py:112-192 with the same NumPy calls."""
    M, dim = t["com"].shape
    nmass = t["mass"].copy()
    com = t["com"].copy()
    for k in range(M):
        if not t["is_leaf"][k]:
            continue
        idx = t["perm"][t["start"][k]:t["start"][k] + t["count"][k]]
        m, x = mass[idx], pos[idx]
        tot = np.sum(m)
        nmass[k] = tot
        com[k] = np.sum(m[:, np.newaxis] * x, axis=0) / tot if tot > 0 else t["center"][k]
    for k in range(M - 1, -1, -1):
        if t["is_leaf"][k]:
            continue
        tot = 0.0
        wc = np.zeros(dim)
        for ch in range(t["first_child"][k], t["first_child"][k] + t["child_count"][k]):
            if nmass[ch] > 0:
                tot += nmass[ch]
                wc += nmass[ch] * com[ch]
        nmass[k] = tot
        com[k] = wc / tot if tot > 0 else t["center"][k]
    return nmass, com


# ---------------------------------------------------------------- traversal
def acc_python(dx, m, eps, G):
    """kernels.py:57-69 with NumPy scalars."""
    r2 = np.sum(dx ** 2)
    r2s = r2 + np.asarray(eps, np.float64) ** 2
    if r2s == 0:
        return np.zeros_like(dx)
    return (G * np.asarray(m, np.float64) / (r2s ** 1.5)) * dx


def leaf_numba_formula(t, k, tp, tidx, pos, mass, eps, G):
    """The arithmetic of bh.py:249-278 in float64 scalars; returns None where
    Numba would raise ZeroDivisionError."""
    acc = np.zeros(len(tp))
    e2 = eps * eps
    for i in range(t["start"][k], t["start"][k] + t["count"][k]):
        p = t["perm"][i]
        if p == tidx:
            continue
        dx = pos[p] - tp
        rsq = 0.0
        for d in range(len(tp)):
            rsq += dx[d] * dx[d]
        rsq += e2
        r = np.sqrt(rsq)
        if r == 0.0:
            return None
        rinv = 1.0 / r
        mag = G * mass[p] * (rinv * rinv * rinv)
        for d in range(len(tp)):
            acc[d] += mag * dx[d]
    return acc


def leaf_python_formula(t, k, tp, tidx, pos, mass, eps, G):
    acc = np.zeros(len(tp))
    for i in range(t["start"][k], t["start"][k] + t["count"][k]):
        p = t["perm"][i]
        if p == tidx:
            continue
        acc += acc_python(pos[p] - tp, mass[p], eps, G)
    return acc


def walk(t, tp, tidx, pos, mass, eps, theta, G, formula):
    acc = np.zeros(len(tp))
    lst = []
    stack = [0]
    while stack:
        k = stack.pop()
        if t["mass"][k] == 0:
            continue
        if t["first_child"][k] == -1:
            lst.append(k)
            a = None
            if formula == "numba":
                a = leaf_numba_formula(t, k, tp, tidx, pos, mass, eps, G)
            if a is None:
                a = leaf_python_formula(t, k, tp, tidx, pos, mass, eps, G)
            acc += a
        else:
            dist = np.linalg.norm(tp - t["com"][k])
            if dist == 0 or (2 * t["half_size"][k]) / dist >= theta:
                c = t["first_child"][k]
                for _ in range(t["child_count"][k]):
                    stack.append(c)
                    c += 1
            else:
                lst.append(k)
                acc += acc_python(t["com"][k] - tp, t["mass"][k], eps, G)
    return acc, lst


# ---------------------------------------------------------------- cases
def cases():
    out = []
    # tests/conftest.py:14-21 simple_system
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    out.append(("simple2", p, np.ones(2), [32, 1]))
    # tests/conftest.py:24-30 random_system
    np.random.seed(42)
    out.append(("random10", np.random.randn(10, 3), np.ones(10), [1, 4, 32]))
    # tests/conftest.py:33-42 hierarchical_system
    p = np.array([[-0.1, 0.0, 0.0], [0.1, 0.0, 0.0], [10.0, 0.0, 0.0]])
    out.append(("hier3", p, np.array([1.0, 1.0, 0.1]), [1, 2]))
    # tests/conftest.py:45-51 planar_system (a degenerate z axis)
    np.random.seed(123)
    p = np.random.randn(8, 3)
    p[:, 2] = 0.0
    out.append(("planar8", p, np.ones(8), [1, 2]))
    # tests/conftest.py:64-70 varied_masses_system
    np.random.seed(456)
    out.append(("varied8", np.random.randn(8, 3),
                np.array([0.1, 1.0, 10.0, 0.5, 2.0, 0.01, 5.0, 0.2]), [2]))
    rng = np.random.default_rng(7)
    out.append(("gauss257", rng.standard_normal((257, 3)), rng.uniform(0.5, 2.0, 257), [4, 16]))
    # coincident particles and tied Morton codes (stability of the sort)
    rng = np.random.default_rng(11)
    base = rng.standard_normal((40, 3))
    p = np.concatenate([base, base[:20], base[:5]])
    out.append(("dups65", p, rng.uniform(0.1, 1.0, len(p)), [4]))
    # a zero-mass leaf and negative masses
    rng = np.random.default_rng(12)
    p = rng.standard_normal((48, 3))
    m = rng.uniform(0.5, 1.5, 48)
    m[np.argsort(p[:, 0])[:6]] = 0.0
    m[5] = -0.25
    out.append(("massedge48", p, m, [3]))
    rng = np.random.default_rng(0)
    out.append(("gauss2000", rng.standard_normal((2000, 3)), rng.uniform(0.5, 1.5, 2000), [32, 8]))
    rng = np.random.default_rng(21)
    out.append(("plane300_2d", rng.standard_normal((300, 2)), np.ones(300), [8]))
    return out


def main():
    meta = dict(numpy=np.__version__, machine=platform.machine(),
                cpu=[l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][:1],
                blas=str(np.show_config(mode="dicts")["Build Dependencies"]["blas"].get("version")))
    thetas = [0.0, 0.3, 0.5, 1.0]
    for name, pos, mass, leaves in cases():
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        n, dim = pos.shape
        rng = np.random.default_rng(99)
        ext = rng.uniform(-3, 3, size=(min(32, 4 + n // 20), dim))
        pos2 = pos + 0.01 * rng.standard_normal(pos.shape)
        rec = dict(pos=pos, mass=mass, ext=ext, pos2=pos2, thetas=np.array(thetas),
                   meta=np.array(repr(meta)))
        for leaf in leaves:
            t = tree_arrays(pos, mass, leaf)
            for k, v in t.items():
                rec[f"L{leaf}_{k}"] = v
            rm, rc = refit_arrays(t, pos2, mass)
            rec[f"L{leaf}_refit_mass"], rec[f"L{leaf}_refit_com"] = rm, rc
            nt = min(n, 96)
            for theta in thetas:
                for eps in (0.0, 0.01):
                    tag = f"L{leaf}_t{theta}_e{eps}"
                    accn = np.zeros((nt, dim))
                    accp = np.zeros((nt, dim))
                    ptr, ids = [0], []
                    for i in range(nt):
                        accn[i], lst = walk(t, pos[i], i, pos, mass, eps, theta, 1.0, "numba")
                        accp[i], _ = walk(t, pos[i], i, pos, mass, eps, theta, 1.0, "python")
                        ids += lst
                        ptr.append(len(ids))
                    rec[f"{tag}_acc_numba"] = accn
                    rec[f"{tag}_acc_python"] = accp
                    rec[f"{tag}_list_ptr"] = np.array(ptr, np.int64)
                    rec[f"{tag}_list_ids"] = np.array(ids, np.int64)
                    acce = np.zeros_like(ext)
                    ptr, ids = [0], []
                    for i in range(len(ext)):
                        acce[i], lst = walk(t, ext[i], -1, pos, mass, eps, theta, 1.0, "numba")
                        ids += lst
                        ptr.append(len(ids))
                    rec[f"{tag}_ext_acc"] = acce
                    rec[f"{tag}_ext_list_ptr"] = np.array(ptr, np.int64)
                    rec[f"{tag}_ext_list_ids"] = np.array(ids, np.int64)
        rec["leaves"] = np.array(leaves, np.int64)
        path = os.path.join(HERE, f"golden_{name}.npz")
        np.savez_compressed(path, **rec)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
