"""ctypes binding of libbhut_hip.so (the C ABI declared in include/bhut_hip.h).

There is no CPU fallback: every compute call goes to the HIP library, and a
missing library or device raises.  Status codes map to the reference's
exception types: invalid arguments and empty inputs raise ``ValueError``
(bhut/api.py:84-152, bhut/tree/build.py:81-82), everything else raises
``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbhut_hip.so")

BH_OK, BH_EINVAL, BH_EHIP, BH_ENOMEM, BH_EEMPTY, BH_ENODEV, BH_EOVERFLOW = 0, -1, -2, -3, -4, -5, -6

i64, i32, dbl, vp = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p


class F32Opts(ctypes.Structure):
    _fields_ = [("group_size", i32), ("kahan", i32), ("stack_cap", i32), ("multipole", i32)]


# name -> (restype, argtypes); mirrors include/bhut_hip.h
_SIGS = {
    "bh_last_error": (ctypes.c_char_p, []),
    "bh_version": (ctypes.c_int, []),
    "bh_device_count": (ctypes.c_int, []),
    "bh_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    "bh_ctx_destroy": (None, [vp]),
    "bh_ctx_synchronize": (ctypes.c_int, [vp]),
    "bh_ctx_stream": (vp, [vp]),
    "bh_ctx_set_stream": (ctypes.c_int, [vp, vp]),
    "bh_ctx_trim": (ctypes.c_int, [vp]),
    "bh_ctx_set_list_budget": (ctypes.c_int, [vp, i64]),
    "bh_ctx_profile": (ctypes.c_int, [vp, i32]),
    "bh_ctx_profile_summary": (ctypes.c_int, [vp, ctypes.c_char_p, i64]),
    "bh_tree_build": (ctypes.c_int, [vp, vp, vp, i64, i32, i32, ctypes.POINTER(vp)]),
    "bh_tree_build_dev": (ctypes.c_int, [vp, vp, vp, i64, i32, i32, ctypes.POINTER(vp)]),
    "bh_tree_import": (ctypes.c_int, [vp, i64, i64, i32, i32] + [vp] * 11 + [ctypes.POINTER(vp)]),
    "bh_tree_clone": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "bh_forest_build": (ctypes.c_int, [vp, vp, vp, vp, i64, i32, i32, ctypes.POINTER(vp)]),
    "bh_forest_build_dev": (ctypes.c_int, [vp, vp, vp, vp, i64, i32, i32, ctypes.POINTER(vp)]),
    "bh_tree_systems": (ctypes.c_int, [vp, ctypes.POINTER(i64), vp, vp]),
    "bh_tree_free": (None, [vp]),
    "bh_tree_shape": (ctypes.c_int, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i32),
                                     ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    "bh_tree_export": (ctypes.c_int, [vp] * 11),
    "bh_tree_f32_stack": (ctypes.c_int, [vp, vp]),
    "bh_tree_should_refit": (ctypes.c_int, [vp, vp, i64, i32, ctypes.POINTER(i32)]),
    "bh_tree_should_refit_dev": (ctypes.c_int, [vp, vp, i64, i32, ctypes.POINTER(i32)]),
    "bh_tree_refit": (ctypes.c_int, [vp, vp, vp]),
    "bh_tree_refit_dev": (ctypes.c_int, [vp, vp, vp]),
    "bh_accel_self": (ctypes.c_int, [vp, dbl, dbl, dbl, vp]),
    "bh_accel_self_dev": (ctypes.c_int, [vp, dbl, dbl, dbl, vp]),
    "bh_accel_targets": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp]),
    "bh_accel_targets_dev": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp]),
    "bh_accel_targets_dev_noalloc": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp]),
    "bh_accel_targets_streamed": (ctypes.c_int, [vp, vp, i64, i64, dbl, dbl, dbl, i32, vp, vp]),
    "bh_accel_f32_self": (ctypes.c_int, [vp, dbl, dbl, dbl, vp, vp, vp]),
    "bh_accel_f32_self_dev": (ctypes.c_int, [vp, dbl, dbl, dbl, vp, vp, vp]),
    "bh_accel_f32_targets": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp, vp, vp]),
    "bh_accel_f32_targets_dev": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp, vp, vp]),
    "bh_potential": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp]),
    "bh_potential_dev": (ctypes.c_int, [vp, vp, i64, dbl, dbl, dbl, vp]),
    "bh_interaction_lists": (ctypes.c_int, [vp, vp, i64, dbl, vp, vp, i64, ctypes.POINTER(i64)]),
    "bh_counters": (ctypes.c_int, [vp, vp, i64, dbl, vp]),
    "bh_direct": (ctypes.c_int, [vp, vp, vp, i64, vp, i64, i32, i32, dbl, dbl, vp, vp]),
    "bh_direct_dev": (ctypes.c_int, [vp, vp, vp, i64, vp, i64, i32, i32, dbl, dbl, vp, vp]),
    "bh_sim_create": (ctypes.c_int, [vp, vp, vp, vp, i64, i32, i32, dbl, dbl, dbl, i32, vp, ctypes.POINTER(vp)]),
    "bh_sim_run": (ctypes.c_int, [vp, dbl, i64, i32, i32]),
    "bh_sim_accelerations": (ctypes.c_int, [vp, vp]),
    "bh_sim_set_list_capacity": (ctypes.c_int, [vp, i64, i64]),
    "bh_sim_info": (ctypes.c_int, [vp, vp]),
    "bh_sim_destroy": (None, [vp]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load libbhut_hip.so and declare its signatures."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                    " or `make -C bhut_amd/csrc`")
            # One HIP runtime per process: PyTorch-ROCm ships its own
            # libamdhip64 (needed as "libamdhip64.so", soname ".so.7").  When
            # torch is loaded first, this library's ".so.7" dependency binds
            # to that same copy; loaded the other way round the process would
            # hold two runtimes and torch would find no GPU.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    msg = load().bh_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    if rc == BH_OK:
        return
    msg = last_error()
    if rc in (BH_EINVAL, BH_EEMPTY):
        raise ValueError(msg)
    raise RuntimeError(f"libbhut_hip error {rc}: {msg}")


class Context:
    """One device, one stream, one cache of device allocations."""

    def __init__(self, device: int = 0):
        lib = load()
        h = vp()
        check(lib.bh_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    def synchronize(self) -> None:
        check(load().bh_ctx_synchronize(self.handle))

    def set_stream(self, stream_ptr) -> None:
        check(load().bh_ctx_set_stream(self.handle, vp(stream_ptr) if stream_ptr else None))

    def set_list_budget(self, nbytes: int) -> None:
        """Bytes of FP32 interaction lists held at once (0: default)."""
        check(load().bh_ctx_set_list_budget(self.handle, int(nbytes)))

    def trim(self) -> None:
        check(load().bh_ctx_trim(self.handle))

    def profile(self, enable: bool) -> None:
        check(load().bh_ctx_profile(self.handle, int(bool(enable))))

    def profile_summary(self) -> dict:
        """{kernel name: (launches, total ms)} recorded since profile(True)."""
        buf = ctypes.create_string_buffer(1 << 16)
        check(load().bh_ctx_profile_summary(self.handle, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.bh_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_ctx = None


def ctx() -> Context:
    """The process-wide default context (device from BHUT_DEVICE, else 0)."""
    global _ctx
    if _ctx is None:
        with _lock:
            pass
        _ctx = Context(int(os.environ.get("BHUT_DEVICE", "0")))
    return _ctx


def device_count() -> int:
    return int(load().bh_device_count())


def ptr(a: np.ndarray):
    return vp(a.ctypes.data) if a is not None else None


def f64c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _rows(a: np.ndarray, what: str, n: int | None, dim: int) -> np.ndarray:
    """a as a contiguous float64 (rows, dim) array; ValueError otherwise (the
    C side reads rows * dim values, so a wrong shape must stop here)."""
    a = f64c(a)
    if a.ndim != 2 or a.shape[1] != dim or (n is not None and a.shape[0] != n):
        want = f"({n if n is not None else 'M'}, {dim})"
        raise ValueError(f"{what} must have shape {want}, got {a.shape}")
    return a


def _vec(a: np.ndarray, what: str, n: int) -> np.ndarray:
    a = f64c(a)
    if a.shape != (n,):
        raise ValueError(f"{what} must have shape ({n},), got {a.shape}")
    return a


class DeviceTree:
    """Owner of a bh_tree*; freed with the Python object."""

    def __init__(self, handle: vp, context: Context):
        self.handle = handle
        self.context = context  # keeps the context alive
        n, M, dim, leaf, depth = i64(), i64(), i32(), i32(), i32()
        check(load().bh_tree_shape(handle, ctypes.byref(n), ctypes.byref(M), ctypes.byref(dim),
                                   ctypes.byref(leaf), ctypes.byref(depth)))
        self.n, self.num_nodes, self.dim, self.leaf_size, self.depth = n.value, M.value, dim.value, leaf.value, depth.value

    # -- construction -------------------------------------------------
    @classmethod
    def build(cls, positions: np.ndarray, masses: np.ndarray, leaf_size: int,
              context: Context | None = None) -> "DeviceTree":
        c = context or ctx()
        pos, m = f64c(positions), f64c(masses)
        h = vp()
        check(load().bh_tree_build(c.handle, ptr(pos), ptr(m), pos.shape[0], pos.shape[1], int(leaf_size),
                                   ctypes.byref(h)))
        return cls(h, c)

    @classmethod
    def build_dev(cls, pos_ptr: int, mass_ptr: int, n: int, dim: int, leaf_size: int,
                  context: Context | None = None) -> "DeviceTree":
        c = context or ctx()
        h = vp()
        check(load().bh_tree_build_dev(c.handle, vp(pos_ptr), vp(mass_ptr), int(n), int(dim), int(leaf_size),
                                       ctypes.byref(h)))
        return cls(h, c)

    @classmethod
    def from_arrays(cls, td, positions: np.ndarray, masses: np.ndarray,
                    context: Context | None = None) -> "DeviceTree":
        c = context or ctx()
        arrs = [f64c(td.center), f64c(td.half_size), f64c(td.mass), f64c(td.com)]
        ints = [np.ascontiguousarray(getattr(td, k), dtype=np.int64)
                for k in ("first_child", "child_count", "start", "count", "perm")]
        n, M, dim = len(ints[4]), len(arrs[2]), int(td.dim)
        pos = _rows(positions, "positions", n, dim)
        m = _vec(masses, "masses", n)
        for name, a, shape in (("center", arrs[0], (M, dim)), ("half_size", arrs[1], (M,)),
                               ("com", arrs[3], (M, dim)), ("first_child", ints[0], (M,)),
                               ("child_count", ints[1], (M,)), ("start", ints[2], (M,)),
                               ("count", ints[3], (M,))):
            if a.shape != shape:
                raise ValueError(f"{name} has shape {a.shape}, expected {shape}")
        h = vp()
        check(load().bh_tree_import(c.handle, len(ints[4]), len(arrs[2]), int(td.dim), int(td.leaf_size),
                                    *[ptr(a) for a in arrs], *[ptr(a) for a in ints], ptr(pos), ptr(m),
                                    ctypes.byref(h)))
        return cls(h, c)

    @classmethod
    def build_forest(cls, positions: np.ndarray, masses: np.ndarray, counts, leaf_size: int,
                     context: Context | None = None) -> "DeviceTree":
        """One forest of len(counts) independent systems stored back to back."""
        c = context or ctx()
        pos, m = f64c(positions), f64c(masses)
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        if pos.ndim != 2 or m.shape != (pos.shape[0],) or int(cnt.sum()) != pos.shape[0]:
            raise ValueError("positions (n, dim), masses (n,) and sum(counts) == n required")
        h = vp()
        check(load().bh_forest_build(c.handle, ptr(pos), ptr(m), ptr(cnt), len(cnt), pos.shape[1],
                                     int(leaf_size), ctypes.byref(h)))
        return cls(h, c)

    @classmethod
    def build_forest_dev(cls, pos_ptr: int, mass_ptr: int, counts, dim: int, leaf_size: int,
                         context: Context | None = None) -> "DeviceTree":
        c = context or ctx()
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        h = vp()
        check(load().bh_forest_build_dev(c.handle, vp(pos_ptr), vp(mass_ptr), ptr(cnt), len(cnt), int(dim),
                                         int(leaf_size), ctypes.byref(h)))
        return cls(h, c)

    def systems(self) -> tuple[np.ndarray, np.ndarray]:
        """Particle and node offsets of the systems (nsys + 1 entries each)."""
        ns = i64()
        check(load().bh_tree_systems(self.handle, ctypes.byref(ns), None, None))
        po, no = np.empty(ns.value + 1, np.int64), np.empty(ns.value + 1, np.int64)
        check(load().bh_tree_systems(self.handle, ctypes.byref(ns), ptr(po), ptr(no)))
        return po, no

    def f32_stack(self) -> tuple[int, int]:
        """LDS stack entries the FP32 walk grew to (self, external targets)."""
        out = np.zeros(2, np.int32)
        check(load().bh_tree_f32_stack(self.handle, ptr(out)))
        return int(out[0]), int(out[1])

    def clone(self) -> "DeviceTree":
        h = vp()
        check(load().bh_tree_clone(self.handle, ctypes.byref(h)))
        return DeviceTree(h, self.context)

    # -- queries --------------------------------------------------------
    def export(self) -> dict:
        M, n, d = self.num_nodes, self.n, self.dim
        out = dict(center=np.empty((M, d)), half_size=np.empty(M), mass=np.empty(M), com=np.empty((M, d)),
                   first_child=np.empty(M, np.int64), child_count=np.empty(M, np.int64),
                   start=np.empty(M, np.int64), count=np.empty(M, np.int64), is_leaf=np.empty(M, np.uint8),
                   perm=np.empty(n, np.int64))
        check(load().bh_tree_export(self.handle, *[ptr(out[k]) for k in (
            "center", "half_size", "mass", "com", "first_child", "child_count", "start", "count", "is_leaf",
            "perm")]))
        out["is_leaf"] = out["is_leaf"].astype(bool)
        return out

    def should_refit(self, positions: np.ndarray) -> bool:
        pos = f64c(positions)
        if pos.ndim != 2:
            raise ValueError(f"positions must be 2D array, got shape {pos.shape}")
        ok = i32()
        check(load().bh_tree_should_refit(self.handle, ptr(pos), pos.shape[0], pos.shape[1], ctypes.byref(ok)))
        return bool(ok.value)

    def refit(self, positions: np.ndarray, masses: np.ndarray | None) -> None:
        pos = _rows(positions, "positions", self.n, self.dim)
        m = _vec(masses, "masses", self.n) if masses is not None else None
        check(load().bh_tree_refit(self.handle, ptr(pos), ptr(m)))

    def accel_self(self, theta: float, softening: float, G: float) -> np.ndarray:
        out = np.empty((self.n, self.dim))
        check(load().bh_accel_self(self.handle, float(theta), float(softening), float(G), ptr(out)))
        return out

    # -- device-pointer forms (inputs already resident in device memory) --
    def accel_self_dev(self, out_ptr: int, theta: float, softening: float, G: float) -> None:
        check(load().bh_accel_self_dev(self.handle, float(theta), float(softening), float(G), vp(out_ptr)))

    def accel_targets_dev(self, tgt_ptr: int, m: int, out_ptr: int, theta: float, softening: float,
                          G: float) -> None:
        check(load().bh_accel_targets_dev(self.handle, vp(tgt_ptr), int(m), float(theta), float(softening),
                                          float(G), vp(out_ptr)))

    def accel_targets_dev_noalloc(self, tgt_ptr: int, m: int, out_ptr: int, theta: float, softening: float,
                                  G: float) -> None:
        """accel_targets_dev without device allocation (at most 2**20 targets
        per call; calls on one device must not overlap)."""
        check(load().bh_accel_targets_dev_noalloc(self.handle, vp(tgt_ptr), int(m), float(theta),
                                                  float(softening), float(G), vp(out_ptr)))

    def accel_f32_self_dev(self, out_ptr: int, theta: float, softening: float, G: float,
                           opts: F32Opts | None = None) -> np.ndarray:
        st = np.zeros(4, np.int64)
        check(load().bh_accel_f32_self_dev(self.handle, float(theta), float(softening), float(G),
                                           ctypes.byref(opts) if opts is not None else None, vp(out_ptr), ptr(st)))
        return st

    def accel_f32_targets_dev(self, tgt_ptr: int, m: int, out_ptr: int, theta: float, softening: float, G: float,
                              opts: F32Opts | None = None) -> np.ndarray:
        st = np.zeros(4, np.int64)
        check(load().bh_accel_f32_targets_dev(self.handle, vp(tgt_ptr), int(m), float(theta), float(softening),
                                              float(G), ctypes.byref(opts) if opts is not None else None,
                                              vp(out_ptr), ptr(st)))
        return st

    def refit_dev(self, pos_ptr: int, mass_ptr: int | None) -> None:
        check(load().bh_tree_refit_dev(self.handle, vp(pos_ptr), vp(mass_ptr) if mass_ptr else None))

    def accel_targets(self, targets: np.ndarray, theta: float, softening: float, G: float) -> np.ndarray:
        tgt = _rows(targets, "targets", None, self.dim)
        out = np.empty_like(tgt)
        if len(tgt):
            check(load().bh_accel_targets(self.handle, ptr(tgt), len(tgt), float(theta), float(softening),
                                          float(G), ptr(out)))
        return out

    def accel_targets_streamed(self, targets: np.ndarray, theta: float, softening: float, G: float, *,
                               precision: str = "fp64", chunk: int = 0, opts: F32Opts | None = None) -> np.ndarray:
        """Accelerations at host targets streamed through the device in
        chunks of `chunk` targets on two streams (0: 2**20); bit-identical
        to accel_targets / accel_f32 on all targets at once."""
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")
        tgt = _rows(targets, "targets", None, self.dim)
        out = np.empty_like(tgt)
        if len(tgt):
            check(load().bh_accel_targets_streamed(self.handle, ptr(tgt), len(tgt), int(chunk), float(theta),
                                                   float(softening), float(G), 64 if precision == "fp64" else 32,
                                                   ctypes.byref(opts) if opts is not None else None, ptr(out)))
        return out

    def accel_f32(self, targets, theta: float, softening: float, G: float, opts: F32Opts | None = None,
                  stats: bool = False):
        st = np.zeros(4, np.int64)
        op = ctypes.byref(opts) if opts is not None else None
        if targets is None:
            out = np.empty((self.n, self.dim))
            check(load().bh_accel_f32_self(self.handle, float(theta), float(softening), float(G), op, ptr(out),
                                           ptr(st)))
        else:
            tgt = _rows(targets, "targets", None, self.dim)
            out = np.empty_like(tgt)
            if len(tgt):
                check(load().bh_accel_f32_targets(self.handle, ptr(tgt), len(tgt), float(theta),
                                                  float(softening), float(G), op, ptr(out), ptr(st)))
        return (out, st) if stats else out

    def potential(self, targets, theta: float, softening: float, G: float) -> np.ndarray:
        """Softened potential at the sources (targets None) or at targets."""
        tgt = None if targets is None else _rows(targets, "targets", None, self.dim)
        m = self.n if tgt is None else len(tgt)
        out = np.empty(m)
        if m:
            check(load().bh_potential(self.handle, ptr(tgt), m, float(theta), float(softening), float(G), ptr(out)))
        return out

    def potential_dev(self, tgt_ptr: int | None, m: int, out_ptr: int, theta: float, softening: float,
                      G: float) -> None:
        check(load().bh_potential_dev(self.handle, vp(tgt_ptr) if tgt_ptr else None, int(m), float(theta),
                                      float(softening), float(G), vp(out_ptr)))

    def interaction_lists(self, targets, theta: float):
        """(ptr, ids) of per-target interaction lists in pop order."""
        tgt = None if targets is None else _rows(targets, "targets", None, self.dim)
        m = self.n if tgt is None else len(tgt)
        p = np.zeros(m + 1, np.int64)
        total = i64()
        check(load().bh_interaction_lists(self.handle, ptr(tgt), m, float(theta), ptr(p), None, 0,
                                          ctypes.byref(total)))
        ids = np.zeros(max(total.value, 1), np.int32)
        check(load().bh_interaction_lists(self.handle, ptr(tgt), m, float(theta), ptr(p), ptr(ids),
                                          total.value, ctypes.byref(total)))
        return p, ids[:total.value].astype(np.int64)

    def counters(self, targets, theta: float) -> np.ndarray:
        tgt = None if targets is None else _rows(targets, "targets", None, self.dim)
        m = self.n if tgt is None else len(tgt)
        out = np.zeros((m, 5), np.int64)
        check(load().bh_counters(self.handle, ptr(tgt), m, float(theta), ptr(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.bh_tree_free(self.handle)
                self.handle = None
        except Exception:
            pass


def direct_dev(src_ptr: int, mass_ptr: int, n: int, tgt_ptr: int, m: int, dim: int, self_mode: bool,
               softening: float, G: float, acc_ptr: int | None, pot_ptr: int | None, context=None) -> None:
    c = context or ctx()
    check(load().bh_direct_dev(c.handle, vp(src_ptr), vp(mass_ptr), int(n), vp(tgt_ptr), int(m), int(dim),
                               int(self_mode), float(softening), float(G), vp(acc_ptr) if acc_ptr else None,
                               vp(pot_ptr) if pot_ptr else None))


def direct(src, mass, targets=None, *, softening=0.0, G=1.0, potential=False, context=None):
    """FP64 direct sum on the device: accelerations (and potential)."""
    c = context or ctx()
    s, m = f64c(src), f64c(mass)
    self_mode = targets is None
    t = s if self_mode else f64c(targets)
    acc = np.empty_like(t)
    pot = np.empty(len(t)) if potential else None
    check(load().bh_direct(c.handle, ptr(s), ptr(m), len(s), ptr(t), len(t), s.shape[1], int(self_mode),
                           float(softening), float(G), ptr(acc), ptr(pot)))
    return (acc, pot) if potential else acc


class Simulation:
    """Owner of a bh_sim*: KDK leapfrog over device buffers (pointers of
    float64 (n, dim) positions / velocities and (n,) masses, which must stay
    alive while this object lives)."""

    INFO_KEYS = ("steps", "rebuilds", "refits", "graph_launches", "graph_captures", "list_regrows",
                 "list_region", "item_capacity")

    def __init__(self, pos_ptr: int, vel_ptr: int, mass_ptr: int, n: int, dim: int, *, leaf_size: int = 32,
                 theta: float = 0.5, softening: float = 0.0, G: float = 1.0, precision: int = 32,
                 opts: F32Opts | None = None, context: Context | None = None):
        self.context = context or ctx()
        h = vp()
        check(load().bh_sim_create(self.context.handle, vp(pos_ptr), vp(vel_ptr), vp(mass_ptr), int(n), int(dim),
                                   int(leaf_size), float(theta), float(softening), float(G), int(precision),
                                   ctypes.byref(opts) if opts is not None else None, ctypes.byref(h)))
        self.handle = h

    def run(self, dt: float, steps: int, rebuild_every: int = 0, use_graph: bool = True) -> None:
        check(load().bh_sim_run(self.handle, float(dt), int(steps), int(rebuild_every), int(bool(use_graph))))

    def accelerations_dev(self, out_ptr: int) -> None:
        check(load().bh_sim_accelerations(self.handle, vp(out_ptr)))

    def set_list_capacity(self, region: int, items: int) -> None:
        """FP32 list capacities of captured steps: entries per group of each
        list region, and work items."""
        check(load().bh_sim_set_list_capacity(self.handle, int(region), int(items)))

    def info(self) -> dict:
        out = np.zeros(8, np.int64)
        check(load().bh_sim_info(self.handle, ptr(out)))
        return dict(zip(self.INFO_KEYS, (int(v) for v in out)))

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.bh_sim_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
