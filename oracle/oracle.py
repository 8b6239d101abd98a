"""ctypes view of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import this
module, and only as the checker or the timed CPU baseline. The product path never does.

The library restates particular's f32 ``BruteForce``, octree build and ``BarnesHut``
(particular/src/sequential.rs:181-194, 456-506; particular/src/tree/mod.rs:73-139); each C
function cites its reference lines in ``oracle.c``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcNode(ctypes.Structure):
    _fields_ = [
        ("child", ctypes.c_int32 * 8),
        ("internal", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("begin", ctypes.c_int32),
        ("count", ctypes.c_int32),
        ("bmin", ctypes.c_float * 3),
        ("bmax", ctypes.c_float * 3),
        ("com", ctypes.c_float * 3),
        ("mass", ctypes.c_float),
        ("com64", ctypes.c_double * 3),
        ("mass64", ctypes.c_double),
    ]


NODE_DTYPE = np.dtype(
    [
        ("child", "<i4", (8,)),
        ("internal", "<i4"),
        ("level", "<i4"),
        ("begin", "<i4"),
        ("count", "<i4"),
        ("bmin", "<f4", (3,)),
        ("bmax", "<f4", (3,)),
        ("com", "<f4", (3,)),
        ("mass", "<f4"),
        ("com64", "<f8", (3,)),
        ("mass64", "<f8"),
    ],
    align=True,
)

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        sz = ctypes.c_size_t
        L.orc_accel_pair.argtypes = [fp, fp, ctypes.c_float, ctypes.c_int, fp]
        L.orc_bf_f32.argtypes = [fp, sz, fp, sz, ctypes.c_float, ctypes.c_int, fp, ctypes.c_int]
        L.orc_bf_pairs_f32.argtypes = [fp, sz, sz, ctypes.c_float, ctypes.c_int, fp]
        L.orc_direct_f64.argtypes = [fp, sz, fp, sz, ctypes.c_double, dp, ctypes.c_int]
        L.orc_tree_build.argtypes = [fp, sz, ctypes.c_int]
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_num_nodes.argtypes = [ctypes.c_void_p]
        L.orc_tree_num_nodes.restype = sz
        L.orc_tree_max_level.argtypes = [ctypes.c_void_p]
        L.orc_tree_nodes.argtypes = [ctypes.c_void_p]
        L.orc_tree_nodes.restype = ctypes.POINTER(OrcNode)
        L.orc_tree_order.argtypes = [ctypes.c_void_p]
        L.orc_tree_order.restype = ctypes.POINTER(ctypes.c_uint32)
        L.orc_tree_root_box.argtypes = [ctypes.c_void_p, fp, fp]
        L.orc_tree_skips.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_tree_skips.restype = None
        L.orc_bh_f32.argtypes = [
            ctypes.c_void_p, fp, sz, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, fp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int,
        ]
        L.orc_bh_rel_f32.argtypes = [
            ctypes.c_void_p, fp, sz, ctypes.c_float, ctypes.c_float, fp, ctypes.c_int,
            ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, fp, ctypes.POINTER(ctypes.c_uint32),
            ctypes.c_int,
        ]
        _lib = L
    return _lib


def _f32x4(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("expected an (n, 4) float32 array")
    return a


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def accel_pair(p2, at, softening=0.0, checked=True) -> np.ndarray:
    p2 = np.ascontiguousarray(p2, dtype=np.float32)
    at = np.ascontiguousarray(at, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_accel_pair(_fp(p2), _fp(at), softening, int(checked), _fp(out))
    return out


def brute_force(affected, affecting, softening=0.0, checked=True, threads=0) -> np.ndarray:
    """f32 sequential::BruteForce; returns (n, 4) float32 with w = 0."""
    a = _f32x4(affected)
    b = _f32x4(affecting)
    out = np.zeros((a.shape[0], 4), np.float32)
    rc = lib().orc_bf_f32(_fp(a), a.shape[0], _fp(b), b.shape[0], softening, int(checked),
                          _fp(out), threads)
    if rc:
        raise RuntimeError(f"orc_bf_f32 failed ({rc})")
    return out


def brute_force_pairs(particles, restrict_len=None, softening=0.0, checked=False) -> np.ndarray:
    """f32 sequential::BruteForcePairs (RestrictedBruteForce over the first restrict_len
    particles, default all); returns (n, 4) float32 with w = 0. checked=False is the reference."""
    a = _f32x4(particles)
    n = a.shape[0]
    out = np.zeros((n, 4), np.float32)
    rc = lib().orc_bf_pairs_f32(_fp(a), n, n if restrict_len is None else restrict_len,
                                softening, int(checked), _fp(out))
    if rc:
        raise RuntimeError(f"orc_bf_pairs_f32 failed ({rc})")
    return out


def direct_f64(affected, affecting, softening=0.0, threads=0) -> np.ndarray:
    """FP64 direct sum; returns (n, 3) float64."""
    a = _f32x4(affected)
    b = _f32x4(affecting)
    out = np.zeros((a.shape[0], 3), np.float64)
    rc = lib().orc_direct_f64(_fp(a), a.shape[0], _fp(b), b.shape[0], softening,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), threads)
    if rc:
        raise RuntimeError(f"orc_direct_f64 failed ({rc})")
    return out


class Tree:
    """The restated RootedOrthtree; nodes in the reference's pre-order."""

    def __init__(self, affecting, max_depth: int = 4096):
        self._src = _f32x4(affecting)
        self._h = lib().orc_tree_build(_fp(self._src), self._src.shape[0], max_depth)
        if not self._h:
            raise RuntimeError("orc_tree_build failed (allocation or depth limit)")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_tree_free(h)
            self._h = None

    @property
    def num_nodes(self) -> int:
        return int(lib().orc_tree_num_nodes(self._h))

    @property
    def max_level(self) -> int:
        return int(lib().orc_tree_max_level(self._h))

    def nodes(self) -> np.ndarray:
        n = self.num_nodes
        if n == 0:
            return np.zeros(0, NODE_DTYPE)
        assert ctypes.sizeof(OrcNode) == NODE_DTYPE.itemsize
        ptr = lib().orc_tree_nodes(self._h)
        buf = (ctypes.c_char * (n * NODE_DTYPE.itemsize)).from_address(ctypes.addressof(ptr.contents))
        return np.frombuffer(buf, dtype=NODE_DTYPE).copy()  # one copy out of the C array

    def skips(self) -> np.ndarray:
        """Per node, the index of the first node after its subtree (pre-order)."""
        out = np.zeros(self.num_nodes, np.uint32)
        if out.size:
            lib().orc_tree_skips(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out

    def order(self) -> np.ndarray:
        m = self._src.shape[0]
        if m == 0:
            return np.zeros(0, np.uint32)
        ptr = lib().orc_tree_order(self._h)
        return np.ctypeslib.as_array(ptr, shape=(m,)).copy()

    def root_box(self):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        lib().orc_tree_root_box(self._h, _fp(mn), _fp(mx))
        return mn, mx

    def barnes_hut(self, affected, theta, softening=0.0, checked=True, com64=False, threads=0,
                   return_visits=False, accum64=False):
        """Traversal of sequential::BarnesHut. com64: node data from the FP64 reduction;
        accum64: same decisions, pair terms summed in FP64."""
        a = _f32x4(affected)
        out = np.zeros((a.shape[0], 4), np.float32)
        visits = np.zeros(a.shape[0], np.uint32)
        rc = lib().orc_bh_f32(self._h, _fp(a), a.shape[0], theta, softening, int(checked),
                              int(com64), int(accum64), _fp(out),
                              visits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), threads)
        if rc:
            raise RuntimeError(f"orc_bh_f32 failed ({rc})")
        return (out, visits) if return_visits else out

    def barnes_hut_rel(self, affected, theta, est, order=2, softening=0.0, com64=True,
                       threads=0, return_visits=False, accum64=False, bucket=0):
        """The traversal with the GPU backend's relative criterion (no reference counterpart):
        est[i] = alpha |a~_i| per affected particle, order 2 (plain walk) or 3 (with the learned
        correction), opened nodes of at most `bucket` particles summed particle by particle; see
        orc_bh_rel_f32."""
        a = _f32x4(affected)
        e = np.ascontiguousarray(est, dtype=np.float32)
        if e.shape != (a.shape[0],):
            raise ValueError("est needs one value per affected particle")
        out = np.zeros((a.shape[0], 4), np.float32)
        visits = np.zeros(a.shape[0], np.uint32)
        rc = lib().orc_bh_rel_f32(self._h, _fp(a), a.shape[0], theta, softening, _fp(e),
                                  int(order), int(bucket), _fp(self._src), int(com64),
                                  int(accum64), _fp(out),
                                  visits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), threads)
        if rc:
            raise RuntimeError(f"orc_bh_rel_f32 failed ({rc})")
        return (out, visits) if return_visits else out


def barnes_hut(affected, affecting, theta, softening=0.0, checked=True, com64=False, threads=0):
    """sequential::BarnesHut over Between(affected, affecting): tree from `affecting`."""
    return Tree(affecting).barnes_hut(affected, theta, softening, checked, com64, threads)
