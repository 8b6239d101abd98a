"""ctypes binding of ``libparticular_amd.so`` (the C ABI declared in include/particular_amd.h).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparticular_amd.so")

PNB_OK = 0
PNB_E_ARG = 1
PNB_E_NODEV = 2
PNB_E_OOM = 3
PNB_E_HIP = 4
PNB_E_DEPTH = 5
PNB_E_STATE = 6

PROF_BF = 0
PROF_BH_BUILD = 1
PROF_BH_WALK = 2
PROF_DRIFT = 3
PROF_BATCH = 4
PROF_MLP = 5
PROF_SORT = 6

BH_MLP = 1
BH_REF_COM = 2
BH_RELATIVE = 4

_c = ctypes
_vp = _c.c_void_p
_sz = _c.c_size_t
_i = _c.c_int
_f = _c.c_float
_d = _c.c_double
_u32 = _c.c_uint32
_u64 = _c.c_uint64
_pi = _c.POINTER(_c.c_int)
_pd = _c.POINTER(_c.c_double)
_pu64 = _c.POINTER(_c.c_uint64)

# name -> (restype, argtypes). Pointers to particle data are passed as void* so that host numpy
# buffers and device addresses share one signature.
SIGNATURES = {
    "pnb_device_count": (_i, [_pi]),
    "pnb_create": (_i, [_i, _c.POINTER(_vp)]),
    "pnb_destroy": (None, [_vp]),
    "pnb_last_error": (_c.c_char_p, [_vp]),
    "pnb_status_string": (_c.c_char_p, [_i]),
    "pnb_set_stream": (_i, [_vp, _vp]),
    "pnb_reset_stream": (_i, [_vp]),
    "pnb_get_stream": (_vp, [_vp]),
    "pnb_synchronize": (_i, [_vp]),
    "pnb_profile_enable": (_i, [_vp, _i]),
    "pnb_profile_read": (_i, [_vp, _i, _pd, _pu64]),
    "pnb_bf_accel_f32": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_accel_f32_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_accel_f32_dev_noalloc": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _i, _vp]),
    "pnb_reordered_bf_accel_f32": (_i, [_vp, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_pairs_accel_f32": (_i, [_vp, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_pairs_accel_f32_dev": (_i, [_vp, _vp, _sz, _f, _i, _vp]),
    "pnb_reordered_bf_accel_f32_dev": (_i, [_vp, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_set_tuning": (_i, [_vp, _sz, _i, _i, _i]),
    "pnb_bf_clear_tuning": (_i, [_vp]),
    "pnb_direct_f64_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _d, _vp]),
    "pnb_bf_accel_f64": (_i, [_vp, _vp, _sz, _vp, _sz, _d, _i, _vp]),
    "pnb_bf_accel_f64_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _d, _i, _vp]),
    "pnb_bf_accel_f32_2d": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _i, _vp]),
    "pnb_bf_accel_f32_2d_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _i, _vp]),
    "pnb_bh_accel_f32_2d": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_bh_accel_f32_2d_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_bh_accel_f32": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_bh_set_capturable": (_i, [_vp, _i]),
    "pnb_bh_check": (_i, [_vp]),
    "pnb_bh_accel_f32_dev": (_i, [_vp, _vp, _sz, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_reordered_bh_accel_f32": (_i, [_vp, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_reordered_bh_accel_f32_dev": (_i, [_vp, _vp, _sz, _f, _f, _i, _u32, _vp]),
    "pnb_bh_tree_size": (_i, [_vp, _c.POINTER(_sz), _pi]),
    "pnb_bh_tree_export": (_i, [_vp, _vp, _sz, _vp, _sz]),
    "pnb_bh_node_features": (_i, [_vp, _vp, _sz]),
    "pnb_bh_stats": (_i, [_vp, _pu64, _pu64]),
    "pnb_bh_set_relative": (_i, [_vp, _f, _f, _i, _i]),
    "pnb_batch_set_tuning": (_i, [_vp, _sz, _i]),
    "pnb_batch_init": (_i, [_vp, _vp, _vp, _vp, _vp, _sz, _f, _i]),
    "pnb_batch_run": (_i, [_vp, _sz, _d, _i]),
    "pnb_batch_read": (_i, [_vp, _vp, _vp, _vp]),
    "pnb_batch_info": (_i, [_vp, _c.POINTER(_sz), _c.POINTER(_sz), _c.POINTER(_c.c_ulonglong)]),
    "pnb_mlp_num_params": (_i, []),
    "pnb_mlp_init": (_i, [_vp, _u64]),
    "pnb_mlp_set_params": (_i, [_vp, _vp, _sz]),
    "pnb_mlp_get_params": (_i, [_vp, _vp, _sz]),
    "pnb_mlp_collect": (_i, [_vp, _vp, _sz, _f, _f, _sz, _u64, _sz]),
    "pnb_mlp_clear_dataset": (_i, [_vp]),
    "pnb_mlp_dataset_info": (_i, [_vp, _c.POINTER(_sz), _c.POINTER(_sz), _pu64]),
    "pnb_mlp_dataset_read": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pnb_mlp_dataset_load": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _vp, _vp, _vp]),
    "pnb_mlp_grad": (_i, [_vp, _vp, _vp, _sz]),
    "pnb_mlp_train": (_i, [_vp, _sz, _f, _vp]),
    "pnb_mlp_forward_dev": (_i, [_vp, _vp, _sz, _vp]),
    "pnb_leapfrog_init": (_i, [_vp, _vp, _vp, _vp, _sz, _f, _i]),
    "pnb_leapfrog_run": (_i, [_vp, _sz, _d, _i]),
    "pnb_leapfrog_read": (_i, [_vp, _vp, _vp, _vp]),
    "pnb_leapfrog_views": (_i, [_vp, _c.POINTER(_vp), _c.POINTER(_vp), _c.POINTER(_vp),
                                _c.POINTER(_sz), _c.POINTER(_sz)]),
}

TREE_NODE_DTYPE = None  # numpy dtype of pnb_tree_node, see tree_node_dtype()


def tree_node_dtype():
    import numpy as np

    return np.dtype([("com", "<f4", (3,)), ("mass", "<f4"), ("mass64", "<f8"), ("width", "<f4"),
                     ("skip", "<u4"), ("begin", "<u4"), ("count", "<u4"), ("level", "<u4"),
                     ("internal", "<u4"), ("quad", "<f4", (6,))], align=True)


_lib = None


class PnbError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it is missing (build with ``python -m particular_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: the HIP extension has not been built "
                "(run `python -m particular_amd.build`); there is no CPU fallback")
        # torch ships its own copy of the HIP runtime under the same soname
        # (libamdhip64.so.7). Importing torch first lets the loader bind this library to that
        # copy, so the process holds one HIP runtime and both can use the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, ctx=None) -> None:
    if status != PNB_OK:
        L = lib()
        msg = L.pnb_last_error(ctx).decode() if ctx else ""
        if not msg:
            msg = L.pnb_status_string(status).decode()
        raise PnbError(status, msg)
