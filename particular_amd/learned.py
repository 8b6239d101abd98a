"""Host side of the learned far-field correction (MLP on top of Barnes-Hut).

Thin wrappers over the pnb_mlp_* entry points: parameters, training-set collection (rows and
brute-force-minus-Barnes-Hut targets generated on the GPU), full-batch training, and export of the
training set for checks against the CPU oracle.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

FEAT = 16  # stored feature stride
IN = 16    # features the network reads
OUT = 13   # outputs: quadrupole {xx, yy, zz, xy, xz, yz} (a residual on features 0..5), then
QOUT = 6   # octupole {xxx, yyy, zzz, xxy, xxz, xyy, xyz} (features 9..15 hold 32 x the node's own)
OCT_FEATURE_SCALE = 32.0
OCT_MIN = 3e-3  # the walk adds a node's octupole only when its largest coefficient reaches this


def walk_outputs(y: np.ndarray) -> np.ndarray:
    """The network outputs as the corrected walk applies them: octupole coefficients of a node
    whose largest one is below OCT_MIN are left out (set to zero here)."""
    y = np.array(y, np.float64, copy=True)
    small = np.abs(y[:, QOUT:OUT]).max(1) < OCT_MIN
    y[small, QOUT:OUT] = 0.0
    return y
YS = 16    # stored output stride
GEOM = 8   # per row: centre of mass, M w^2, M w^3, three zeros


def num_params() -> int:
    return int(_lib.lib().pnb_mlp_num_params())


def init(res, seed: int = 0) -> None:
    res.check(_lib.lib().pnb_mlp_init(res.handle, seed))


def set_params(res, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, np.float32).ravel()
    res.check(_lib.lib().pnb_mlp_set_params(res.handle, p.ctypes.data, p.size))


def get_params(res) -> np.ndarray:
    p = np.zeros(num_params(), np.float32)
    res.check(_lib.lib().pnb_mlp_get_params(res.handle, p.ctypes.data, p.size))
    return p


def params_to_dict(p: np.ndarray) -> dict:
    """Splits the flat vector into the oracle's W1, b1, W2, b2, W3, b3 (FP64)."""
    p = np.asarray(p, np.float64)
    shapes = [("W1", (IN, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, OUT)),
              ("b3", (OUT,))]
    out, o = {}, 0
    for name, shp in shapes:
        k = int(np.prod(shp))
        out[name] = p[o:o + k].reshape(shp).copy()
        o += k
    return out


def dict_to_params(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[k], np.float64).ravel()
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3")]).astype(np.float32)


def collect(res, particles_dev, theta: float, waves: int, seed: int, softening: float = 0.0,
            row_cap: int = 4096) -> None:
    """Appends the rows and targets of `waves` seeded wavefronts of one snapshot (device tensor)."""
    res.use_torch_stream()
    res.check(_lib.lib().pnb_mlp_collect(res.handle, particles_dev.data_ptr(),
                                         particles_dev.shape[0], theta, softening, waves, seed,
                                         row_cap))
    res.synchronize()


def clear(res) -> None:
    res.check(_lib.lib().pnb_mlp_clear_dataset(res.handle))


def info(res) -> tuple[int, int, int]:
    r, w, d = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint64()
    res.check(_lib.lib().pnb_mlp_dataset_info(res.handle, ctypes.byref(r), ctypes.byref(w),
                                              ctypes.byref(d)))
    return r.value, w.value, d.value


def dataset(res) -> dict:
    rows, waves, _ = info(res)
    d = {
        "feats": np.zeros((rows, FEAT), np.float32), "geom": np.zeros((rows, GEOM), np.float32),
        "mask": np.zeros(rows, np.uint64), "row_wave": np.zeros(rows, np.uint32),
        "wave_rows": np.zeros(waves + 1, np.uint32), "lanes": np.zeros((waves, 64, 4), np.float32),
        "target": np.zeros((waves, 64, 4), np.float32), "wgt": np.zeros((waves, 64), np.float32),
    }
    res.check(_lib.lib().pnb_mlp_dataset_read(
        res.handle, d["feats"].ctypes.data, d["geom"].ctypes.data, d["mask"].ctypes.data,
        d["row_wave"].ctypes.data, d["wave_rows"].ctypes.data, d["lanes"].ctypes.data,
        d["target"].ctypes.data, d["wgt"].ctypes.data))
    return d


_DATASET_DTYPES = {"feats": np.float32, "geom": np.float32, "mask": np.uint64,
                   "row_wave": np.uint32, "wave_rows": np.uint32, "lanes": np.float32,
                   "target": np.float32, "wgt": np.float32}


def load(res, d: dict) -> None:
    """Appends a training set held in host arrays (the layout `dataset` returns)."""
    # the converted arrays are kept referenced until the call returns
    f = {k: np.ascontiguousarray(d[k], dt) for k, dt in _DATASET_DTYPES.items()}
    if f["feats"].ndim != 2 or f["feats"].shape[1] != FEAT:
        raise ValueError(f"feats must be (rows, {FEAT})")
    if (f["target"].shape[1:] != (64, 4) or f["lanes"].shape[1:] != (64, 4)
            or f["wgt"].shape[1:] != (64,)):
        raise ValueError("lanes and target must be (waves, 64, 4) and wgt (waves, 64)")
    rows, waves = f["feats"].shape[0], f["wave_rows"].shape[0] - 1
    if not (f["geom"].shape == (rows, GEOM) and f["mask"].size == f["row_wave"].size == rows
            and f["target"].shape[0] == f["wgt"].shape[0] == f["lanes"].shape[0] == waves):
        raise ValueError("inconsistent training-set sizes")
    res.check(_lib.lib().pnb_mlp_dataset_load(
        res.handle, f["feats"].ctypes.data, f["geom"].ctypes.data, f["mask"].ctypes.data,
        f["row_wave"].ctypes.data, f["wave_rows"].ctypes.data, rows, waves,
        f["lanes"].ctypes.data, f["target"].ctypes.data, f["wgt"].ctypes.data))


def grad(res) -> tuple[float, np.ndarray]:
    loss = ctypes.c_float()
    g = np.zeros(num_params(), np.float32)
    res.check(_lib.lib().pnb_mlp_grad(res.handle, ctypes.byref(loss), g.ctypes.data, g.size))
    return loss.value, g


def train(res, iterations: int, lr: float = 1e-3) -> np.ndarray:
    hist = np.zeros(iterations, np.float32)
    res.check(_lib.lib().pnb_mlp_train(res.handle, iterations, lr, hist.ctypes.data))
    return hist


def quadrupole_params() -> np.ndarray:
    """Parameters with a zero output layer: the network adds nothing to its residual, the node's
    own normalised quadrupole (features 0..5), so every accepting lane adds the quadrupole term of
    the node's multipole expansion. pnb_mlp_init gives the same outputs (random hidden layers,
    zero output layer) and is where training starts."""
    return np.zeros(num_params(), np.float32)


def octupole_params() -> np.ndarray:
    """Parameters that make the network's octupole outputs (6..12) equal the node's own normalised
    octupole (features 9..15 over 32) and add nothing to the quadrupole residual: each of the seven
    features passes through two hidden units as relu(+f) and relu(-f) (weights 0, +-1 and +-1/32,
    exact in bf16).
    The walk then adds the multipole expansion's quadrupole and octupole terms: the analytic
    reference a trained network is measured against."""
    d = params_to_dict(np.zeros(num_params(), np.float32))
    for j in range(OUT - QOUT):
        d["W1"][9 + j, 2 * j] = 1.0
        d["W1"][9 + j, 2 * j + 1] = -1.0
        d["W2"][2 * j, 2 * j] = 1.0
        d["W2"][2 * j + 1, 2 * j + 1] = 1.0
        d["W3"][2 * j, QOUT + j] = 1.0 / OCT_FEATURE_SCALE
        d["W3"][2 * j + 1, QOUT + j] = -1.0 / OCT_FEATURE_SCALE
    return dict_to_params(d)


TRAINED_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params",
                              "two_plummer_4M.npy")


def trained_params() -> np.ndarray | None:
    """The network trained by particular_amd.nn_bh on two-Plummer snapshots (the file written by
    its --params-out), or None when that file is absent or holds another network's size."""
    if not os.path.exists(TRAINED_PARAMS):
        return None
    p = np.load(TRAINED_PARAMS, allow_pickle=False).astype(np.float32).ravel()
    return p if p.size == num_params() else None


def default_params() -> np.ndarray:
    """The parameters the corrected walk uses unless told otherwise: the trained network when
    it is present, else the quadrupole term."""
    p = trained_params()
    return p if p is not None else quadrupole_params()


def fit(res, iterations: int, lrs=(3e-3, 1e-3, 3e-4, 1e-4), probe: int = 150,
        chunk: int = 250, log=None) -> dict:
    """Full-batch training from the current parameters with the learning rate chosen by probes.

    Every candidate in `lrs` runs `probe` Adam steps from the same start (short probes favour
    small rates that lose later: on two-Plummer sets 40 steps picked 3e-4, which ended at 1.7x
    the loss 3e-3 reaches); the one with the lowest loss then runs `iterations` steps in three
    phases (lr for half of them, lr/10 and
    lr/100 for a quarter each), `chunk` steps per library call with `log(message)` after each
    when given. If the result fits the training set worse than the starting parameters, those are
    restored (a fresh network is the quadrupole term, the fallback). Returns the loss history,
    the chosen rate and the start / end losses."""
    p0 = get_params(res)
    loss0, _ = grad(res)
    best = None
    for lr in lrs:
        set_params(res, p0)
        h = train(res, probe, lr)
        last, _ = grad(res)
        if np.isfinite(last) and (best is None or last < best[1]):
            best = (lr, last)
    set_params(res, p0)
    if best is None:
        return {"history": np.zeros(0, np.float32), "lr": None, "loss_start": loss0,
                "loss_end": loss0, "kept": False}
    lr = best[0]
    phases = [(iterations // 2, lr), (iterations // 4, lr / 10),
              (iterations - iterations // 2 - iterations // 4, lr / 100)]
    parts = []
    done = 0
    for it, r in phases:
        while it > 0:
            k = min(it, max(chunk, 1))
            parts.append(train(res, k, r))
            it -= k
            done += k
            if log:
                log(f"training: {done} of {iterations} steps, lr {r:.1e}, loss {parts[-1][-1]:.4e}")
    hist = np.concatenate(parts or [np.zeros(0)])
    loss1, _ = grad(res)
    kept = bool(np.isfinite(loss1) and loss1 < loss0)
    if not kept:
        set_params(res, p0)
    return {"history": hist, "lr": lr, "loss_start": loss0, "loss_end": loss1 if kept else loss0,
            "kept": kept}


def node_features(res) -> np.ndarray:
    """(nodes, FEAT) network inputs of every node of the last octree built with moments
    (pnb_bh_node_features), pre-order; zeros for leaves and massless nodes."""
    L = _lib.lib()
    n = ctypes.c_size_t()
    lvl = ctypes.c_int()
    res.check(L.pnb_bh_tree_size(res.handle, ctypes.byref(n), ctypes.byref(lvl)))
    out = np.zeros((n.value, FEAT), np.float32)
    res.check(L.pnb_bh_node_features(res.handle, out.ctypes.data if n.value else None, n.value))
    return out


def forward(res, feats_dev, y_dev) -> None:
    """y_dev (rows x YS) = feats_dev[:, :6] + MLP(feats_dev (rows x FEAT)); device tensors."""
    res.use_torch_stream()
    res.check(_lib.lib().pnb_mlp_forward_dev(res.handle, feats_dev.data_ptr(), feats_dev.shape[0],
                                             y_dev.data_ptr()))
