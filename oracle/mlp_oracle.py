"""FP64 CPU oracle of the far-field correction MLP -- TEST INFRASTRUCTURE ONLY.

No reference counterpart (particular has no learned component). The network and its loss are
defined here in plain numpy and pinned by finite differences (tests/test_mlp_oracle.py); the GPU
forward/backward kernels are checked against it.

Network: 16 -> 64 -> 64 -> 13, ReLU, y = relu(relu(x W1 + b1) W2 + b2) W3 + b3 plus x[:, :6] on
y[:, :6] (a residual on the node's normalised quadrupole, inputs 0..5). When ``bf16=True`` the
inputs of every matrix product (x, h1, h2, g_y, g_z2, g_z1 and W1..W3) are rounded to bfloat16
(round to nearest even) as the MFMA kernels round them; the residual and all sums are FP64. The
residual does not depend on the parameters, so it leaves every gradient formula as it is.

A row is an internal node accepted by some lanes of a wavefront: row r belongs to wave
``wave[r]``, applies to the lanes set in ``mask[r]`` (64-bit), and has geometry ``geom[r]`` =
(centre of mass, M w^2, M w^3, ...). Its output y[r] holds a second moment Q = M w^2 y[r, :6]
({xx, yy, zz, xy, xz, yz}) and a traceless third moment O = M w^3 y[r, 6:13] ({xxx, yyy, zzz,
xxy, xxz, xyy, xyz}); lane l of wave w at ``lanes[w, l, :3]`` (softening^2 in ``lanes[w, l, 3]``)
predicts ``pred[w, l] = sum_{r in w, l in mask[r]} field(Q_r, O_r, d)`` with d = centre - lane,
field = -Q d |d|^-5 + 2.5 (d.Q.d) d |d|^-7 + 3 F |d|^-7 - 7 (d.F) d |d|^-9, F_i = O_ijk d_j d_k,
for target ``t[w, l]`` with weight ``wgt[w, l]``; loss = sum wgt * |pred - t|^2 / (number of
lanes with wgt > 0).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 16, 64, 13
QOUT = 6  # quadrupole outputs; the rest are octupole coefficients


def bf16_round(a) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)

    def he(fan_in, shape):
        return rng.normal(0.0, np.sqrt(2.0 / fan_in), shape)

    return {
        "W1": he(IN, (IN, HID)), "b1": np.zeros(HID),
        "W2": he(HID, (HID, HID)), "b2": np.zeros(HID),
        "W3": he(HID, (HID, OUT)) * 0.1, "b3": np.zeros(OUT),
    }


def forward(p: dict, x: np.ndarray, bf16: bool = False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    z1 = r(x) @ r(p["W1"]) + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = r(h1) @ r(p["W2"]) + p["b2"]
    h2 = np.maximum(z2, 0.0)
    y = r(h2) @ r(p["W3"]) + p["b3"]
    y[:, :QOUT] += np.asarray(x, np.float64)[:, :QOUT]
    return y, (z1, h1, z2, h2)


def lane_bits(mask: np.ndarray) -> np.ndarray:
    """(rows, 64) bool matrix of the lanes set in each 64-bit mask."""
    m = np.asarray(mask, np.uint64)
    return ((m[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)


def field_basis(d: np.ndarray, inv: np.ndarray) -> np.ndarray:
    """B (..., 3, 6) with field(Q, d) = B @ q for q = {xx, yy, zz, xy, xz, yz}; inv = 1 / |d|
    (softened)."""
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    inv2 = inv * inv
    inv5 = inv2 * inv2 * inv
    z = np.zeros_like(dx)
    c = np.stack([dx * dx, dy * dy, dz * dz, 2 * dx * dy, 2 * dx * dz, 2 * dy * dz], -1)
    m = np.stack([np.stack([dx, z, z, dy, dz, z], -1),
                  np.stack([z, dy, z, dx, z, dz], -1),
                  np.stack([z, z, dz, z, dx, dy], -1)], -2)
    return inv5[..., None, None] * (2.5 * inv2[..., None, None] * d[..., :, None] * c[..., None, :] - m)


def oct_field_basis(d: np.ndarray, inv: np.ndarray) -> np.ndarray:
    """B (..., 3, 7) with the octupole field 3 F |d|^-7 - 7 (d.F) d |d|^-9 = B @ o for the seven
    coefficients o = {xxx, yyy, zzz, xxy, xxz, xyy, xyz} of a traceless symmetric O (the others
    from O_iik = 0), F_i = O_ijk d_j d_k; inv = 1 / |d| (softened). Built from the full tensor."""
    idx = {(0, 0, 0): 0, (1, 1, 1): 1, (2, 2, 2): 2, (0, 0, 1): 3, (0, 0, 2): 4, (0, 1, 1): 5,
           (0, 1, 2): 6}
    cols = []
    for j in range(7):
        o = np.zeros(7)
        o[j] = 1.0
        O = np.zeros((3, 3, 3))
        full = {**{k: o[v] for k, v in idx.items()},
                (0, 2, 2): -o[0] - o[5], (1, 2, 2): -o[3] - o[1], (1, 1, 2): -o[4] - o[2]}
        for i in range(3):
            for a in range(3):
                for b in range(3):
                    O[i, a, b] = full[tuple(sorted((i, a, b)))]
        F = np.einsum("iab,...a,...b->...i", O, d, d)
        G = (F * d).sum(-1)
        inv2 = inv * inv
        inv7 = inv2 ** 3 * inv
        cols.append(inv7[..., None] * (3.0 * F - 7.0 * (G * inv2)[..., None] * d))
    return np.stack(cols, -1)


def _row_lane_basis(wave, mask, geom, lanes):
    """Per (row, lane): the quadrupole basis scaled by M w^2 and the octupole basis scaled by M w^3
    (rows x 64 x 3 x 13) and the lane bits."""
    bits = lane_bits(mask)
    geom = np.asarray(geom, np.float64)
    d = geom[:, None, :3] - np.asarray(lanes, np.float64)[wave][:, :, :3]
    r2 = (d * d).sum(-1) + np.asarray(lanes, np.float64)[wave][:, :, 3]
    inv = np.where(bits, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    B = np.concatenate([field_basis(d, inv) * geom[:, 3, None, None, None],
                        oct_field_basis(d, inv) * geom[:, 4, None, None, None]], -1)
    return B * bits[:, :, None, None], bits


def predict(y, wave, mask, geom, lanes, num_waves):
    B, _ = _row_lane_basis(wave, mask, geom, lanes)
    contrib = np.einsum("rlij,rj->rli", B, y)
    pred = np.zeros((num_waves, 64, 3))
    np.add.at(pred, wave, contrib)
    return pred


def loss_and_grads(p, x, wave, mask, geom, lanes, target, wgt, bf16=False):
    num_waves = target.shape[0]
    y, (z1, h1, z2, h2) = forward(p, x, bf16)
    B, _ = _row_lane_basis(wave, mask, geom, lanes)
    pred = np.zeros((num_waves, 64, 3))
    np.add.at(pred, wave, np.einsum("rlij,rj->rli", B, y))
    count = max(int((wgt > 0).sum()), 1)
    diff = pred - target
    loss = float((wgt[..., None] * diff * diff).sum() / count)
    g_pred = 2.0 * wgt[..., None] * diff / count
    g_y = np.einsum("rlij,rli->rj", B, g_pred[wave])
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    g = {}
    g["W3"] = r(h2).T @ r(g_y)
    g["b3"] = g_y.sum(0)
    g_z2 = (r(g_y) @ r(p["W3"]).T) * (z2 > 0)
    g["W2"] = r(h1).T @ r(g_z2)
    g["b2"] = g_z2.sum(0)
    g_z1 = (r(g_z2) @ r(p["W2"]).T) * (z1 > 0)
    g["W1"] = r(x).T @ r(g_z1)
    g["b1"] = g_z1.sum(0)
    return loss, g, pred


def adam_step(p, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in FP64 (state: dict with 'm', 'v', 't')."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        m = state.setdefault("m", {}).setdefault(k, np.zeros_like(p[k]))
        v = state.setdefault("v", {}).setdefault(k, np.zeros_like(p[k]))
        m[...] = b1 * m + (1 - b1) * g[k]
        v[...] = b2 * v + (1 - b2) * g[k] * g[k]
        mh = m / (1 - b1**t)
        vh = v / (1 - b2**t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p
