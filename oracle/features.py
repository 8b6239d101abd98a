"""Node features and the quadrupole field of the learned Barnes-Hut correction in FP64 -- TEST
INFRASTRUCTURE ONLY.

No reference counterpart (particular has no learned component). Restates, in numpy, what the
library computes per internal octree node (particular_amd/csrc/bh.hip: node_features, the
moments of bh_up_kernel) and the field every accepting lane adds (csrc/mlp_common.hpp:
quad_field), so that the GPU's training rows and node moments can be checked against an exported
octree. The field formula is pinned against exact sums over a cluster in
tests/test_features_oracle.py: with the node's own quadrupole it is the monopole's residual up to
the next (octupole) order.
"""
from __future__ import annotations

import numpy as np


def quadrupole(pos: np.ndarray, mass: np.ndarray, width: float):
    """Normalised traceless quadrupole (3 T - tr(T) I) / (M w^2) about the centre of mass, as
    {xx, yy, zz, xy, xz, yz}; with T = sum m (x - c)(x - c)^T."""
    pos = np.asarray(pos, np.float64)
    m = np.asarray(mass, np.float64)
    M = m.sum()
    c = (m[:, None] * pos).sum(0) / M
    d = pos - c
    T = (m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
    Q = (3.0 * T - np.trace(T) * np.eye(3)) / (M * width * width)
    return np.array([Q[0, 0], Q[1, 1], Q[2, 2], Q[0, 1], Q[0, 2], Q[1, 2]])


def quad_matrix(q6) -> np.ndarray:
    xx, yy, zz, xy, xz, yz = q6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]], np.float64)


def quad_field(q6_abs, d, eps2: float = 0.0) -> np.ndarray:
    """Field at a point of a second moment Q (absolute units), d = centre of mass - point:
    -Q d |d|^-5 + 2.5 (d.Q.d) d |d|^-7 with |d| softened as sqrt(|d|^2 + eps2)."""
    Q = quad_matrix(q6_abs)
    d = np.asarray(d, np.float64)
    r = np.sqrt(d @ d + eps2)
    qd = Q @ d
    return -qd / r**5 + 2.5 * (d @ qd) * d / r**7


def node_children(nodes, k: int):
    """Indices of the children of node k in a pre-order node array with skip pointers."""
    out, c = [], k + 1
    while c < nodes["skip"][k]:
        out.append(c)
        c = max(int(nodes["skip"][c]), c + 1)
    return out


def node_features(nodes, k: int) -> np.ndarray:
    """Features f[0..10] of node k of an exported tree (zeros for a leaf or a massless node)."""
    n = nodes[k]
    f = np.zeros(11)
    if not n["internal"] or not n["mass"] > 0:
        return f
    f[0:6] = n["quad"]
    f[6] = 1.0 + np.log2(max(float(n["mass"]) / float(nodes["mass"][0]), 1e-30)) / 32.0
    f[7] = np.log2(float(n["count"])) / 32.0
    f[8] = float(n["level"]) / 64.0
    ch = node_children(nodes, k)
    f[9] = max(float(nodes["mass64"][c]) for c in ch) / float(n["mass64"])
    f[10] = len(ch) / 8.0
    return f
