"""Node features and the quadrupole field of the learned Barnes-Hut correction in FP64 -- TEST
INFRASTRUCTURE ONLY.

No reference counterpart (particular has no learned component). Restates, in numpy, what the
library computes per internal octree node (particular_amd/csrc/bh.hip: node_features, the
moments of bh_up_kernel) and the field every accepting lane adds (csrc/mlp_common.hpp:
quad_field), so that the GPU's training rows, node moments and features can be checked against
an exported octree and sums over its bodies. The field formula is pinned against exact sums over a cluster in
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


def third_moments(pos: np.ndarray, mass: np.ndarray) -> np.ndarray:
    """U = sum m (x - c)(x - c)(x - c) about the centre of mass c, as {xxx, yyy, zzz, xxy, xxz,
    xyy, yyz, xzz, yzz, xyz}."""
    pos = np.asarray(pos, np.float64)
    m = np.asarray(mass, np.float64)
    c = (m[:, None] * pos).sum(0) / m.sum()
    x, y, z = (pos - c).T
    return np.array([(m * a * b * e).sum() for a, b, e in
                     ((x, x, x), (y, y, y), (z, z, z), (x, x, y), (x, x, z), (x, y, y), (y, y, z),
                      (x, z, z), (y, z, z), (x, y, z))])


def octupole7(u10: np.ndarray, mass: float, width: float) -> np.ndarray:
    """Normalised octupole (5/2) U~ / (M w^3) of third moments u10 (third_moments order), U~ the
    traceless part U_ijk - (d_ij V_k + d_ik V_j + d_jk V_i) / 5 with V_k = U_llk, built as the
    full tensor; as the seven coefficients {xxx, yyy, zzz, xxy, xxz, xyy, xyz}."""
    idx = {(0, 0, 0): 0, (1, 1, 1): 1, (2, 2, 2): 2, (0, 0, 1): 3, (0, 0, 2): 4, (0, 1, 1): 5,
           (1, 1, 2): 6, (0, 2, 2): 7, (1, 2, 2): 8, (0, 1, 2): 9}
    U = np.zeros((3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                U[i, j, k] = u10[idx[tuple(sorted((i, j, k)))]]
    V = np.einsum("llk->k", U)
    d = np.eye(3)
    O = U - (np.einsum("ij,k->ijk", d, V) + np.einsum("ik,j->ijk", d, V)
             + np.einsum("jk,i->ijk", d, V)) / 5.0
    f = 2.5 / (mass * width**3)
    return f * np.array([O[0, 0, 0], O[1, 1, 1], O[2, 2, 2], O[0, 0, 1], O[0, 0, 2], O[0, 1, 1],
                         O[0, 1, 2]])


def oct_field(o7, d, eps2: float = 0.0) -> np.ndarray:
    """Field at a point of the octupole term with coefficients o7 (absolute units, i.e. times
    M w^3): the gradient of O_ijk r_i r_j r_k / |r|^7 (r = point - centre of mass), written with
    d = centre of mass - point as 3 F |d|^-7 - 7 (d.F) d |d|^-9, F_i = O_ijk d_j d_k."""
    o = np.asarray(o7, np.float64)
    O = np.zeros((3, 3, 3))
    full = {(0, 0, 0): o[0], (1, 1, 1): o[1], (2, 2, 2): o[2], (0, 0, 1): o[3], (0, 0, 2): o[4],
            (0, 1, 1): o[5], (0, 1, 2): o[6], (0, 2, 2): -o[0] - o[5], (1, 2, 2): -o[3] - o[1],
            (1, 1, 2): -o[4] - o[2]}
    for i in range(3):
        for a in range(3):
            for b in range(3):
                O[i, a, b] = full[tuple(sorted((i, a, b)))]
    d = np.asarray(d, np.float64)
    r = np.sqrt(d @ d + eps2)
    F = np.einsum("iab,a,b->i", O, d, d)
    return 3.0 * F / r**7 - 7.0 * (d @ F) * d / r**9


OCT_FEATURE_SCALE = 32.0  # the network reads 32 x the normalised octupole as f9..15


def node_features(nodes, k: int, sp: np.ndarray | None = None) -> np.ndarray:
    """Features f[0..15] of node k of an exported tree (zeros for a leaf or a massless node).

    Without `sp`, f0..8 from the exported node data (f9..15 left zero). With `sp` (the bodies in
    tree order, x y z mu), every feature from FP64 sums over the node's own bodies, the tree's
    structure (range, level) taken from the export: the mass share, the quadrupole and the
    normalised octupole (f9..15 = 32 x octupole7 of third_moments) are then computed independently
    of the library's bottom-up moments."""
    n = nodes[k]
    f = np.zeros(16)
    if not n["internal"] or not n["mass"] > 0:
        return f
    if sp is None:
        f[0:6] = n["quad"]
        f[6] = 1.0 + np.log2(max(float(n["mass"]) / float(nodes["mass"][0]), 1e-30)) / 32.0
    else:
        b = int(n["begin"])
        own = np.asarray(sp[b:b + int(n["count"])], np.float64)
        M = own[:, 3].sum()
        root_mass = np.asarray(sp, np.float64)[:, 3].sum()
        w = float(n["width"])
        f[0:6] = quadrupole(own[:, :3], own[:, 3], w)
        f[6] = 1.0 + np.log2(max(M / root_mass, 1e-30)) / 32.0
        f[9:16] = OCT_FEATURE_SCALE * octupole7(third_moments(own[:, :3], own[:, 3]), M, w)
    f[7] = np.log2(float(n["count"])) / 32.0
    f[8] = float(n["level"]) / 64.0
    return f
