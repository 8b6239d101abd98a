"""Row features of the learned Barnes-Hut correction in FP64 -- TEST INFRASTRUCTURE ONLY.

No reference counterpart (particular has no learned component). Restates, in numpy, what the
library computes for one (wavefront, accepted internal node) row (particular_amd/csrc/bh.hip,
row_features and wave_frame), so the GPU's training rows can be checked against the exported
octree; and the node moments from the node's particles, so the exported quadrupoles can be
checked. The multipole claim behind features 7..12 (the quadrupole term of a node seen from c is
(M w^2 / d^4) (-f7..9 + 1.5 f10..12)) is pinned against exact sums in tests/test_features_oracle.py.
"""
from __future__ import annotations

import numpy as np


ROW_SPREAD = 0.15  # rows with wavefront rms radius / d above this are not corrected


def wave_frame(pos: np.ndarray):
    """Centroid and rms radius of a wavefront's particles (k x 3)."""
    pos = np.asarray(pos, np.float64)
    c = pos.mean(0)
    return c, float(np.sqrt(((pos - c) ** 2).sum(1).mean()))


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


def row_features(com, mass, width, q6, centre, wrad, total_mass, lanes):
    """Features f[0..12] and correction scale M w^2 / d^4 of one row (zeros and 0 when d == 0 or
    the wavefront is too spread out for the node: rms radius > ROW_SPREAD d)."""
    r = np.asarray(com, np.float64) - np.asarray(centre, np.float64)
    d2 = float(r @ r)
    f = np.zeros(13)
    if not d2 > 0 or wrad * wrad > ROW_SPREAD * ROW_SPREAD * d2:
        return f, 0.0
    d = np.sqrt(d2)
    u = r / d
    s = width / d
    Q = quad_matrix(q6)
    qu = Q @ u
    quu = float(u @ qu)
    f[0:3] = u
    f[3] = s
    f[4] = 1.0 + np.log2(max(mass / total_mass, 1e-30)) / 32.0
    f[5] = wrad / d
    f[6] = lanes / 64.0
    f[7:10] = qu - quu * u
    f[10:13] = quu * u
    return f, mass / d2 * s * s


def quadrupole_term(f, scale):
    """The quadrupole correction implied by a row's features and scale M w^2 / d^4:
    scale (-f7..9 + 1.5 f10..12)."""
    return scale * (-np.asarray(f[7:10]) + 1.5 * np.asarray(f[10:13]))
