"""Relative acceleration error metrics.

rms_rel = sqrt(mean_i |a_i - r_i|^2 / |r_i|^2) and max_rel = max_i |a_i - r_i| / |r_i|, both over
particles with |r_i| > 0. (The reference's own tests use a componentwise ratio,
particular/src/gravity/newtonian/mod.rs:268-271; these per-particle vector norms are stricter and
do not blow up on near-zero components.)
"""
from __future__ import annotations

import numpy as np


def rel_errors(a, ref) -> np.ndarray:
    a = np.asarray(a, np.float64)[:, :3]
    ref = np.asarray(ref, np.float64)[:, :3]
    num = np.linalg.norm(a - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    ok = den > 0
    return num[ok] / den[ok]


def rms_rel(a, ref) -> float:
    e = rel_errors(a, ref)
    return float(np.sqrt(np.mean(e * e))) if e.size else 0.0


def max_rel(a, ref) -> float:
    e = rel_errors(a, ref)
    return float(e.max()) if e.size else 0.0


def summary(a, ref) -> dict:
    return {"rms_rel": rms_rel(a, ref), "max_rel": max_rel(a, ref)}
