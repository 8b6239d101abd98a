"""Pins the feature/field oracle (oracle/features.py) with the mathematics of the multipole
expansion: for a cluster seen from a distant point, the exact field minus the monopole term is
the quadrupole field of the cluster's second moment, up to O((w/d)^3) relative to the monopole;
and the field basis of the MLP oracle agrees with that field."""
import numpy as np

from oracle import features as F
from oracle import mlp_oracle as M


def exact_field(x, pos, mass):
    d = pos - x
    r = np.sqrt((d * d).sum(1))
    return (mass[:, None] * d / r[:, None] ** 3).sum(0)


def test_quadrupole_field_is_the_monopole_residual():
    rng = np.random.default_rng(7)
    for trial in range(20):
        k = int(rng.integers(2, 40))
        w = 1.0
        pos = rng.uniform(-0.5, 0.5, (k, 3)) * w + rng.normal(size=3)
        mass = rng.uniform(0.1, 1.0, k)
        Mt = mass.sum()
        com = (mass[:, None] * pos).sum(0) / Mt
        q_abs = Mt * w * w * F.quadrupole(pos, mass, w)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for d in (8.0, 16.0, 32.0):
            x = com - d * u
            mono = Mt * (com - x) / d**3
            resid = exact_field(x, pos, mass) - mono
            quad = F.quad_field(q_abs, com - x)
            bound = 4.0 * (w / d) ** 3 * np.linalg.norm(mono)  # octupole order
            assert np.linalg.norm(resid - quad) < bound, (trial, d)
            if d == 8.0 and np.linalg.norm(resid) > 10 * bound:
                assert np.linalg.norm(resid - quad) < 0.3 * np.linalg.norm(resid)


def test_mlp_oracle_basis_is_the_quadrupole_field():
    rng = np.random.default_rng(9)
    for _ in range(20):
        q = rng.normal(size=6)
        d = rng.normal(size=3) * 3
        eps2 = float(rng.uniform(0, 0.1))
        inv = 1.0 / np.sqrt(d @ d + eps2)
        B = M.field_basis(d[None], np.array([inv]))[0]
        np.testing.assert_allclose(B @ q, F.quad_field(q, d, eps2), rtol=1e-12, atol=1e-14)


def test_quadrupole_is_traceless_and_parallel_axis_consistent():
    rng = np.random.default_rng(8)
    pos = rng.normal(size=(50, 3))
    mass = rng.uniform(0.1, 1.0, 50)
    q = F.quadrupole(pos, mass, 2.0)
    assert abs(q[0] + q[1] + q[2]) < 1e-12
    # parallel-axis composition: moments of two halves combine into the moments of the whole
    a, b = slice(0, 20), slice(20, 50)

    def T(p, m):
        c = (m[:, None] * p).sum(0) / m.sum()
        d = p - c
        return (m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0), c

    Ta, ca = T(pos[a], mass[a])
    Tb, cb = T(pos[b], mass[b])
    Tw, cw = T(pos, mass)
    Ma, Mb = mass[a].sum(), mass[b].sum()
    comb = Ta + Ma * np.outer(ca - cw, ca - cw) + Tb + Mb * np.outer(cb - cw, cb - cw)
    np.testing.assert_allclose(comb, Tw, rtol=1e-12, atol=1e-12)
