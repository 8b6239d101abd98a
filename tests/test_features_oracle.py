"""Pins the row-feature oracle (oracle/features.py) with the mathematics of the multipole
expansion: for a cluster seen from a distant point, the exact field minus the monopole term is
the quadrupole term that features 7..12 encode, up to O((w/d)^3) relative to the monopole."""
import numpy as np

from oracle import features as F


def exact_field(x, pos, mass):
    d = pos - x
    r = np.sqrt((d * d).sum(1))
    return (mass[:, None] * d / r[:, None] ** 3).sum(0)


def test_quadrupole_term_is_the_monopole_residual():
    rng = np.random.default_rng(7)
    for trial in range(20):
        k = int(rng.integers(2, 40))
        w = 1.0
        pos = rng.uniform(-0.5, 0.5, (k, 3)) * w + rng.normal(size=3)
        mass = rng.uniform(0.1, 1.0, k)
        M = mass.sum()
        com = (mass[:, None] * pos).sum(0) / M
        q6 = F.quadrupole(pos, mass, w)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for d in (8.0, 16.0, 32.0):
            x = com - d * u  # field point: the node is at +u from it
            f, scale = F.row_features(com, M, w, q6, x, 0.0, M, 64)
            mono = M * (com - x) / d**3
            resid = exact_field(x, pos, mass) - mono
            quad = F.quadrupole_term(f, scale)
            # next term (octupole) is O((w/d)^3) of the monopole
            bound = 4.0 * (w / d) ** 3 * np.linalg.norm(mono)
            assert np.linalg.norm(resid - quad) < bound, (trial, d)
            # and the quadrupole term itself is resolved (not trivially small) at d = 8
            if d == 8.0 and np.linalg.norm(resid) > 10 * bound:
                assert np.linalg.norm(resid - quad) < 0.3 * np.linalg.norm(resid)


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


def test_features_of_a_point_at_the_centre_are_zero():
    f, s = F.row_features([1.0, 2.0, 3.0], 1.0, 1.0, np.zeros(6), [1.0, 2.0, 3.0], 0.1, 1.0, 3)
    assert s == 0.0 and not f.any()
