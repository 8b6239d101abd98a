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


def test_octupole_field_is_the_quadrupole_residual():
    """For a cluster seen from a distant point, the exact field minus its monopole and quadrupole
    terms is the octupole field of its normalised octupole (octupole7 of third_moments, times
    M w^3), up to O((w/d)^4) relative to the monopole: pins third_moments, octupole7 (its trace
    removal and the 5/2) and oct_field together; at d = 128 w the residual it explains is well
    above the tolerance."""
    rng = np.random.default_rng(7)
    explained = 0
    for trial in range(20):
        k = int(rng.integers(3, 40))
        w = 1.0
        pos = rng.uniform(-0.5, 0.5, (k, 3)) * w + rng.normal(size=3)
        mass = rng.uniform(0.1, 1.0, k)
        Mt = mass.sum()
        com = (mass[:, None] * pos).sum(0) / Mt
        q_abs = Mt * w * w * F.quadrupole(pos, mass, w)
        o_abs = Mt * w**3 * F.octupole7(F.third_moments(pos, mass), Mt, w)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for d in (16.0, 128.0):
            x = com - d * u
            mono = Mt * (com - x) / d**3
            resid = exact_field(x, pos, mass) - mono - F.quad_field(q_abs, com - x)
            octf = F.oct_field(o_abs, com - x)
            # the hexadecapole order: (residual - octupole) d^4 / |monopole| stays O(0.1) as d
            # grows, while the residual itself falls only as d^-3
            bound = 0.3 * (w / d) ** 4 * np.linalg.norm(mono)
            assert np.linalg.norm(resid - octf) < bound, (trial, d)
            if d == 128.0 and np.linalg.norm(resid) > 5 * bound:
                explained += 1
    assert explained >= 15


def test_third_moments_compose_by_the_parallel_axis_rule():
    """U of a union from its parts' U, second moments T and centres: U = sum over parts of
    U_p + u_i T_jk + u_j T_ik + u_k T_ij + M_p u_i u_j u_k with u = c_p - c (the composition the
    library's bottom-up pass uses)."""
    rng = np.random.default_rng(12)
    pos = rng.normal(size=(60, 3))
    mass = rng.uniform(0.1, 1.0, 60)
    full = F.third_moments(pos, mass)
    c = (mass[:, None] * pos).sum(0) / mass.sum()
    acc = np.zeros(10)
    trip = ((0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 0, 1), (0, 0, 2), (0, 1, 1), (1, 1, 2),
            (0, 2, 2), (1, 2, 2), (0, 1, 2))
    for part in (slice(0, 17), slice(17, 41), slice(41, 60)):
        p, m = pos[part], mass[part]
        cp = (m[:, None] * p).sum(0) / m.sum()
        d = p - cp
        T = (m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
        u = cp - c
        Up = F.third_moments(p, m)
        for q, (i, j, k) in enumerate(trip):
            acc[q] += (Up[q] + u[i] * T[j, k] + u[j] * T[i, k] + u[k] * T[i, j]
                       + m.sum() * u[i] * u[j] * u[k])
    np.testing.assert_allclose(acc, full, rtol=1e-10, atol=1e-10)


def test_octupole7_removes_traces():
    """A tensor made only of traces, U_ijk = d_ij a_k + d_ik a_j + d_jk a_i, has no octupole; the
    result is linear in U and scales as 1 / (M w^3)."""
    rng = np.random.default_rng(13)
    a = rng.normal(size=3)
    iso = np.array([3 * a[0], 3 * a[1], 3 * a[2], a[1], a[2], a[0], a[2], a[0], a[1], 0.0])
    np.testing.assert_allclose(F.octupole7(iso, 1.0, 1.0), 0.0, atol=1e-12)
    u = F.third_moments(rng.normal(size=(30, 3)), rng.uniform(0.1, 1.0, 30))
    np.testing.assert_allclose(F.octupole7(u + iso, 2.0, 0.5), F.octupole7(u, 1.0, 1.0) * 4.0,
                               rtol=1e-12)


def test_mlp_oracle_octupole_basis_is_the_octupole_field():
    rng = np.random.default_rng(10)
    for _ in range(20):
        o = rng.normal(size=7)
        d = rng.normal(size=3) * 3
        eps2 = float(rng.uniform(0, 0.1))
        inv = 1.0 / np.sqrt(d @ d + eps2)
        B = M.oct_field_basis(d[None], np.array([inv]))[0]
        np.testing.assert_allclose(B @ o, F.oct_field(o, d, eps2), rtol=1e-12, atol=1e-14)
