"""Leapfrog step loop on the GPU (FP64 positions/velocities, brute-force forces)."""
import ctypes

import numpy as np
import pytest

import particular_amd as pa
from particular_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def init(res, s, checked=1):
    L = _lib.lib()
    pos = np.ascontiguousarray(s.pos)
    vel = np.ascontiguousarray(s.vel)
    mu = np.ascontiguousarray(s.particles[:, 3])
    res.check(L.pnb_leapfrog_init(res.handle, pos.ctypes.data, vel.ctypes.data, mu.ctypes.data,
                                  s.n, s.softening, checked))


def read(res, n):
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    a = np.zeros((n, 4), np.float32)
    res.check(_lib.lib().pnb_leapfrog_read(res.handle, x.ctypes.data, v.ctypes.data, a.ctypes.data))
    return x, v, a


def run(res, steps, dt, graph):
    res.check(_lib.lib().pnb_leapfrog_run(res.handle, steps, dt, int(graph)))
    res.synchronize()


def test_matches_numpy_kdk_with_oracle_forces(res, orc):
    s = workloads.massive_massless_case(n_massive=512, n_total=1500, softening=1e-2)
    res.reset_stream()
    init(res, s)
    dt, steps = 1e-3, 20
    run(res, steps, dt, False)
    xg, vg, ag = read(res, s.n)

    x, v = s.pos.copy(), s.vel.copy()
    mu = s.particles[:, 3]
    massive = mu != 0

    def force(x):
        p = workloads.pack(x, mu)
        return orc.brute_force(p, p[massive], softening=s.softening)[:, :3].astype(np.float64)

    a = force(x)
    for _ in range(steps):
        v += a * (0.5 * dt)
        x += v * dt
        a = force(x)
        v += a * (0.5 * dt)
    assert np.max(np.abs(xg - x)) < 1e-9
    assert np.max(np.abs(vg - v)) < 1e-6 * np.max(np.abs(v))
    np.testing.assert_allclose(ag[:, :3], a, rtol=1e-4, atol=1e-6 * np.abs(a).max())


def test_energy_conserved_plummer(res):
    s = workloads.direct_sum_case(2048, seed=31)
    s.softening = 1e-2
    res.reset_stream()
    init(res, s)

    def energy(x, v):
        m = s.particles[:, 3].astype(np.float64)
        d = x[:, None, :] - x[None, :, :]
        r = np.sqrt((d * d).sum(-1) + s.softening**2)
        np.fill_diagonal(r, np.inf)
        return 0.5 * (m * (v * v).sum(1)).sum() - 0.5 * (m[:, None] * m[None, :] / r).sum()

    e0 = energy(s.pos, s.vel)
    run(res, 200, 1e-3, True)
    x, v, _ = read(res, s.n)
    assert abs(energy(x, v) / e0 - 1) < 1e-4


def test_graph_replays_match_eager(res):
    s = workloads.massive_massless_case(n_massive=3000, n_total=9000)
    res.reset_stream()
    dt, k = 1e-3, 3
    init(res, s)
    eager = []
    for _ in range(3):
        run(res, k, dt, False)
        eager.append(read(res, s.n))
    init(res, s)
    other = workloads.direct_sum_case(700, seed=2).particles
    for i in range(3):
        run(res, k, dt, True)
        g = read(res, s.n)
        for a, b in zip(g, eager[i]):
            np.testing.assert_array_equal(a, b)
        # unrelated eager work on the same context and a host synchronisation between replays
        pa.BruteForce(res, pa.Acceleration.checked()).compute(other)
        res.synchronize()


def test_state_errors(res):
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.pnb_create(0, ctypes.byref(h)) == 0
    try:
        assert L.pnb_leapfrog_run(h, 1, 1e-3, 0) == _lib.PNB_E_STATE
        assert L.pnb_leapfrog_read(h, None, None, None) == _lib.PNB_E_STATE
        assert L.pnb_leapfrog_init(h, None, None, None, 0, 0.0, 1) == _lib.PNB_E_ARG
    finally:
        L.pnb_destroy(h)
