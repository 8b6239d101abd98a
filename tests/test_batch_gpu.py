"""Batched ensemble on the GPU: per-system forces and trajectories."""
import ctypes

import numpy as np
import pytest

import particular_amd as pa
from particular_amd import _lib, metrics, workloads

pytestmark = pytest.mark.gpu


def ensemble(num=24, n_min=20, n_max=3000, seed=5):
    return workloads.ensemble_case(num_systems=num, n_min=n_min, n_max=n_max, seed=seed)


def init(res, e, checked=1):
    L = _lib.lib()
    pos = np.ascontiguousarray(e.pos)
    vel = np.ascontiguousarray(e.vel)
    mu = np.ascontiguousarray(e.particles[:, 3])
    off = np.ascontiguousarray(e.offsets, dtype=np.int64)
    res.check(L.pnb_batch_init(res.handle, pos.ctypes.data, vel.ctypes.data, mu.ctypes.data,
                               off.ctypes.data, e.num_systems, e.softening, checked))


def read(res, n):
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    a = np.zeros((n, 4), np.float32)
    res.check(_lib.lib().pnb_batch_read(res.handle, x.ctypes.data, v.ctypes.data, a.ctypes.data))
    return x, v, a


def run(res, steps, dt, graph):
    res.check(_lib.lib().pnb_batch_run(res.handle, steps, dt, int(graph)))
    res.synchronize()


@pytest.mark.parametrize("per_lane", [1, 2, 4, 8])
def test_initial_forces_match_oracle(res, orc, per_lane):
    e = ensemble()
    res.reset_stream()
    res.check(_lib.lib().pnb_batch_set_tuning(res.handle, 0, per_lane))
    init(res, e)
    _, _, a = read(res, e.particles.shape[0])
    for k in range(e.num_systems):
        lo, hi = e.offsets[k], e.offsets[k + 1]
        p = e.particles[lo:hi]
        ref = orc.brute_force(p, p, softening=e.softening)
        assert metrics.rms_rel(a[lo:hi], ref) < 1e-5 and metrics.max_rel(a[lo:hi], ref) < 1e-4


def test_trajectories_equal_single_system_loop(res):
    """Same arithmetic and summation order as the brute-force kernel: bitwise equal states."""
    e = ensemble(num=6, n_min=100, n_max=2000, seed=9)
    res.reset_stream()
    res.check(_lib.lib().pnb_batch_set_tuning(res.handle, 0, 2))
    init(res, e)
    run(res, 5, 1e-3, False)
    xb, vb, ab = read(res, e.particles.shape[0])
    L = _lib.lib()
    res.set_bf_tuning(0, 1, 2, 1)  # scalar-load variant, 2 per lane, no affecting split
    try:
        for k in range(e.num_systems):
            lo, hi = e.offsets[k], e.offsets[k + 1]
            pos = np.ascontiguousarray(e.pos[lo:hi])
            vel = np.ascontiguousarray(e.vel[lo:hi])
            mu = np.ascontiguousarray(e.particles[lo:hi, 3])
            res.check(L.pnb_leapfrog_init(res.handle, pos.ctypes.data, vel.ctypes.data,
                                          mu.ctypes.data, hi - lo, e.softening, 1))
            res.check(L.pnb_leapfrog_run(res.handle, 5, 1e-3, 0))
            x = np.zeros((hi - lo, 3))
            a = np.zeros((hi - lo, 4), np.float32)
            res.check(L.pnb_leapfrog_read(res.handle, x.ctypes.data, None, a.ctypes.data))
            np.testing.assert_array_equal(x, xb[lo:hi])
            np.testing.assert_array_equal(a, ab[lo:hi])
    finally:
        res.clear_bf_tuning()


def test_graph_replays_match_eager(res):
    e = ensemble(num=40, n_min=16, n_max=1500, seed=3)
    res.reset_stream()
    init(res, e)
    eager = []
    for _ in range(3):
        run(res, 2, 1e-3, False)
        eager.append(read(res, e.particles.shape[0]))
    init(res, e)
    other = workloads.direct_sum_case(500, seed=4).particles
    for i in range(3):
        run(res, 2, 1e-3, True)
        g = read(res, e.particles.shape[0])
        for a, b in zip(g, eager[i]):
            np.testing.assert_array_equal(a, b)
        pa.BruteForce(res, pa.Acceleration.checked()).compute(other)
        res.synchronize()


def test_info_and_errors(res):
    e = ensemble(num=5, n_min=10, n_max=100, seed=1)
    res.reset_stream()
    init(res, e)
    n = ctypes.c_size_t()
    items = ctypes.c_size_t()
    inter = ctypes.c_ulonglong()
    res.check(_lib.lib().pnb_batch_info(res.handle, ctypes.byref(n), ctypes.byref(items),
                                        ctypes.byref(inter)))
    sizes = np.diff(e.offsets)
    assert n.value == sizes.sum() and inter.value == int((sizes.astype(np.int64) ** 2).sum())
    L = _lib.lib()
    bad = np.array([0, 5, 3], np.int64)
    z = np.zeros(15)
    assert L.pnb_batch_init(res.handle, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                            bad.ctypes.data, 2, 0.0, 1) == _lib.PNB_E_ARG
    assert L.pnb_batch_set_tuning(res.handle, 0, 3) == _lib.PNB_E_ARG
