"""Brute force on the GPU through the C ABI, against the pinned CPU oracle.

Bars (relative acceleration error, per-particle vector norm): rms <= 1e-5 against the f32
restatement of sequential::BruteForce and against FP64; max <= 1e-4 at these sizes.
"""
import numpy as np
import pytest

import particular_amd as pa
from particular_amd import metrics, workloads
from tests.shared import circular_orbit, kat_particles

pytestmark = pytest.mark.gpu

RMS_BAR = 1e-5
MAX_BAR = 1e-4


def bf(res, storage, interaction=None):
    return pa.BruteForce(res, interaction or pa.Acceleration.checked()).compute(storage)


def test_kat_reordered(res, orc):
    p = kat_particles()
    got = bf(res, pa.Reordered.new(p))
    exp = orc.direct_f64(p, p[p[:, 3] != 0])
    assert got.shape == (6, 3)
    assert metrics.max_rel(got, exp) < 1e-6


def test_direct_sum_4096_plummer(res, orc):
    s = workloads.direct_sum_case(4096)
    got = bf(res, s.particles)
    ref32 = orc.brute_force(s.particles, s.particles)
    ref64 = orc.direct_f64(s.particles, s.particles)
    print("vs f32 oracle", metrics.summary(got, ref32), "vs f64", metrics.summary(got, ref64),
          "oracle vs f64", metrics.summary(ref32, ref64))
    assert metrics.rms_rel(got, ref32) < RMS_BAR and metrics.max_rel(got, ref32) < MAX_BAR
    assert metrics.rms_rel(got, ref64) < RMS_BAR and metrics.max_rel(got, ref64) < MAX_BAR


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("per_lane", [1, 2, 4, 8])
@pytest.mark.parametrize("j_split", [1, 3, 16])
def test_every_kernel_shape(res, orc, variant, per_lane, j_split):
    s = workloads.direct_sum_case(3000, seed=21)
    b = s.particles[:2500]
    ref = orc.direct_f64(s.particles, b)
    res.set_bf_tuning(0, variant, per_lane, j_split)
    try:
        got = bf(res, pa.Between(s.particles, b))
    finally:
        res.clear_bf_tuning()
    assert metrics.rms_rel(got, ref) < RMS_BAR and metrics.max_rel(got, ref) < MAX_BAR


@pytest.mark.parametrize("n,m", [(1, 1), (3, 5), (257, 255), (1000, 1023), (2049, 1), (5, 4097)])
def test_ragged_sizes(res, orc, n, m):
    rng = np.random.default_rng(n + 7 * m)
    a = rng.normal(size=(n, 4)).astype(np.float32)
    b = rng.normal(size=(m, 4)).astype(np.float32)
    b[:, 3] = rng.uniform(0.5, 1.5, m)
    got = bf(res, pa.Between(a, b))
    assert metrics.max_rel(got, orc.direct_f64(a, b)) < MAX_BAR


def test_empty_inputs(res):
    a = np.ones((4, 4), np.float32)
    empty = np.zeros((0, 4), np.float32)
    assert bf(res, pa.Between(empty, a)).shape == (0, 3)
    np.testing.assert_array_equal(bf(res, pa.Between(a, empty)), np.zeros((4, 3), np.float32))
    massless = a.copy()
    massless[:, 3] = 0
    np.testing.assert_array_equal(bf(res, pa.Reordered.new(massless)), np.zeros((4, 3)))


def test_coincident_particles_checked(res, orc):
    s = workloads.direct_sum_case(600, seed=4)
    p = np.concatenate([s.particles, s.particles[:50]])  # exact duplicates
    got = bf(res, p)
    assert np.all(np.isfinite(got))
    ref = orc.brute_force(p, p)
    assert metrics.rms_rel(got, ref) < RMS_BAR and metrics.max_rel(got, ref) < MAX_BAR


def test_unchecked_equals_checked_for_distinct_positions(res):
    s = workloads.direct_sum_case(1000, seed=6)
    a = s.particles
    b = s.particles[::-1].copy() + np.float32(1e-3)
    b[:, 3] = s.particles[:, 3]
    np.testing.assert_array_equal(bf(res, pa.Between(a, b), pa.Acceleration.unchecked()),
                                  bf(res, pa.Between(a, b), pa.Acceleration.checked()))


@pytest.mark.parametrize("eps", [1e-3, 0.1])
def test_softened(res, orc, eps):
    s = workloads.direct_sum_case(2048, seed=8)
    got = bf(res, s.particles, pa.AccelerationSoftened.checked(eps))
    ref = orc.brute_force(s.particles, s.particles, softening=eps)
    assert metrics.rms_rel(got, ref) < RMS_BAR and metrics.max_rel(got, ref) < MAX_BAR
    ref64 = orc.direct_f64(s.particles, s.particles, softening=eps)
    assert metrics.rms_rel(got, ref64) < RMS_BAR


def test_reordered_host_and_device_agree(res, orc):
    import torch

    s = workloads.massive_massless_case(n_massive=4096, n_total=20000)
    r = pa.Reordered.new(s.particles)
    host = bf(res, r, pa.AccelerationSoftened.checked(s.softening))
    between = bf(res, pa.Between(s.particles, r.affecting()), pa.AccelerationSoftened.checked(s.softening))
    np.testing.assert_array_equal(host, between)
    dev = bf(res, pa.Reordered.new(torch.from_numpy(s.particles).cuda()),
             pa.AccelerationSoftened.checked(s.softening))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy()[:, :3], host)
    ref = orc.brute_force(s.particles, r.affecting(), softening=s.softening)
    assert metrics.rms_rel(host, ref) < RMS_BAR


def test_massive_massless_at_scale_sampled(res, orc):
    """1,048,576 particles (65,536 massive, eps = 1e-3): the full evaluation on the GPU, checked
    on a seeded sample of 2,048 affected particles against the oracle and FP64."""
    import torch

    s = workloads.massive_massless_case()
    dev = torch.from_numpy(s.particles).cuda()
    out = bf(res, pa.Reordered.new(dev), pa.AccelerationSoftened.checked(s.softening))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    idx = np.random.default_rng(0).choice(s.n, 2048, replace=False)
    massive = s.particles[s.particles[:, 3] != 0]
    ref32 = orc.brute_force(s.particles[idx], massive, softening=s.softening)
    ref64 = orc.direct_f64(s.particles[idx], massive, softening=s.softening)
    print("vs f32 oracle", metrics.summary(got[idx], ref32), "vs f64", metrics.summary(got[idx], ref64),
          "oracle vs f64", metrics.summary(ref32, ref64))
    assert metrics.rms_rel(got[idx], ref64) < RMS_BAR
    assert metrics.rms_rel(got[idx], ref32) < 2 * metrics.rms_rel(ref32, ref64) + RMS_BAR


def test_direct_f64_reference_matches_oracle(res, orc):
    import torch

    s = workloads.direct_sum_case(2000, seed=12)
    d = torch.from_numpy(s.particles).cuda()
    out = torch.empty((s.n, 3), dtype=torch.float64, device="cuda")
    res.use_torch_stream()
    res.check(pa._lib.lib().pnb_direct_f64_dev(res.handle, d.data_ptr(), s.n, d.data_ptr(), s.n,
                                               0.0, out.data_ptr()))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), orc.direct_f64(s.particles, s.particles),
                               rtol=1e-11, atol=1e-300)


def test_circular_orbit_gpu(res):
    """circular_orbit! on the GPU path, 100 orbits as the reference's GPU test
    (particular/src/gravity/newtonian/mod.rs:511-518)."""
    alg = pa.BruteForce(res, pa.Acceleration.checked())

    def accel(p):
        out = np.zeros((2, 4), np.float32)
        alg._reordered_host(np.ascontiguousarray(p), out)
        return out

    ed, ee = circular_orbit(accel, 100)
    assert ed < 1e-2 and ee < 1e-2


def test_f64_brute_force(res):
    """float64 particles take the FP64 kernel (the reference's DVec3 path)."""
    rng = np.random.default_rng(77)
    a = rng.normal(size=(700, 4))
    b = rng.normal(size=(900, 4))
    b[:, 3] = rng.uniform(0.1, 2.0, 900)
    b[:3, :3] = a[:3, :3]  # coincident pairs: checked drops them
    got = bf(res, pa.Between(a, b), pa.AccelerationSoftened.checked(0.0))
    d = b[None, :, :3] - a[:, None, :3]
    r2 = (d * d).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(r2 > 0, b[None, :, 3] / (r2 * np.sqrt(r2)), 0.0)
    ref = (d * s[..., None]).sum(1)
    assert got.dtype == np.float64
    assert metrics.max_rel(got, ref) < 1e-12
    soft = bf(res, pa.Reordered.new(b), pa.AccelerationSoftened.checked(0.05))
    d = b[None, :, :3] - b[:, None, :3]
    rs = (d * d).sum(-1) + 0.05**2
    ref = (d * (b[None, :, 3] / (rs * np.sqrt(rs)))[..., None]).sum(1)
    assert metrics.max_rel(soft, ref) < 1e-12
