"""BruteForcePairs on the GPU (pnb_bf_pairs_accel_f32*) against the oracle.

The pair kernel sums different rounding orders from the reference's sequential
RestrictedBruteForce (particular/src/sequential.rs:215-263), so the bars are the brute-force ones:
rms <= 1e-5 and max <= 1e-4 relative against the restated reference, and against the FP64 direct
sum. Sizes cover every tiling edge: one body, sizes around a tile (512) and a block of eight
tiles (4,096), and several blocks (the off-diagonal block pairs and the partial sums per block).
"""
import numpy as np
import pytest

import particular_amd as pa
from particular_amd import _lib, metrics, workloads
from tests.shared import kat_particles

pytestmark = pytest.mark.gpu


def pairs(res, storage, interaction=None):
    return pa.BruteForcePairs(res, interaction or pa.Acceleration.checked()).compute(storage)


def test_kat_reordered(res, orc):
    """The reference's acceleration_error! data through Reordered (sequential.rs:368-385)."""
    p = kat_particles()
    got = pairs(res, pa.Reordered.new(p))
    aff = p[:, 3] != 0
    order = np.concatenate([np.nonzero(aff)[0], np.nonzero(~aff)[0]])
    ref_re = orc.brute_force_pairs(p[order], restrict_len=int(aff.sum()))
    ref = np.empty_like(ref_re)
    ref[order] = ref_re
    assert metrics.max_rel(got, ref) < 1e-5
    assert metrics.max_rel(got, orc.direct_f64(p, p)) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 511, 512, 513, 4095, 4096, 4097, 8193, 20000])
def test_sizes_against_oracle(res, orc, n):
    s = workloads.direct_sum_case(n, seed=300 + n)
    p = s.particles.copy()
    p[:, 3] *= np.random.default_rng(n).uniform(0.5, 2.0, n).astype(np.float32)  # unequal masses
    got = pairs(res, p)
    if n == 1:
        np.testing.assert_array_equal(got, np.zeros((1, 3), np.float32))
        return
    ref = orc.brute_force_pairs(p) if n <= 4097 else orc.brute_force(p, p)
    e = metrics.summary(got, ref)
    e64 = metrics.summary(got, orc.direct_f64(p, p))
    print(f"n={n}: vs oracle {e}, vs FP64 {e64}", flush=True)
    assert e["rms_rel"] < 1e-5 and e["max_rel"] < 1e-4
    assert e64["rms_rel"] < 1e-5 and e64["max_rel"] < 1e-4


def test_equals_one_sided_brute_force(res):
    s = workloads.direct_sum_case(10000, seed=41)
    got = pairs(res, s.particles)
    one = pa.BruteForce(res, pa.Acceleration.checked()).compute(s.particles)
    assert metrics.rms_rel(got, one) < 1e-6


def test_softened_and_unchecked(res, orc):
    s = workloads.direct_sum_case(3000, seed=42)
    p = s.particles
    for eps in (1e-3, 0.1):
        got = pairs(res, p, pa.AccelerationSoftened.checked(eps))
        ref = orc.brute_force(p, p, softening=eps)
        assert metrics.rms_rel(got, ref) < 1e-5, eps
    # distinct positions: the unchecked form gives the same sums
    np.testing.assert_array_equal(pairs(res, p, pa.Acceleration.unchecked()), pairs(res, p))


def test_coincident_particles(res, orc):
    """checked drops pairs at zero distance (as the one-sided kernel); unchecked keeps the
    reference's pair form, whose coincident pairs are non-finite."""
    s = workloads.direct_sum_case(700, seed=43)
    p = np.concatenate([s.particles, s.particles[:5]])
    got = pairs(res, p)
    assert np.all(np.isfinite(got))
    assert metrics.rms_rel(got, orc.brute_force(p, p, checked=True)) < 1e-5
    assert not np.all(np.isfinite(pairs(res, p, pa.Acceleration.unchecked())))


def test_storages_and_device(res, orc):
    """Slice, Ordered, Reordered and a device tensor give one answer (pairs of two massless
    particles add zero)."""
    import torch

    s = workloads.direct_sum_case(5000, seed=44)
    p = s.particles.copy()
    p[::3, 3] = 0.0
    a = pairs(res, p)
    b = pairs(res, pa.Ordered.new(p))
    perm = np.concatenate([np.nonzero(p[:, 3] != 0)[0], np.nonzero(p[:, 3] == 0)[0]])
    c = pairs(res, pa.Reordered.new(p))
    d = pairs(res, torch.from_numpy(p).cuda()).cpu().numpy()[:, :3]
    assert metrics.rms_rel(b, a[perm]) < 1e-6
    assert metrics.rms_rel(c, a) < 1e-6
    np.testing.assert_array_equal(d, a)
    ref = orc.brute_force(p, p[p[:, 3] != 0])
    assert metrics.rms_rel(a, ref) < 1e-5
    with pytest.raises(TypeError):
        pairs(res, pa.Between(p, p))


def test_deterministic(res):
    s = workloads.direct_sum_case(9000, seed=45)
    np.testing.assert_array_equal(pairs(res, s.particles), pairs(res, s.particles))


def test_argument_errors(res):
    L = _lib.lib()
    assert L.pnb_bf_pairs_accel_f32(res.handle, None, 5, 0.0, 1, None) == _lib.PNB_E_ARG
    p = np.zeros((4, 4), np.float32)
    out = np.zeros((4, 4), np.float32)
    assert L.pnb_bf_pairs_accel_f32(res.handle, p.ctypes.data, 4, -1.0, 1,
                                    out.ctypes.data) == _lib.PNB_E_ARG
    assert L.pnb_bf_pairs_accel_f32(res.handle, p.ctypes.data, (1 << 22) + 1, 0.0, 1,
                                    out.ctypes.data) == _lib.PNB_E_ARG
    assert L.pnb_bf_pairs_accel_f32(res.handle, None, 0, 0.0, 1, None) == _lib.PNB_OK
