"""Host logic of bench.py (no GPU): the equal-error interpolation behind every reported speed-up,
and the refusal of more than one GPU."""
import subprocess
import sys
import os

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rows(pairs):
    return [{"rms_rel": e, "ms_per_eval": t} for e, t in pairs]


def test_time_at_error_is_exact_on_power_laws():
    """Time a power of the error: log-log interpolation returns it exactly between samples."""
    c, k = 3.0, -0.7
    errs = [3e-5, 1e-4, 4e-4, 2e-3]
    curve = rows([(e, c * e ** k) for e in errs][::-1])  # order of the rows does not matter
    for target in (5e-5, 1e-4, 7.3e-4):
        assert np.isclose(bench.time_at_error(curve, target), c * target ** k, rtol=1e-12)


def test_time_at_error_outside_the_curve_is_none():
    curve = rows([(1e-4, 10.0), (1e-3, 2.0)])
    assert bench.time_at_error(curve, 5e-5) is None
    assert bench.time_at_error(curve, 2e-3) is None
    assert np.isclose(bench.time_at_error(curve, 1e-4), 10.0, rtol=1e-12)  # an end point


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bh_roofline_picks_the_equal_error_row_and_the_profile():
    """The walk's roofline entry: the plain geometric row closest to 1e-4 rms, its accepted-term
    rate over the brute-force kernel's pair rate, and the counters of the profiled angle."""
    def row(theta, rms, crit="geometric", mlp=False):
        return {"theta": theta, "rms_rel": rms, "criterion": crit, "learned_correction": mlp,
                "accepted_G_terms_per_s": 1000.0 * theta, "walk_ms": 50.0,
                "wave_steps_per_wave": 1.0, "accepted_terms_per_particle": 18884.87}
    rows_ = [row(0.5, 8.5e-4), row(0.2291, 0.98e-4), row(0.3, 1.0e-4, mlp=True),
             row(1.0, 1.0e-4, crit="relative")]
    bf = {"config": {"interactions_per_step": 6.8719476736e10}, "roofline": {"avg_kernel_ms": 17.0}}
    r = bench.bh_roofline(rows_, bf)
    assert r["theta"] == 0.2291
    pair_rate = 6.8719476736e10 / 17.0e-3 / 1e9
    assert np.isclose(r["peak"], pair_rate) and np.isclose(r["frac"], 229.1 / pair_rate)
    prof = bench.load_profile("bh_walk_two_plummer_4M")
    assert prof is not None and r["counters_theta"] == 0.2291
    assert np.isclose(r["valu_lane_ops_per_accepted_term"], 6.109e10 * 64 / (18884.87 * 4194304))
    assert 0 < r["l2_hit_rate"] < 1 and r["traffic"] > 1e10
