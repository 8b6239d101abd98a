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
