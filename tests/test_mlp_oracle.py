"""Pins the MLP oracle (oracle/mlp_oracle.py): analytic gradients against central finite
differences of the loss, the bf16 rounding against exact bf16 values, and the lane aggregation
against an explicit loop."""
import numpy as np

from oracle import mlp_oracle as M


def problem(seed=0, waves=3, rows_per_wave=9):
    """Rows of nodes a few units from lanes clustered near the origin."""
    rng = np.random.default_rng(seed)
    rows = waves * rows_per_wave
    x = rng.normal(size=(rows, M.IN))
    wave = np.repeat(np.arange(waves), rows_per_wave)
    mask = rng.integers(0, 2**63, rows, dtype=np.int64).astype(np.uint64) | np.uint64(1)
    geom = np.zeros((rows, 8))
    u = rng.normal(size=(rows, 3))
    geom[:, :3] = 3.0 * u / np.linalg.norm(u, axis=1, keepdims=True)
    geom[:, 3] = rng.uniform(0.5, 2.0, rows)  # M w^2
    geom[:, 4] = geom[:, 3] * rng.uniform(0.5, 1.5, rows)  # M w^3
    lanes = np.zeros((waves, 64, 4))
    lanes[..., :3] = rng.normal(size=(waves, 64, 3)) * 0.3
    lanes[..., 3] = 1e-4
    target = rng.normal(size=(waves, 64, 3)) * 0.01
    wgt = rng.uniform(0.0, 1.0, (waves, 64))
    wgt[0, :5] = 0.0
    return x, wave, mask, geom, lanes, target, wgt


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2**-7, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.0e38, 1e-30])
    got = M.bf16_round(vals)
    # 8 significant bits: 1 + 2^-7 is representable; 1 + 2^-8 ties to even (1.0);
    # 1 + 3*2^-8 ties between 1 + 2^-7 (odd) and 1 + 2^-6 (even) and goes up
    np.testing.assert_array_equal(got[:5], [1.0, 1.0 + 2**-7, 1.0, 1.0 + 2**-6, -2.5])
    assert abs(got[6] / 1e-30 - 1) < 2**-8


def test_aggregation_matches_loop():
    from oracle import features as F

    x, wave, mask, geom, lanes, target, wgt = problem(1)
    p = M.init_params(1)
    p["W3"] = np.random.default_rng(5).normal(size=p["W3"].shape)
    y, _ = M.forward(p, x)
    pred = M.predict(y, wave, mask, geom, lanes, target.shape[0])
    ref = np.zeros_like(pred)
    for r in range(x.shape[0]):
        for lane in range(64):
            if (int(mask[r]) >> lane) & 1:
                lp = lanes[wave[r], lane]
                d = geom[r, :3] - lp[:3]
                ref[wave[r], lane] += (F.quad_field(geom[r, 3] * y[r, :6], d, lp[3])
                                       + F.oct_field(geom[r, 4] * y[r, 6:], d, lp[3]))
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(pred, ref, rtol=1e-12, atol=1e-13)


def test_gradients_match_finite_differences():
    x, wave, mask, geom, lanes, target, wgt = problem(2)
    p = M.init_params(2)
    p["b1"] += 0.05  # keep pre-activations away from the ReLU kink for the probes
    p["b2"] += 0.05
    _, g, _ = M.loss_and_grads(p, x, wave, mask, geom, lanes, target, wgt)
    rng = np.random.default_rng(3)
    h = 1e-6
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in p[name].shape)
            q = {k: v.copy() for k, v in p.items()}
            q[name][idx] += h
            lp, _, _ = M.loss_and_grads(q, x, wave, mask, geom, lanes, target, wgt)
            q[name][idx] -= 2 * h
            lm, _, _ = M.loss_and_grads(q, x, wave, mask, geom, lanes, target, wgt)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, g[name][idx])


def test_adam_reduces_loss():
    x, wave, mask, geom, lanes, target, wgt = problem(4)
    target *= 0.01
    p = M.init_params(4)
    st = {}
    l0, g, _ = M.loss_and_grads(p, x, wave, mask, geom, lanes, target, wgt)
    for _ in range(60):
        _, g, _ = M.loss_and_grads(p, x, wave, mask, geom, lanes, target, wgt)
        p = M.adam_step(p, g, st, lr=3e-3)
    l1, _, _ = M.loss_and_grads(p, x, wave, mask, geom, lanes, target, wgt)
    assert l1 < 0.5 * l0
