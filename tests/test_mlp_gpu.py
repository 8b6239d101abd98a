"""Learned far-field correction on the GPU: MFMA forward/backward against the FP64 oracle with the
same bf16 rounding points, training, training-set generation, and the corrected walk."""
import numpy as np
import pytest

import particular_amd as pa
from oracle import mlp_oracle as M
from particular_amd import learned, metrics, workloads

pytestmark = pytest.mark.gpu


def oracle_params(seed):
    p = M.init_params(seed)
    p["b1"] = np.random.default_rng(seed + 100).normal(0, 0.1, 64)
    p["b2"] = np.random.default_rng(seed + 200).normal(0, 0.1, 64)
    p["b3"] = np.random.default_rng(seed + 300).normal(0, 0.1, 3)
    return p


def set_oracle_params(res, p):
    flat = learned.dict_to_params(p)
    learned.set_params(res, flat)
    return learned.params_to_dict(flat.astype(np.float64))  # exactly what the GPU holds


def random_dataset(seed, waves=40, rows_per_wave=50):
    rng = np.random.default_rng(seed)
    rows = waves * rows_per_wave
    feats = np.zeros((rows, 8), np.float32)
    feats[:, :7] = rng.normal(size=(rows, 7))
    return {
        "feats": feats,
        "scale": rng.uniform(0.5, 2.0, rows).astype(np.float32),
        "mask": rng.integers(0, 2**63, rows, dtype=np.int64).astype(np.uint64),
        "row_wave": np.repeat(np.arange(waves), rows_per_wave).astype(np.uint32),
        "wave_rows": (np.arange(waves + 1) * rows_per_wave).astype(np.uint32),
        "target": rng.normal(size=(waves, 64, 4)).astype(np.float32),
        "wgt": rng.uniform(0.1, 1.0, (waves, 64)).astype(np.float32),
    }


def test_forward_matches_oracle(res):
    import torch

    p = set_oracle_params(res, oracle_params(1))
    x = np.random.default_rng(2).normal(size=(1000, 8)).astype(np.float32)
    x[:, 7] = 0
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((1000, 4), dtype=torch.float32, device="cuda")
    learned.forward(res, xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()[:, :3]
    ref, _ = M.forward(p, x[:, :7].astype(np.float64), bf16=True)
    err = np.abs(y - ref)
    scale = np.abs(ref).max()
    print("forward max abs err", err.max(), "scale", scale, "median", np.median(err))
    assert np.median(err) < 1e-4 * scale and err.max() < 3e-2 * scale


def test_gradients_match_oracle(res):
    p = set_oracle_params(res, oracle_params(3))
    d = random_dataset(4)
    learned.clear(res)
    learned.load(res, d)
    loss, g = learned.grad(res)
    ref_loss, ref_g, _ = M.loss_and_grads(p, d["feats"][:, :7].astype(np.float64),
                                          d["row_wave"].astype(np.int64), d["mask"],
                                          d["scale"].astype(np.float64),
                                          d["target"][..., :3].astype(np.float64),
                                          d["wgt"].astype(np.float64), bf16=True)
    gd = learned.params_to_dict(g)
    print("loss", loss, ref_loss)
    assert abs(loss / ref_loss - 1) < 1e-3
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        rel = np.linalg.norm(gd[k] - ref_g[k]) / np.linalg.norm(ref_g[k])
        print(k, rel)
        assert rel < 2e-2, k


def test_training_fits_a_teacher(res):
    teacher = oracle_params(5)
    d = random_dataset(6, waves=64, rows_per_wave=40)
    y, _ = M.forward(teacher, d["feats"][:, :7].astype(np.float64))
    pred = M.predict(y, d["row_wave"].astype(np.int64), d["mask"], d["scale"].astype(np.float64), 64)
    d["target"][..., :3] = pred
    d["wgt"][:] = 1.0 / (1.0 + (pred ** 2).sum(-1))
    learned.init(res, 7)
    learned.clear(res)
    learned.load(res, d)
    hist = learned.train(res, 300, 3e-3)
    print("loss", hist[0], hist[-1])
    assert hist[-1] < 0.05 * hist[0]


def test_collect_rows_and_targets(res, orc):
    import torch

    s = workloads.direct_sum_case(16384, seed=41)
    dev = torch.from_numpy(s.particles).cuda()
    learned.clear(res)
    learned.collect(res, dev, 0.8, waves=16, seed=3)
    rows, waves, dropped = learned.info(res)
    assert waves == 16 and rows > 16 and dropped == 0
    d = learned.dataset(res)
    assert np.all(np.diff(d["wave_rows"].astype(np.int64)) >= 0) and d["wave_rows"][-1] == rows
    assert np.all(d["scale"] > 0) and np.all(d["mask"] != 0)
    f = d["feats"]
    np.testing.assert_allclose(np.linalg.norm(f[:, :3], axis=1), 1.0, rtol=1e-5)
    # Accepted internal nodes: some lane at distance d_l >= width / theta accepted the node, and
    # that lane is within 8 rms radii of the wavefront centre (64 lanes), so the centre distance
    # d_c >= width / theta - 8 * wrad, i.e. width/d_c <= theta * (1 + 8 * wrad/d_c).
    assert np.all(f[:, 3] > 0)
    assert np.all(f[:, 3] <= 0.8 * (1 + 8 * f[:, 5]) * (1 + 1e-5))
    # targets are brute force minus Barnes-Hut: |target| relative to |a| is a monopole-sized error
    a_ref = 1.0 / np.sqrt(d["wgt"][d["wgt"] > 0])
    t = np.linalg.norm(d["target"][..., :3], axis=-1)[d["wgt"] > 0]
    rel = t / a_ref
    assert 1e-4 < np.sqrt(np.mean(rel**2)) < 0.2


def test_walk_correction_matches_rows(res):
    """The corrected walk adds exactly sum(scale * MLP(row)) over each lane's rows."""
    import torch

    s = workloads.direct_sum_case(8192, seed=43)
    dev = torch.from_numpy(s.particles).cuda()
    p = set_oracle_params(res, oracle_params(9))
    # the recorded rows of every wavefront (all 128 waves, in tree order)
    learned.clear(res)
    learned.collect(res, dev, 0.7, waves=s.n // 64, seed=0)
    d = learned.dataset(res)
    plain = pa.BarnesHut(res, 0.7, pa.Acceleration.checked()).compute(s.particles)
    nn = pa.BarnesHut(res, 0.7, pa.Acceleration.checked(), mlp=True).compute(s.particles)
    _, order = pa.tree_export(res, s.n)
    y, _ = M.forward(p, d["feats"][:, :7].astype(np.float64), bf16=True)
    pred = M.predict(y, d["row_wave"].astype(np.int64), d["mask"], d["scale"].astype(np.float64),
                     s.n // 64).reshape(-1, 3)
    corr = (nn - plain)[order]
    scale = np.abs(corr).max()
    err = np.abs(corr - pred)
    print("correction scale", scale, "max err", err.max())
    assert err.max() < 3e-2 * scale and np.median(err) < 1e-3 * scale
    # a zero last layer leaves the walk bitwise unchanged
    p["W3"][:] = 0
    p["b3"][:] = 0
    set_oracle_params(res, p)
    nn0 = pa.BarnesHut(res, 0.7, pa.Acceleration.checked(), mlp=True).compute(s.particles)
    np.testing.assert_array_equal(nn0, plain)


def test_mlp_flag_needs_initialised_network():
    r = pa.HipResources(0)
    try:
        p = workloads.direct_sum_case(256, seed=1).particles
        with pytest.raises(pa.PnbError) as e:
            pa.BarnesHut(r, 0.5, pa.Acceleration.checked(), mlp=True).compute(p)
        assert e.value.status == pa._lib.PNB_E_STATE
    finally:
        r.close()
