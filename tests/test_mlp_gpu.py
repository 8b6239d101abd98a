"""Learned far-field correction on the GPU: MFMA forward/backward against the FP64 oracle with the
same bf16 rounding points, training, training-set generation (rows against a CPU emulation of the
walk and the feature oracle, node quadrupoles against their particles), and the corrected walk."""
import numpy as np
import pytest

import particular_amd as pa
from oracle import features as F
from oracle import mlp_oracle as M
from particular_amd import learned, workloads

pytestmark = pytest.mark.gpu

NF = M.IN  # features the network reads (of learned.FEAT stored per row)


def oracle_params(seed):
    p = M.init_params(seed)
    p["b1"] = np.random.default_rng(seed + 100).normal(0, 0.1, 64)
    p["b2"] = np.random.default_rng(seed + 200).normal(0, 0.1, 64)
    p["b3"] = np.random.default_rng(seed + 300).normal(0, 0.1, M.OUT)
    return p


def set_oracle_params(res, p):
    flat = learned.dict_to_params(p)
    learned.set_params(res, flat)
    return learned.params_to_dict(flat.astype(np.float64))  # exactly what the GPU holds


def random_dataset(seed, waves=40, rows_per_wave=50):
    rng = np.random.default_rng(seed)
    rows = waves * rows_per_wave
    feats = np.zeros((rows, learned.FEAT), np.float32)
    feats[:, :NF] = rng.normal(size=(rows, NF))
    u = rng.normal(size=(rows, 3))
    geom = np.zeros((rows, learned.GEOM), np.float32)
    geom[:, :3] = 3.0 * u / np.linalg.norm(u, axis=1, keepdims=True)
    geom[:, 3] = rng.uniform(0.5, 2.0, rows)  # M w^2
    geom[:, 4] = geom[:, 3] * rng.uniform(0.5, 1.5, rows)  # M w^3
    lanes = np.zeros((waves, 64, 4), np.float32)
    lanes[..., :3] = rng.normal(size=(waves, 64, 3)) * 0.3
    lanes[..., 3] = 1e-4
    return {
        "feats": feats,
        "geom": geom,
        "mask": rng.integers(0, 2**63, rows, dtype=np.int64).astype(np.uint64),
        "row_wave": np.repeat(np.arange(waves), rows_per_wave).astype(np.uint32),
        "wave_rows": (np.arange(waves + 1) * rows_per_wave).astype(np.uint32),
        "lanes": lanes,
        "target": (rng.normal(size=(waves, 64, 4)) * 0.01).astype(np.float32),
        "wgt": rng.uniform(0.1, 1.0, (waves, 64)).astype(np.float32),
    }


def oracle_args(d):
    return (d["row_wave"].astype(np.int64), d["mask"], d["geom"].astype(np.float64),
            d["lanes"].astype(np.float64))


def test_forward_matches_oracle(res):
    import torch

    p = set_oracle_params(res, oracle_params(1))
    x = np.zeros((1000, learned.FEAT), np.float32)
    x[:, :NF] = np.random.default_rng(2).normal(size=(1000, NF))
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((1000, learned.YS), dtype=torch.float32, device="cuda")
    learned.forward(res, xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    ref, _ = M.forward(p, x[:, :NF].astype(np.float64), bf16=True)
    err = np.abs(y[:, :M.OUT] - ref)
    scale = np.abs(ref).max()
    print("forward max abs err", err.max(), "scale", scale, "median", np.median(err))
    assert np.median(err) < 1e-4 * scale and err.max() < 3e-2 * scale
    assert not y[:, M.OUT:].any()


def test_gradients_match_oracle(res):
    p = set_oracle_params(res, oracle_params(3))
    d = random_dataset(4)
    learned.clear(res)
    learned.load(res, d)
    loss, g = learned.grad(res)
    ref_loss, ref_g, _ = M.loss_and_grads(p, d["feats"][:, :NF].astype(np.float64),
                                          *oracle_args(d),
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
    teacher["W3"] = np.random.default_rng(6).normal(0, 0.2, teacher["W3"].shape)
    d = random_dataset(6, waves=64, rows_per_wave=40)
    y, _ = M.forward(teacher, d["feats"][:, :NF].astype(np.float64))
    pred = M.predict(y, *oracle_args(d), 64)
    d["target"][..., :3] = pred
    d["wgt"][:] = 1.0 / (1e-4 + (pred ** 2).sum(-1))
    learned.init(res, 7)
    learned.clear(res)
    learned.load(res, d)
    hist = learned.train(res, 300, 3e-3)
    print("loss", hist[0], hist[-1])
    assert hist[-1] < 0.05 * hist[0]


def test_collect_rows_and_targets(res):
    import torch

    s = workloads.direct_sum_case(16384, seed=41)
    dev = torch.from_numpy(s.particles).cuda()
    learned.clear(res)
    learned.collect(res, dev, 0.8, waves=16, seed=3)
    rows, waves, dropped = learned.info(res)
    assert waves == 16 and rows > 16 and dropped == 0
    d = learned.dataset(res)
    assert np.all(np.diff(d["wave_rows"].astype(np.int64)) >= 0) and d["wave_rows"][-1] == rows
    assert np.all(d["geom"][:, 3] > 0) and np.all(d["mask"] != 0)
    f = d["feats"]
    # Every body of a node lies within w of its centre of mass along each axis, so the entries
    # of Q = (3T - tr(T) I) / (M w^2) are at most 2 (diagonal) and 3 (off-diagonal)
    assert np.all(np.abs(f[:, 0:3]) <= 2.0) and np.all(np.abs(f[:, 3:6]) <= 3.0)
    np.testing.assert_allclose(f[:, 0] + f[:, 1] + f[:, 2], 0.0, atol=1e-5)  # traceless
    assert np.all((f[:, 6] > 0) & (f[:, 6] <= 1.0 + 1e-6))  # mass fraction
    assert np.all(f[:, 7] >= 1.0 / 32.0 - 1e-6)             # at least two bodies
    # 32 x normalised octupole (5/2) U~ / (M w^3): every body is within sqrt(3) w of the centre of
    # mass, so |U_ijk| <= 3^1.5 M w^3, |V_k| <= 3^2.5 M w^3 and |U~_ijk| <= 3^1.5 + 3^2.5 * 3 / 5
    assert np.all(np.abs(f[:, 9:16]) < 32 * 2.5 * (3**1.5 + 0.6 * 3**2.5))
    assert np.abs(f[:, 9:16]).max() > 0
    np.testing.assert_allclose(d["lanes"][..., 3], 0.0)  # softening^2
    # targets are brute force minus Barnes-Hut: |target| relative to |a| is a monopole-sized error
    a_ref = 1.0 / np.sqrt(d["wgt"][d["wgt"] > 0])
    t = np.linalg.norm(d["target"][..., :3], axis=-1)[d["wgt"] > 0]
    rel = t / a_ref
    assert 1e-4 < np.sqrt(np.mean(rel**2)) < 0.2


def cpu_rows(nodes, pts, theta):
    """Rows of one wavefront by the reference's acceptance rule in f32 (sequential.rs:485-494):
    the internal nodes some lane accepts, in pre-order, with the accepting-lane mask."""
    f32 = np.float32
    th2 = f32(theta) * f32(theta)
    com = nodes["com"].astype(np.float32)
    w2 = (nodes["width"] * nodes["width"]).astype(np.float32)
    internal = nodes["internal"] == 1
    masks = {}
    for lane, p in enumerate(pts.astype(np.float32)):
        k = 0
        while k < len(nodes):
            d = com[k] - p[:3]
            d2 = f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2])
            if d2 == 0:
                k = nodes["skip"][k]
            elif internal[k] and f32(th2 * d2) < w2[k]:
                k += 1
            else:
                if internal[k]:
                    masks[k] = masks.get(k, 0) | (1 << lane)
                k = nodes["skip"][k]
    return [(k, masks[k]) for k in sorted(masks)]


def test_collected_rows_match_cpu_walk(res):
    """Row lists and masks exactly as the reference rule implies; node features and geometry as
    the oracle computes them from the exported tree."""
    import torch

    s = workloads.two_plummer_case(2048, seed=45)
    dev = torch.from_numpy(s.particles).cuda()
    learned.init(res, 1)
    learned.clear(res)
    theta = 0.6
    learned.collect(res, dev, theta, waves=s.n // 64, seed=0)
    d = learned.dataset(res)
    nodes, order = pa.tree_export(res, s.n)
    sp = s.particles[order]
    for w in (0, 5, 17, 31):
        b, e = int(d["wave_rows"][w]), int(d["wave_rows"][w + 1])
        pts = sp[64 * w:64 * w + 64]
        np.testing.assert_array_equal(d["lanes"][w, :, :3], pts[:, :3])
        want = cpu_rows(nodes, pts, theta)
        assert e - b == len(want) > 0
        assert [int(m) for m in d["mask"][b:e]] == [m for _, m in want], w
        for r, (k, m) in zip(range(b, e), want):
            n = nodes[k]
            np.testing.assert_allclose(d["feats"][r, :NF], F.node_features(nodes, k, sp),
                                       rtol=2e-5, atol=2e-6)
            np.testing.assert_array_equal(d["geom"][r, :3], n["com"])
            assert abs(d["geom"][r, 3] / (float(n["mass"]) * float(n["width"]) ** 2) - 1) < 1e-6
            assert abs(d["geom"][r, 4] / (float(n["mass"]) * float(n["width"]) ** 3) - 1) < 1e-6


def test_node_quadrupoles_match_particles(res):
    import torch

    s = workloads.two_plummer_case(8192, seed=46)
    dev = torch.from_numpy(s.particles).cuda()
    learned.init(res, 1)
    pa.BarnesHut(res, 0.5, pa.Acceleration.checked(), mlp=True).compute(pa.Between(dev, dev))
    nodes, order = pa.tree_export(res, s.n)
    sp = s.particles[order].astype(np.float64)
    internal = np.flatnonzero(nodes["internal"] == 1)
    rng = np.random.default_rng(0)
    worst = 0.0
    for k in np.concatenate([internal[:50], rng.choice(internal, 200, replace=False)]):
        n = nodes[k]
        part = sp[n["begin"]:n["begin"] + n["count"]]
        want = F.quadrupole(part[:, :3], part[:, 3], float(n["width"]))
        err = np.abs(n["quad"] - want).max()
        worst = max(worst, err)
        assert err < 1e-5 + 1e-5 * np.abs(want).max(), (k, n["quad"], want)
    assert not nodes["quad"][nodes["internal"] == 0].any()
    print("worst quadrupole error", worst)


@pytest.mark.parametrize("case", ["two_plummer", "uniform_with_massless"])
def test_node_features_match_bodies(res, case):
    """Every input of the node network (pnb_bh_node_features) against FP64 sums over the node's
    own bodies (oracle/features.py): mass share, quadrupole, and the normalised octupole from
    third moments summed body by body -- independent of the bottom-up moments the library builds
    them from. Structure (ranges, levels) comes from the export, which the tree tests pin."""
    import torch

    if case == "two_plummer":
        p = workloads.two_plummer_case(8192, seed=48).particles
    else:
        p = workloads.direct_sum_case(6000, seed=49).particles.copy()
        p[::4, 3] = 0.0  # massless bodies inside massive nodes, and some massless nodes
    dev = torch.from_numpy(p).cuda()
    learned.init(res, 1)
    pa.BarnesHut(res, 0.5, pa.Acceleration.checked(), mlp=True).compute(pa.Between(dev, dev))
    nodes, order = pa.tree_export(res, p.shape[0])
    got = learned.node_features(res)
    assert got.shape == (nodes.shape[0], learned.FEAT)
    assert not got[nodes["internal"] == 0].any()
    sp = p[order]
    internal = np.flatnonzero(nodes["internal"] == 1)
    rng = np.random.default_rng(0)
    pick = np.unique(np.concatenate([internal[:40], internal[-40:],
                                     rng.choice(internal, 300, replace=False)]))
    worst = np.zeros(16)
    for k in pick:
        want = F.node_features(nodes, int(k), sp)
        err = np.abs(got[k] - want)
        worst = np.maximum(worst, err)
        np.testing.assert_allclose(got[k], want, rtol=1e-5, atol=2e-6, err_msg=str(k))
    print("worst feature error", worst)
    assert np.abs(got[pick, 9:16]).max() > 1e-2  # the octupole features are not all zero


def test_walk_correction_matches_rows(res):
    """The corrected walk adds, for every lane, the field of M w^2 MLP(node) of each internal node
    it accepted: the oracle's prediction from the recorded rows."""
    import torch

    s = workloads.direct_sum_case(8192, seed=43)
    dev = torch.from_numpy(s.particles).cuda()
    p = oracle_params(9)
    p["W3"] = np.random.default_rng(10).normal(0, 0.2, p["W3"].shape)
    p = set_oracle_params(res, p)
    # the recorded rows of every wavefront (all 128 waves, in tree order)
    learned.clear(res)
    learned.collect(res, dev, 0.7, waves=s.n // 64, seed=0)
    d = learned.dataset(res)
    plain = pa.BarnesHut(res, 0.7, pa.Acceleration.checked()).compute(s.particles)
    nn = pa.BarnesHut(res, 0.7, pa.Acceleration.checked(), mlp=True).compute(s.particles)
    _, order = pa.tree_export(res, s.n)
    y, _ = M.forward(p, d["feats"][:, :NF].astype(np.float64), bf16=True)
    y = learned.walk_outputs(y)  # the walk's rule for small octupoles
    pred = M.predict(y, *oracle_args(d), s.n // 64).reshape(-1, 3)
    corr = (nn - plain)[order][:, :3]
    scale = np.abs(corr).max()
    err = np.abs(corr - pred)
    print("correction scale", scale, "max err", err.max())
    assert scale > 0
    assert err.max() < 3e-2 * scale and np.median(err) < 1e-3 * scale


def test_quadrupole_network_reduces_error(res, orc):
    """The network set to output each node's own quadrupole (learned.quadrupole_params) turns the
    walk into quadrupole Barnes-Hut: it must cut the error against FP64 well below plain."""
    s = workloads.two_plummer_case(32768, seed=47)
    idx = np.random.default_rng(1).choice(s.n, 512, replace=False)
    ref = orc.direct_f64(s.particles[idx], s.particles)
    learned.set_params(res, learned.quadrupole_params())

    def rms(a):
        e = np.linalg.norm(a[idx, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return float(np.sqrt(np.mean(e**2)))

    for theta in (0.5, 0.8):
        plain = pa.BarnesHut(res, theta, pa.Acceleration.checked()).compute(s.particles)
        quad = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=True).compute(s.particles)
        print(f"theta {theta}: plain rms {rms(plain):.3e}, quadrupole network rms {rms(quad):.3e}")
        assert rms(quad) < 0.5 * rms(plain)


def test_octupole_network_reduces_error_below_the_quadrupole(res, orc):
    """The network set to pass each node's own normalised octupole to its octupole outputs
    (learned.octupole_params) adds the multipole expansion's octupole term: at the same theta its
    error against FP64 must fall well below the quadrupole term's."""
    s = workloads.two_plummer_case(32768, seed=47)
    idx = np.random.default_rng(1).choice(s.n, 512, replace=False)
    ref = orc.direct_f64(s.particles[idx], s.particles)

    def rms(a):
        e = np.linalg.norm(a[idx, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return float(np.sqrt(np.mean(e**2)))

    # the terms beyond the octupole grow with theta: measured 0.42 at 0.5, 0.67 at 0.8
    for theta, bar in ((0.5, 0.5), (0.8, 0.75)):
        learned.set_params(res, learned.quadrupole_params())
        quad = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=True).compute(s.particles)
        learned.set_params(res, learned.octupole_params())
        octu = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=True).compute(s.particles)
        print(f"theta {theta}: quadrupole rms {rms(quad):.3e}, octupole rms {rms(octu):.3e}")
        assert rms(octu) < bar * rms(quad)


def test_untrained_network_is_the_quadrupole_term(res):
    """A fresh network has a zero output layer, so it adds nothing to its residual: the corrected
    walk is quadrupole Barnes-Hut (the all-zero learned.quadrupole_params), bit for bit, and not
    plain Barnes-Hut."""
    s = workloads.direct_sum_case(8192, seed=44)
    learned.init(res, 3)
    nn0 = pa.BarnesHut(res, 0.7, pa.Acceleration.checked(), mlp=True).compute(s.particles)
    learned.set_params(res, learned.quadrupole_params())
    quad = pa.BarnesHut(res, 0.7, pa.Acceleration.checked(), mlp=True).compute(s.particles)
    plain = pa.BarnesHut(res, 0.7, pa.Acceleration.checked()).compute(s.particles)
    np.testing.assert_array_equal(nn0, quad)
    assert np.abs(nn0 - plain).max() > 0


@pytest.mark.parametrize("n", [3000, 120000])
def test_moments_build_keeps_the_tree(res, n):
    """A build with node moments (the corrected walk's) and a plain build of the same bodies give
    the same nodes bit for bit: centres of mass, masses (f32 and FP64), widths, skips, ranges and
    levels. The bottom-up pass has an instance with and one without the moments; 3,000 bodies
    keep every level in the single-workgroup runs of small levels, 120,000 use the per-level
    launches too. The bodies include massless ones (zero-mass nodes take the mean position)."""
    p = workloads.direct_sum_case(n, seed=51).particles.copy()
    p[::5, 3] = 0.0  # every fifth body massless
    learned.init(res, 5)
    pa.BarnesHut(res, 0.6, pa.Acceleration.checked(), mlp=True).compute(p)
    with_moments, _ = pa.tree_export(res, n)
    pa.BarnesHut(res, 0.6, pa.Acceleration.checked()).compute(p)
    plain, _ = pa.tree_export(res, n)
    assert with_moments.shape == plain.shape and plain.shape[0] > n // 2
    assert (plain["mass"] == 0).any() and (plain["mass"] > 0).any()
    for f in ("com", "mass", "mass64", "width", "skip", "begin", "count", "level", "internal"):
        np.testing.assert_array_equal(with_moments[f], plain[f], err_msg=f)


def test_mlp_flag_needs_initialised_network():
    r = pa.HipResources(0)
    try:
        p = workloads.direct_sum_case(256, seed=1).particles
        with pytest.raises(pa.PnbError) as e:
            pa.BarnesHut(r, 0.5, pa.Acceleration.checked(), mlp=True).compute(p)
        assert e.value.status == pa._lib.PNB_E_STATE
    finally:
        r.close()


def test_trained_network_from_the_quadrupole_start(res, orc):
    """The whole learned path: a fresh network (zero output layer, i.e. the quadrupole term)
    trained with the GPU forward/backward/Adam kernels on the brute-force minus Barnes-Hut
    residuals the walk itself collects, then judged on a held-out snapshot against the FP64 direct
    sum. Its octupole outputs start at zero, so whatever it gains over the quadrupole term is
    learned from the residuals: measured 0.63x (theta 0.5) and 0.70x (theta 0.8) of the
    quadrupole term's error after 300 steps on three snapshots (the 4M-body comparison is
    nn_bh's)."""
    import torch

    n = 16384
    learned.init(res, 0)
    learned.clear(res)
    for seed in (71, 72, 73):
        s = workloads.two_plummer_case(n, seed=seed)
        dev = torch.from_numpy(s.particles).cuda()
        for theta in (0.5, 0.8):
            learned.collect(res, dev, theta, waves=n // 64, seed=seed)
    out = learned.fit(res, 300)
    print("training loss", out["loss_start"], "->", out["loss_end"], "kept", out["kept"])
    assert out["kept"] and out["loss_end"] < 0.7 * out["loss_start"]
    trained = learned.get_params(res)

    h = workloads.two_plummer_case(n, seed=80)
    idx = np.random.default_rng(3).choice(n, 512, replace=False)
    ref = orc.direct_f64(h.particles[idx], h.particles)

    def rms(a):
        e = np.linalg.norm(a[idx, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return float(np.sqrt(np.mean(e**2)))

    for theta, bar in ((0.5, 0.5), (0.8, 0.75)):
        plain = pa.BarnesHut(res, theta, pa.Acceleration.checked()).compute(h.particles)
        learned.set_params(res, trained)
        nn = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=True).compute(h.particles)
        learned.set_params(res, learned.quadrupole_params())
        quad = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=True).compute(h.particles)
        print(f"theta {theta}: plain {rms(plain):.3e}, trained {rms(nn):.3e}, "
              f"quadrupole {rms(quad):.3e}")
        assert rms(nn) < bar * rms(plain) and rms(nn) < 0.85 * rms(quad)
