"""Barnes-Hut on the GPU through the C ABI, against the pinned CPU oracle.

Structure: the GPU octree must equal the restated reference octree node for node (same count,
order, levels, particle ranges, internal/leaf split and bit-identical f32 widths).
Accelerations: against the oracle traversal with FP64 centres of mass (the same node data) the
bar is rms <= 1e-6; against the reference-exact f32 centres of mass rms <= 1e-5.
"""
import numpy as np
import pytest

import particular_amd as pa
from particular_amd import metrics, workloads

from .treecheck import assert_reference_com, assert_same_tree, bh, check_against_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (3, 3), (100, 4), (5000, 5)])
def test_tree_matches_oracle_plummer(res, orc, n, seed):
    p = workloads.direct_sum_case(n, seed=seed).particles
    bh(res, p, 0.5)
    assert_same_tree(res, orc, p)


def test_tree_hand_worked_chain(res, orc):
    p = np.array([[0, 0, 0, 1], [0.1, 0, 0, 1], [1, 1, 1, 1]], np.float32)
    bh(res, p, 0.5)
    gn, order = pa.tree_export(res, 3)
    assert gn.shape[0] == 7 and order.tolist() == [2, 1, 0]
    assert gn["level"].tolist() == [0, 1, 1, 2, 3, 4, 4]
    assert_same_tree(res, orc, p)


def test_tree_with_duplicates(res, orc):
    p = workloads.direct_sum_case(800, seed=6).particles
    p = np.concatenate([p, p[:40], p[10:12], p[10:11]])
    bh(res, p, 0.5)
    assert_same_tree(res, orc, p)
    same = np.repeat(p[:1], 9, axis=0)  # all particles at one position: a single leaf root
    bh(res, same, 0.5)
    t = assert_same_tree(res, orc, same)
    assert t.num_nodes == 1


def test_tree_massless_and_cancelling_masses(res, orc):
    """Nodes of zero total mass take the plain mean of their particles (impls/mod.rs:127-128):
    a massless set (every node), and pairs of opposite masses whose nodes sum to exactly zero
    while their children do not."""
    p = workloads.direct_sum_case(600, seed=14).particles
    massless = p.copy()
    massless[:, 3] = 0.0
    bh(res, massless, 0.5)
    assert_same_tree(res, orc, massless)
    pairs = p.copy()
    pairs[:, 3] = np.float32(0.25)
    pairs[1::2, 3] = np.float32(-0.25)
    pairs[1::2, :3] = pairs[0::2, :3] + np.float32(1e-3)  # each pair shares its small nodes
    bh(res, pairs, 0.5)
    t = assert_same_tree(res, orc, pairs)
    gn, _ = pa.tree_export(res, pairs.shape[0])
    zero = (gn["mass64"] == 0.0) & (gn["internal"] == 1)
    assert zero.sum() > 10, zero.sum()  # internal nodes whose masses cancel exactly


@pytest.mark.parametrize("scale", [1e-7, 1e-16, 1e-30])
def test_tree_deep_runs(res, orc, scale):
    """Distinct positions closer than 21 (and 42, 84) levels of subdivision resolve."""
    rng = np.random.default_rng(7)
    base = workloads.direct_sum_case(300, seed=8).particles
    cluster = np.zeros((12, 4), np.float32)
    cluster[:, :3] = np.float32(0.25) + (rng.integers(0, 4, (12, 3)) * scale).astype(np.float32)
    cluster[:, 3] = 1.0 / 300
    tiny = np.zeros((6, 4), np.float32)
    tiny[:, 0] = (np.arange(6) * scale).astype(np.float32)
    tiny[:, 3] = 1.0 / 300
    p = np.concatenate([base, cluster, tiny, cluster[:3]])
    bh(res, p, 0.5)
    t = assert_same_tree(res, orc, p)
    print("max level", t.max_level)


# Sorts of at most 16,384 keys rank by counting; the trees below hold more bodies, so their keys
# take the multi-tile LSD radix sort that builds every large tree (stable order of equal keys,
# deep-run refinement after it).
RADIX_BASE = 17000


def test_tree_with_duplicates_radix_path(res, orc):
    p = workloads.direct_sum_case(RADIX_BASE, seed=16).particles
    p = np.concatenate([p, p[:400], p[10:12], p[10:11], p[::7][:300], p[5:6], p[5:6]])
    assert p.shape[0] > 16384
    bh(res, p, 0.5)
    t = assert_same_tree(res, orc, p)
    assert_reference_com(res, p, t)


def test_tree_massless_and_cancelling_masses_radix_path(res, orc):
    p = workloads.direct_sum_case(RADIX_BASE, seed=17).particles
    massless = p.copy()
    massless[:, 3] = 0.0
    bh(res, massless, 0.5)
    t = assert_same_tree(res, orc, massless)
    assert_reference_com(res, massless, t)
    pairs = p.copy()
    pairs[:, 3] = np.float32(0.25)
    pairs[1::2, 3] = np.float32(-0.25)
    pairs[1::2, :3] = pairs[0::2, :3] + np.float32(1e-3)
    bh(res, pairs, 0.5)
    t = assert_same_tree(res, orc, pairs)
    gn, _ = pa.tree_export(res, pairs.shape[0])
    assert ((gn["mass64"] == 0.0) & (gn["internal"] == 1)).sum() > 10
    assert_reference_com(res, pairs, t)


@pytest.mark.parametrize("scale", [1e-7, 1e-16, 1e-30])
def test_tree_deep_runs_radix_path(res, orc, scale):
    rng = np.random.default_rng(70)
    base = workloads.direct_sum_case(RADIX_BASE, seed=18).particles
    cluster = np.zeros((40, 4), np.float32)
    cluster[:, :3] = np.float32(0.25) + (rng.integers(0, 4, (40, 3)) * scale).astype(np.float32)
    cluster[:, 3] = 1.0 / RADIX_BASE
    tiny = np.zeros((9, 4), np.float32)
    tiny[:, 0] = (np.arange(9) * scale).astype(np.float32)
    tiny[:, 3] = 1.0 / RADIX_BASE
    # the deep structures interleaved with the base, so equal 21-level keys arrive out of order
    p = np.concatenate([base[:8500], cluster, base[8500:], tiny, cluster[:5]])
    bh(res, p, 0.5)
    t = assert_same_tree(res, orc, p)
    print("max level", t.max_level)
    assert_reference_com(res, p, t)


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.5, 0.8])
def test_accel_matches_oracle(res, orc, theta):
    s = workloads.direct_sum_case(20000, seed=9)
    check_against_oracle(res, orc.Tree(s.particles), s.particles, theta)


def test_theta_zero_is_brute_force(res):
    s = workloads.direct_sum_case(3000, seed=10)
    got = bh(res, s.particles, 0.0)
    ref = pa.BruteForce(res, pa.Acceleration.checked()).compute(s.particles)
    assert metrics.rms_rel(got, ref) < 1e-6


def test_between_distinct_sets_and_reordered(res, orc):
    s = workloads.massive_massless_case(n_massive=5000, n_total=30000, softening=1e-3)
    inter = pa.AccelerationSoftened.checked(s.softening)
    r = pa.Reordered.new(s.particles)
    massive = r.affecting()
    got = bh(res, r, 0.5, inter)
    t = orc.Tree(massive)
    ref = t.barnes_hut(s.particles, 0.5, softening=s.softening, com64=True, accum64=True)
    assert metrics.rms_rel(got, ref) < 5e-7
    got2 = bh(res, pa.Between(s.particles, massive), 0.5, inter)
    np.testing.assert_array_equal(got, got2)
    import torch

    dev = bh(res, pa.Reordered.new(torch.from_numpy(s.particles).cuda()), 0.5, inter)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy()[:, :3], got)


def test_edge_cases(res, orc):
    a = workloads.direct_sum_case(100, seed=11).particles
    empty = np.zeros((0, 4), np.float32)
    np.testing.assert_array_equal(bh(res, pa.Between(a, empty), 0.5), np.zeros((100, 3)))
    assert bh(res, pa.Between(empty, a), 0.5).shape == (0, 3)
    one = a[:1].copy()
    np.testing.assert_allclose(bh(res, pa.Between(a, one), 0.7),
                               orc.barnes_hut(a, one, 0.7)[:, :3], rtol=1e-6)
    massless = a.copy()
    massless[:, 3] = 0
    np.testing.assert_array_equal(bh(res, pa.Reordered.new(massless), 0.5), np.zeros((100, 3)))
    # the tree of a massless slice uses plain means (impls/mod.rs:127-128)
    got = bh(res, massless, 0.5)
    np.testing.assert_array_equal(got, np.zeros((100, 3)))
    bad = a.copy()
    bad[5, 0] = np.nan
    with pytest.raises(pa.PnbError) as e:
        bh(res, bad, 0.5)
    assert e.value.status == pa._lib.PNB_E_ARG


def test_two_plummer_reduced(res, orc):
    s = workloads.two_plummer_case(131072)
    t = orc.Tree(s.particles)
    idx = np.random.default_rng(1).choice(s.n, 4096, replace=False)
    for theta in (0.3, 0.5, 0.8):
        check_against_oracle(res, t, s.particles, theta, idx, "two-plummer")
    gn, _ = pa.tree_export(res)
    assert gn.shape[0] == t.num_nodes


def test_stats_count_accepted_terms(res, orc):
    s = workloads.direct_sum_case(4000, seed=12)
    bh(res, s.particles, 0.6)
    steps, accepted = pa.bh_stats(res)
    _, visits = orc.Tree(s.particles).barnes_hut(s.particles, 0.6, com64=True, return_visits=True)
    assert accepted == int(visits.sum())
    assert steps >= accepted / 64


def test_wide_record_offsets_match_narrow(res, monkeypatch):
    """Trees of 2^27 nodes or more are walked with 64-bit record offsets; PNB_WIDE_WALK=1 takes
    that form for a small tree, which must give bit-identical results (plain and corrected, with
    quadrupole outputs and with octupole outputs)."""
    from particular_amd import learned

    s = workloads.direct_sum_case(30000, seed=61)
    learned.init(res, 0)
    runs = {}
    for wide in ("0", "1"):
        monkeypatch.setenv("PNB_WIDE_WALK", wide)
        runs[wide] = [pa.BarnesHut(res, th, pa.Acceleration.checked()).compute(s.particles)
                      for th in (0.5, 0.8)]
        # the corrected walk with quadrupole outputs only, then with octupole outputs
        for params in (learned.quadrupole_params(), learned.octupole_params()):
            learned.set_params(res, params)
            runs[wide].append(pa.BarnesHut(res, 0.5, pa.Acceleration.checked(), mlp=True)
                              .compute(s.particles))
    for a, b in zip(runs["0"], runs["1"]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(runs["0"][2] - runs["0"][0]).max() > 0  # the correction did something
    assert np.abs(runs["0"][3] - runs["0"][2]).max() > 0  # and the octupole too


@pytest.mark.parametrize("m", [4096, 4097, 9000, 16384, 16385])
def test_sort_size_classes_with_ties(res, orc, m):
    """Tree order around the sort's size classes (ranking by counting up to 16,384 keys, the count
    split over workgroups above 4,096; radix passes above), with most bodies sharing a position
    with others (equal keys, which keep their input order), and an affected set of another size
    class sorted along the tree."""
    rng = np.random.default_rng(m)
    p = np.zeros((m, 4), np.float32)
    p[:, :3] = rng.integers(0, 12, size=(m, 3)).astype(np.float32) * 0.25  # 1,728 points
    p[:, 3] = rng.uniform(0.5, 1.5, m).astype(np.float32)
    bh(res, p, 0.5)
    assert_same_tree(res, orc, p)
    a = workloads.direct_sum_case(m // 2 + 3, seed=m).particles
    got = bh(res, pa.Between(a, p), 0.5)
    ref = orc.Tree(p).barnes_hut(a, 0.5, com64=True, accum64=True)[:, :3]
    # per particle, relative to its acceleration's magnitude (a lattice of coincident bodies gives
    # components that cancel to a small fraction of it)
    row = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert metrics.rms_rel(got, ref) < 5e-7 and row.max() < 1e-5, (metrics.rms_rel(got, ref),
                                                                     row.max())


@pytest.mark.parametrize("n", [64, 4096, 70000])
def test_wavefront_width_changes_no_result(res, monkeypatch, n):
    """Each lane adds exactly the nodes of its own traversal, in pre-order, whatever particles
    share its wavefront: walks with 1 to 64 particles per wavefront (PNB_WALK_LANES; small
    problems take narrow ones by default) give bit-identical plain results, for one set and for
    two. The corrected walk adds the same terms, but in partial sums folded every 16 wavefront
    steps, whose grouping follows the wavefront's steps: its results agree to rounding."""
    from particular_amd import learned

    s = workloads.direct_sum_case(n, seed=62)
    other = workloads.direct_sum_case(max(n // 3, 5), seed=63).particles
    learned.init(res, 0)
    learned.set_params(res, learned.quadrupole_params())
    runs = {}
    for lanes in ("64", "16", "3", "1"):
        monkeypatch.setenv("PNB_WALK_LANES", lanes)
        runs[lanes] = [pa.BarnesHut(res, th, pa.Acceleration.checked(), mlp=m).compute(st)
                       for th, m, st in ((0.5, False, s.particles), (0.5, True, s.particles),
                                         (0.7, False, pa.Between(other, s.particles)))]
    for lanes in ("16", "3", "1"):
        for k, (a, b) in enumerate(zip(runs["64"], runs[lanes])):
            if k == 1:  # corrected: the partial sums' grouping
                row = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
                print(f"{lanes} lanes: corrected walk, largest row difference {row.max():.2e}")
                assert metrics.rms_rel(a[:, :3], b[:, :3]) < 1e-6 and row.max() < 1e-5
            else:
                np.testing.assert_array_equal(a, b, err_msg=f"{lanes} lanes")
    monkeypatch.delenv("PNB_WALK_LANES")
    default = pa.BarnesHut(res, 0.5, pa.Acceleration.checked()).compute(s.particles)
    np.testing.assert_array_equal(default, runs["64"][0])


@pytest.mark.parametrize("m", [5, 64, 65, 257])
def test_small_trees_match_oracle(res, orc, m):
    """Trees from a handful of bodies: level lists with one or two internal levels, a root that
    is the only internal node, and affected sets larger than the tree."""
    a = workloads.direct_sum_case(300, seed=90 + m).particles
    src = a[:m].copy()
    bh(res, src, 0.5)
    assert_same_tree(res, orc, src)
    for theta in (0.0, 0.5):
        got = bh(res, pa.Between(a, src), theta)
        ref = orc.Tree(src).barnes_hut(a, theta, com64=True, accum64=True)[:, :3]
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-6)


def test_all_bodies_at_one_position(res, orc):
    """The whole affecting set at one point is one leaf (no internal node, no level list)."""
    src = np.zeros((50, 4), np.float32)
    src[:, :3] = (0.25, -1.0, 3.0)
    src[:, 3] = 1.0 / 50
    a = workloads.direct_sum_case(200, seed=97).particles
    got = bh(res, pa.Between(a, src), 0.5)
    ref = orc.Tree(src).barnes_hut(a, 0.5, com64=True, accum64=True)[:, :3]
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)
    self_acc = bh(res, src, 0.5)
    np.testing.assert_array_equal(self_acc, np.zeros((50, 3)))


# ---- the relative opening criterion (PNB_BH_RELATIVE; no reference counterpart) ----


def rel_bh(res, theta, alpha, mlp=False, estimate_theta=1.2, interaction=None,
           per_particle=False, bucket=24):
    return pa.BarnesHut(res, theta, interaction or pa.Acceleration.checked(), mlp=mlp,
                        relative=alpha, estimate_theta=estimate_theta,
                        estimate_per_particle=per_particle, bucket=bucket)


def oracle_estimate(tree, particles, alpha, estimate_theta):
    """alpha |a~| per particle from the oracle's plain traversal at estimate_theta (the GPU's
    per-particle estimate pass; the two differ only in the last bits of the summed
    accelerations)."""
    a, vis = tree.barnes_hut(particles, estimate_theta, com64=True, accum64=True,
                             return_visits=True)
    return np.float32(alpha) * np.linalg.norm(a[:, :3], axis=1).astype(np.float32), vis


def oracle_wave_estimate(tree, particles, alpha, estimate_theta):
    """The GPU's default estimate for Between(p, p), whose walk order is the tree order: alpha |a~|
    at the centroid of every 64 consecutive particles in tree order, given to each of them (in
    the caller's order)."""
    order = tree.order()
    q = particles[order].astype(np.float64)
    waves = (len(q) + 63) // 64
    cen = np.zeros((waves, 4), np.float32)
    for w in range(waves):
        cen[w, :3] = q[64 * w:64 * w + 64, :3].mean(0)
    a, vis = tree.barnes_hut(cen, estimate_theta, com64=True, accum64=True, return_visits=True)
    ew = np.float32(alpha) * np.linalg.norm(a[:, :3], axis=1).astype(np.float32)
    est = np.zeros(len(q), np.float32)
    est[order] = ew[np.arange(len(q)) // 64]
    return est, vis


@pytest.mark.parametrize("alpha,per_particle", [(1e-3, False), (1e-4, False), (1e-4, True)])
def test_relative_criterion_matches_oracle(res, orc, alpha, per_particle):
    """Plain walk with the relative rule (p = 2) against the oracle's restatement fed the oracle's
    own estimate (per wavefront centroid, the default, or per particle). A decision can flip where
    the two sides' last-bit rounding (the GPU's v_rsq, its f32 centroids and compensated sums)
    straddles the threshold; each flip moves one particle by at most about alpha |a|, hence the
    alpha term in the bar. The accepted-term counters (estimate pass + walk) agree to 1e-4."""
    p = workloads.direct_sum_case(20000, seed=41).particles
    t = orc.Tree(p)
    got = rel_bh(res, 1.0, alpha, per_particle=per_particle).compute(p)
    _, accepted = pa.bh_stats(res)
    est, est_vis = (oracle_estimate if per_particle else oracle_wave_estimate)(t, p, alpha, 1.2)
    ref, vis = t.barnes_hut_rel(p, 1.0, est, order=2, com64=True, accum64=True,
                                return_visits=True, bucket=24)
    e = metrics.summary(got, ref)
    print("alpha", alpha, "per particle", per_particle, "vs oracle", e, "accepted", accepted,
          int(est_vis.sum() + vis.sum()))
    assert e["rms_rel"] < 5e-7 + 0.02 * alpha and e["max_rel"] < 1e-4 + 3 * alpha
    assert abs(accepted - int(est_vis.sum() + vis.sum())) <= 1e-4 * accepted
    exact = orc.direct_f64(p, p)
    assert metrics.rms_rel(got, exact) < 5 * alpha


def test_relative_criterion_limits_are_geometric_walks(res):
    """Without buckets, alpha = 1e30 never opens by the relative rule: the geometric walk at the
    same theta, bit-identical for the plain walk, to rounding for the corrected one (its geometric
    form sums in partial sums, its relative form term by term); alpha = 0 opens every massive
    internal node: bit-identical to theta = 0 (where the corrected walk, accepting only leaves,
    equals the plain one). With buckets, alpha = 0 sums the same pair terms as theta = 0 in
    another order."""
    from particular_amd import learned

    p = workloads.direct_sum_case(30000, seed=43).particles
    learned.init(res, 0)
    learned.set_params(res, learned.quadrupole_params())
    chk = pa.Acceleration.checked()
    every = pa.BarnesHut(res, 0.0, chk).compute(p)
    for mlp in (False, True):
        geo = pa.BarnesHut(res, 0.6, chk, mlp=mlp).compute(p)
        big = rel_bh(res, 0.6, 1e30, mlp=mlp, bucket=0).compute(p)
        if mlp:
            row = np.linalg.norm(big - geo, axis=1) / np.linalg.norm(geo, axis=1)
            assert metrics.rms_rel(big, geo) < 1e-6 and row.max() < 1e-5, row.max()
        else:
            np.testing.assert_array_equal(big, geo)
        np.testing.assert_array_equal(rel_bh(res, 0.6, 0.0, mlp=mlp, bucket=0).compute(p), every)
        np.testing.assert_allclose(rel_bh(res, 0.6, 0.0, mlp=mlp).compute(p), every, rtol=1e-5,
                                   atol=1e-7)


def test_relative_corrected_walk(res, orc):
    """With the learned quadrupole network the rule bounds the octupole (p = 3): at the same
    alpha the corrected walk is more accurate than the plain relative walk and accepts fewer
    terms; against the oracle's p = 3 decisions the plain-part sums agree."""
    from particular_amd import learned

    p = workloads.direct_sum_case(20000, seed=45).particles
    learned.init(res, 0)
    learned.set_params(res, learned.quadrupole_params())
    exact = orc.direct_f64(p, p)
    alpha = 1e-4
    plain = rel_bh(res, 1.0, alpha).compute(p)
    _, acc_plain = pa.bh_stats(res)
    corr = rel_bh(res, 1.0, alpha, mlp=True).compute(p)
    _, acc_corr = pa.bh_stats(res)
    e_plain, e_corr = metrics.rms_rel(plain, exact), metrics.rms_rel(corr, exact)
    print("plain", e_plain, acc_plain, "corrected", e_corr, acc_corr)
    assert e_corr < e_plain and acc_corr < acc_plain
    assert e_corr < 5 * alpha
    # the decisions: the oracle's p = 3 traversal accepts the same number of terms
    t = orc.Tree(p)
    est, est_vis = oracle_wave_estimate(t, p, alpha, 1.2)
    _, vis = t.barnes_hut_rel(p, 1.0, est, order=3, com64=True, return_visits=True, bucket=24)
    assert abs(acc_corr - int(est_vis.sum() + vis.sum())) <= 1e-4 * acc_corr


def test_relative_wide_offsets_match_narrow(res, monkeypatch):
    from particular_amd import learned

    p = workloads.direct_sum_case(30000, seed=47).particles
    learned.init(res, 0)
    learned.set_params(res, learned.quadrupole_params())
    runs = {}
    for wide in ("0", "1"):
        monkeypatch.setenv("PNB_WIDE_WALK", wide)
        runs[wide] = [rel_bh(res, 1.0, 1e-3, mlp=m, per_particle=pp).compute(p)
                      for m in (False, True) for pp in (False, True)]
    for a, b in zip(runs["0"], runs["1"]):
        np.testing.assert_array_equal(a, b)


def test_relative_storages_and_edge_cases(res, orc):
    """Between distinct sets (softened, massless affected), Reordered, device storage, an empty
    affecting set, a single affecting body, a whole affecting set at one point, bad parameters."""
    import torch

    s = workloads.massive_massless_case(n_massive=5000, n_total=30000, softening=1e-3)
    inter = pa.AccelerationSoftened.checked(s.softening)
    r = pa.Reordered.new(s.particles)
    massive = r.affecting()
    # per-particle estimates: the walk order of a distinct affected set is its own key order
    got = rel_bh(res, 1.0, 1e-4, interaction=inter, per_particle=True).compute(r)
    t = orc.Tree(massive)
    est, _ = t.barnes_hut(s.particles, 1.2, softening=s.softening, com64=True, accum64=True,
                          return_visits=True)
    est = np.float32(1e-4) * np.linalg.norm(est[:, :3], axis=1).astype(np.float32)
    ref = t.barnes_hut_rel(s.particles, 1.0, est, order=2, softening=s.softening, com64=True,
                           accum64=True, bucket=24)
    assert metrics.rms_rel(got, ref) < 5e-7 + 2e-6
    got2 = rel_bh(res, 1.0, 1e-4, interaction=inter, per_particle=True).compute(
        pa.Between(s.particles, massive))
    np.testing.assert_array_equal(got, got2)
    dev = rel_bh(res, 1.0, 1e-4, interaction=inter, per_particle=True).compute(
        pa.Reordered.new(torch.from_numpy(s.particles).cuda()))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy()[:, :3], got)
    # per-wavefront estimates on the same storages: the same storage-independence, and the
    # error against the FP64 direct sum of a plain relative walk
    wav = rel_bh(res, 1.0, 1e-4, interaction=inter).compute(r)
    np.testing.assert_array_equal(
        rel_bh(res, 1.0, 1e-4, interaction=inter).compute(pa.Between(s.particles, massive)), wav)
    idx = np.random.default_rng(3).choice(len(s.particles), 2000, replace=False)
    exact = orc.direct_f64(s.particles[idx], massive, softening=s.softening)
    assert metrics.rms_rel(wav[idx], exact) < 5e-4

    a = workloads.direct_sum_case(100, seed=48).particles
    empty = np.zeros((0, 4), np.float32)
    np.testing.assert_array_equal(rel_bh(res, 1.0, 1e-4).compute(pa.Between(a, empty)),
                                  np.zeros((100, 3)))
    one = a[:1].copy()
    np.testing.assert_allclose(rel_bh(res, 1.0, 1e-4).compute(pa.Between(a, one)),
                               orc.barnes_hut(a, one, 0.7)[:, :3], rtol=1e-6)
    src = np.zeros((50, 4), np.float32)
    src[:, :3] = (0.25, -1.0, 3.0)
    src[:, 3] = 1.0 / 50
    np.testing.assert_allclose(rel_bh(res, 1.0, 1e-4).compute(pa.Between(a, src)),
                               orc.Tree(src).barnes_hut(a, 0.5, com64=True, accum64=True)[:, :3],
                               rtol=1e-6, atol=1e-7)
    L = pa._lib.lib()
    for alpha, te, bucket in ((-1.0, 1.2, 16), (float("nan"), 1.2, 16), (1e-4, 0.0, 16),
                              (1e-4, float("inf"), 16), (1e-4, 1.2, -1), (1e-4, 1.2, 32)):
        assert L.pnb_bh_set_relative(res.handle, alpha, te, 0, bucket) == pa._lib.PNB_E_ARG


@pytest.mark.parametrize("n", [400_000, 2_200_000])
def test_relative_walk_large(res, orc, n):
    """400K bodies (6,250 wavefronts: launched heaviest first, ordered by the multi-tile sort) and
    2.2M (above 128 wavefronts per CU: tree order); both walks bounded against the FP64 direct sum
    on a sample, repeatable bit for bit, and the corrected one more accurate than the plain one at
    the same alpha."""
    from particular_amd import learned

    s = workloads.two_plummer_case(n, seed=77)
    idx = np.random.default_rng(2).choice(s.n, 2048, replace=False)
    exact = orc.direct_f64(s.particles[idx], s.particles)
    learned.init(res, 0)
    learned.set_params(res, learned.quadrupole_params())
    errs = {}
    for mlp in (False, True):
        alg = rel_bh(res, 1.0, 1e-4, mlp=mlp)
        a = alg.compute(s.particles)
        np.testing.assert_array_equal(alg.compute(s.particles), a)
        errs[mlp] = metrics.rms_rel(a[idx], exact)
    print(n, "relative walk rms (plain, corrected)", errs)
    assert errs[False] < 5e-4 and errs[True] < errs[False]
