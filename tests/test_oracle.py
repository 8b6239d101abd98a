"""Pins the CPU oracle (oracle/oracle.c) before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement fails it: a dropped or squared
mu, a wrong sign, a wrong fold order, a wrong child-bit rule, a wrong node order or width, a
missing d^2 == 0 skip. The expected values come from closed forms, hand-worked trees, FP64
arithmetic done independently in numpy, or invariants -- never from the oracle's own formulas.
"""
import numpy as np
import pytest

from particular_amd import metrics, workloads
from tests.shared import circular_orbit, kat_particles




def numpy_direct_f64(affected, affecting, eps=0.0):
    """Exhaustive FP64 pair sum with numpy broadcasting (independent of the C code)."""
    a = affected[:, None, :3].astype(np.float64)
    b = affecting[None, :, :3].astype(np.float64)
    d = b - a
    r2 = (d * d).sum(-1)
    rs = r2 + eps * eps
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(r2 > 0, affecting[None, :, 3].astype(np.float64) / (rs * np.sqrt(rs)), 0.0)
    return (d * s[..., None]).sum(1)


def test_kat_brute_force(orc):
    p = kat_particles()
    massive = p[p[:, 3] != 0]
    got = orc.brute_force(p, massive)
    # the reference test's own expected value: dir * m * (1/mag2).sqrt() / mag2, skipping mag2 == 0
    exp = np.zeros((6, 3))
    for i, p1 in enumerate(p.astype(np.float64)):
        for p2 in massive.astype(np.float64):
            d = p2[:3] - p1[:3]
            m2 = d @ d
            if m2 != 0:
                exp[i] += d * p2[3] * np.sqrt(1.0 / m2) / m2
    assert metrics.max_rel(got, exp) < 1e-6
    # the reference's own componentwise criterion (mod.rs:268-274) at its tolerance
    err = np.linalg.norm(1.0 - got[:, :3] / exp, axis=1)
    assert np.all(err <= 1e-2)


def test_two_body_closed_form(orc):
    for r, mu in [(1.0, 1.0), (3.0, 2.5), (0.01, 1e-6), (1e3, 1e9)]:
        a = orc.accel_pair([r, 0, 0, mu], [0, 0, 0])
        np.testing.assert_allclose(a, [mu / r**2, 0, 0], rtol=3e-7)
        a = orc.accel_pair([0, 0, 0, mu], [0, -r, 0])  # pulled towards the body: +y
        np.testing.assert_allclose(a, [0, mu / r**2, 0], rtol=3e-7)


@pytest.mark.parametrize("n,m,eps", [(1, 1, 0.0), (7, 13, 0.0), (64, 64, 0.0), (33, 17, 0.05)])
def test_bf_vs_exhaustive_f64(orc, n, m, eps):
    rng = np.random.default_rng(n * 1000 + m)
    a = rng.normal(size=(n, 4)).astype(np.float32)
    b = rng.normal(size=(m, 4)).astype(np.float32)
    b[:, 3] = rng.uniform(0.1, 2.0, m)
    got = orc.brute_force(a, b, softening=eps)
    exp = numpy_direct_f64(a, b, eps)
    assert metrics.max_rel(got, exp) < 2e-5
    np.testing.assert_allclose(orc.direct_f64(a, b, eps), exp, rtol=1e-12, atol=1e-300)


def test_bf_mu_enters_linearly(orc):
    s = workloads.direct_sum_case(256, seed=3)
    b2 = s.particles.copy()
    b2[:, 3] *= 2.0  # exact in f32: every rounded step scales by exactly 2
    np.testing.assert_array_equal(orc.brute_force(s.particles, b2),
                                  2.0 * orc.brute_force(s.particles, s.particles))


def test_bf_fold_order_is_index_order(orc):
    # contributions +1e8, +1, -1e8 along x: the reference's left fold gives (1e8 + 1) - 1e8 = 0 in
    # f32, any other order that adds the 1 last gives 1.
    at = np.array([[0, 0, 0, 0]], np.float32)
    b = np.array([[1, 0, 0, 1e8], [1, 0, 0, 1], [-1, 0, 0, 1e8]], np.float32)
    assert orc.brute_force(at, b)[0, 0] == 0.0
    assert orc.brute_force(at, b[[0, 2, 1]])[0, 0] == 1.0


def test_checked_and_unchecked_semantics(orc):
    at = np.array([[1, 2, 3, 5]], np.float32)
    b = np.array([[1, 2, 3, 5], [2, 2, 3, 1]], np.float32)
    a = orc.brute_force(at, b, checked=True)
    np.testing.assert_allclose(a[0, :3], [1, 0, 0], rtol=1e-7)  # self pair contributes nothing
    with np.errstate(invalid="ignore"):
        u = orc.brute_force(at, b, checked=False)
    assert np.isnan(u[0, 0])  # d * (m / (0 * 0)) = 0 * inf


def test_momentum_conservation_equal_mass(orc):
    s = workloads.direct_sum_case(512, seed=5)
    a = orc.brute_force(s.particles, s.particles).astype(np.float64)[:, :3]
    total = (a * s.particles[:, 3:4]).sum(0)
    scale = (np.abs(a) * s.particles[:, 3:4]).sum(0)
    assert np.all(np.abs(total) < 1e-5 * scale)


def test_large_softening_limit(orc):
    s = workloads.direct_sum_case(128, seed=9)
    eps = 1e4
    got = orc.brute_force(s.particles, s.particles, softening=eps).astype(np.float64)[:, :3]
    p = s.particles.astype(np.float64)
    exp = ((p[None, :, :3] - p[:, None, :3]) * p[None, :, 3:4]).sum(1) / eps**3
    np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-12 * np.abs(exp).max())


# ---- BruteForcePairs (sequential.rs:215-305) ----------------------------------------------------

def test_pairs_kat_reordered(orc):
    """The reference's acceleration_error! runs BruteForcePairs on the Reordered storage: pairs
    over the reordered particles (massive first) restricted to the massive ones, outputs mapped
    back to the original order (sequential.rs:368-385). Against the test's own FP64 expected
    value, at the reference's tolerance and at rounding level."""
    p = kat_particles()
    aff = p[:, 3] != 0
    order = np.concatenate([np.nonzero(aff)[0], np.nonzero(~aff)[0]])
    got_re = orc.brute_force_pairs(p[order], restrict_len=int(aff.sum()))
    got = np.empty_like(got_re)
    got[order] = got_re
    massive = p[aff].astype(np.float64)
    exp = np.zeros((6, 3))
    for i, p1 in enumerate(p.astype(np.float64)):
        for p2 in massive:
            d = p2[:3] - p1[:3]
            m2 = d @ d
            if m2 != 0:
                exp[i] += d * p2[3] * np.sqrt(1.0 / m2) / m2
    assert np.all(np.linalg.norm(1.0 - got[:, :3] / exp, axis=1) <= 1e-2)
    assert metrics.max_rel(got, exp) < 1e-6


@pytest.mark.parametrize("n,eps", [(2, 0.0), (9, 0.0), (64, 0.0), (37, 0.05)])
def test_pairs_vs_exhaustive_f64(orc, n, eps):
    """Unequal masses (a swapped mu_i / mu_j or a wrong reaction sign fails this)."""
    rng = np.random.default_rng(100 + n)
    p = rng.normal(size=(n, 4)).astype(np.float32)
    p[:, 3] = rng.uniform(0.1, 3.0, n)
    got = orc.brute_force_pairs(p, softening=eps)
    exp = numpy_direct_f64(p, p, eps)
    assert metrics.max_rel(got, exp) < 2e-5


def test_pairs_restricted_is_ordered_brute_force(orc):
    """RestrictedBruteForce over massive-first particles with restrict_len = massive count is
    Between(particles, massive): the pairs of two massless particles are never formed."""
    s = workloads.direct_sum_case(300, seed=21)
    p = s.particles.copy()
    p[200:, 3] = 0.0
    got = orc.brute_force_pairs(p, restrict_len=200)
    exp = numpy_direct_f64(p, p[:200])
    assert metrics.max_rel(got, exp) < 2e-5
    # coincident massless particles: the restricted form never pairs them, so no NaN
    q = np.concatenate([p, p[250:252]])
    with np.errstate(invalid="ignore", divide="ignore"):
        got = orc.brute_force_pairs(q, restrict_len=200)
    assert np.all(np.isfinite(got))


def test_pairs_checked_and_unchecked(orc):
    """The reference's pair form has no zero-distance check: a coincident pair is non-finite
    (1 / 0 then 0 * inf); checked drops the pair as the one-sided brute force does."""
    p = np.array([[1, 2, 3, 5], [1, 2, 3, 2], [2, 2, 3, 1]], np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = orc.brute_force_pairs(p)
    assert not np.all(np.isfinite(u[:2, :3]))
    c = orc.brute_force_pairs(p, checked=True)
    np.testing.assert_allclose(c, orc.brute_force(p, p, checked=True), rtol=1e-6)


def test_pairs_reaction_order(orc):
    """Particle k receives the reactions of i < k in order i = 0 .. k-1, then its own j > k fold
    (sequential.rs:249-258): +1e8, +1, -1e8 reactions sum to 0 only in that order."""
    # particle 3 at the origin; particles 0, 1, 2 at x = +1, +1, -1 with masses 1e8, 1, 1e8
    p = np.array([[1, 0, 0, 1e8], [1, 0, 0, 1], [-1, 0, 0, 1e8], [0, 0, 0, 1]], np.float32)
    got = orc.brute_force_pairs(p)
    assert got[3, 0] == 0.0
    got = orc.brute_force_pairs(p[[0, 2, 1, 3]])
    assert got[3, 0] == 1.0


# ---- octree ---------------------------------------------------------------------------------

def test_tree_hand_worked_two_bodies(orc):
    p = np.array([[0, 0, 0, 1], [1, 1, 1, 1]], np.float32)
    t = orc.Tree(p)
    nd = t.nodes()
    assert t.num_nodes == 3
    mn, mx = t.root_box()
    np.testing.assert_array_equal(mn, [0, 0, 0])
    np.testing.assert_array_equal(mx, [1, 1, 1])
    assert nd["internal"].tolist() == [1, 0, 0]
    # (1,1,1) is not below the centre on any axis -> orthant 0; (0,0,0) -> orthant 7
    assert nd["child"][0].tolist() == [1, -1, -1, -1, -1, -1, -1, 2]
    np.testing.assert_array_equal(nd["com"][0], [0.5, 0.5, 0.5])
    assert nd["mass"][0] == 2.0
    np.testing.assert_array_equal(nd["bmin"][1], [0.5, 0.5, 0.5])
    np.testing.assert_array_equal(nd["bmax"][2], [0.5, 0.5, 0.5])
    assert t.order().tolist() == [1, 0]


def test_tree_hand_worked_chain(orc):
    # (0,0,0) and (0.1,0,0) share orthant 7 at levels 1..3 and split at level 4, where the centre
    # is 0.0625: x = 0 -> bit 1 (orthant 7), x = 0.1 -> bit 0 (orthant 6).
    p = np.array([[0, 0, 0, 1], [0.1, 0, 0, 1], [1, 1, 1, 1]], np.float32)
    t = orc.Tree(p)
    nd = t.nodes()
    assert t.num_nodes == 7 and t.max_level == 4
    assert nd["level"].tolist() == [0, 1, 1, 2, 3, 4, 4]
    assert nd["internal"].tolist() == [1, 0, 1, 1, 1, 0, 0]
    assert nd["count"].tolist() == [3, 1, 2, 2, 2, 1, 1]
    widths = nd["bmax"][:, 0] - nd["bmin"][:, 0]
    np.testing.assert_array_equal(widths[[0, 2, 3, 4]], [1.0, 0.5, 0.25, 0.125])
    assert nd["child"][4].tolist() == [-1] * 6 + [5, 6]
    assert t.order().tolist() == [2, 1, 0]


def test_tree_invariants_plummer(orc):
    s = workloads.direct_sum_case(3000, seed=11)
    p = s.particles
    # a few exact duplicates exercise the equal-coordinates leaf rule (tree/mod.rs:112-114)
    p = np.concatenate([p, p[:5], p[:2]])
    t = orc.Tree(p)
    nd = t.nodes()
    order = t.order()
    assert sorted(order.tolist()) == list(range(len(p)))
    tp = p[order].astype(np.float64)
    leaves = 0
    for k, n in enumerate(nd):
        b, c = n["begin"], n["count"]
        mem = tp[b:b + c]
        distinct = np.unique(mem[:, :3], axis=0).shape[0]
        assert bool(n["internal"]) == (distinct > 1)
        if not n["internal"]:
            leaves += c
            assert all(ch == -1 for ch in n["child"])
        else:
            kids = [ch for ch in n["child"] if ch >= 0]
            assert sum(nd["count"][ch] for ch in kids) == c
            assert all(nd["level"][ch] == n["level"] + 1 for ch in kids)
            assert kids == sorted(kids) and kids[0] == k + 1  # depth-first pre-order
        if n["level"] > 0:
            assert np.all(mem[:, :3] >= n["bmin"] - 1e-6) and np.all(mem[:, :3] <= n["bmax"] + 1e-6)
        np.testing.assert_allclose(n["mass64"], mem[:, 3].sum(), rtol=1e-12)
        np.testing.assert_allclose(n["com64"], (mem[:, :3] * mem[:, 3:4]).sum(0) / mem[:, 3].sum(),
                                   rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(n["com"], n["com64"], rtol=1e-4, atol=1e-5)
    assert leaves == len(p)


def test_bh_theta_zero_equals_brute_force(orc):
    s = workloads.direct_sum_case(1024, seed=13)
    bf = orc.brute_force(s.particles, s.particles)
    bh = orc.barnes_hut(s.particles, s.particles, 0.0)
    assert metrics.rms_rel(bh, bf) < 1e-6


def test_bh_single_body_is_exact(orc):
    b = np.array([[0.3, -1.0, 2.0, 7.0]], np.float32)
    a = workloads.direct_sum_case(64, seed=2).particles
    for theta in (0.0, 0.5, 10.0):
        np.testing.assert_array_equal(orc.barnes_hut(a, b, theta), orc.brute_force(a, b))


def test_bh_huge_theta_is_root_monopole(orc):
    s = workloads.direct_sum_case(500, seed=17)
    far = np.array([[40, 0, 0, 0], [0, -55, 3, 0], [10, 10, 10, 0]], np.float32)
    got = orc.barnes_hut(far, s.particles, 1e30)
    p = s.particles.astype(np.float64)
    com = (p[:, :3] * p[:, 3:4]).sum(0) / p[:, 3].sum()
    d = com[None] - far[:, :3].astype(np.float64)
    exp = d * p[:, 3].sum() / (np.linalg.norm(d, axis=1) ** 3)[:, None]
    assert metrics.max_rel(got, exp) < 1e-4


def test_bh_self_skip_on_zero_distance(orc):
    # an affected particle sitting exactly on a node's centre of mass skips that node entirely
    b = np.array([[1, 0, 0, 1], [-1, 0, 0, 1]], np.float32)
    at = np.array([[0, 0, 0, 0]], np.float32)
    np.testing.assert_array_equal(orc.barnes_hut(at, b, 1e30)[0], [0, 0, 0, 0])
    np.testing.assert_allclose(orc.barnes_hut(at, b, 0.0)[0], [0, 0, 0, 0], atol=0)


def test_bh_error_vs_f64_is_monopole_sized(orc):
    s = workloads.direct_sum_case(2048, seed=19)
    ref = orc.direct_f64(s.particles, s.particles)
    e3 = metrics.rms_rel(orc.barnes_hut(s.particles, s.particles, 0.3), ref)
    e8 = metrics.rms_rel(orc.barnes_hut(s.particles, s.particles, 0.8), ref)
    assert 1e-5 < e3 < 5e-3 and e3 < e8 < 5e-2




# ---- the relative opening criterion (the GPU backend's PNB_BH_RELATIVE; no reference
# counterpart): open an internal node also when M w^p / r^(p+2) > est = alpha |a~| ----


@pytest.mark.parametrize("order", [2, 3])
def test_relative_criterion_limits(orc, order):
    """est = +inf never opens by the relative rule (the geometric traversal, bit for bit); est = 0
    opens every massive internal node (theta = 0: every leaf accepted)."""
    s = workloads.direct_sum_case(1500, seed=31).particles
    t = orc.Tree(s)
    geo, vg = t.barnes_hut(s, 0.6, com64=True, return_visits=True)
    rel, vr = t.barnes_hut_rel(s, 0.6, np.full(len(s), np.inf, np.float32), order=order,
                               return_visits=True)
    np.testing.assert_array_equal(rel, geo)
    np.testing.assert_array_equal(vr, vg)
    all_, va = t.barnes_hut(s, 0.0, com64=True, return_visits=True)
    rel0, v0 = t.barnes_hut_rel(s, 0.6, np.zeros(len(s), np.float32), order=order,
                                return_visits=True)
    np.testing.assert_array_equal(rel0, all_)
    np.testing.assert_array_equal(v0, va)


@pytest.mark.parametrize("order", [2, 3])
def test_relative_criterion_threshold_closed_form(orc, order):
    """A compact cluster seen from outside along a ray: the root is accepted exactly beyond the
    distance r* where M w^p / r*^(p+2) = alpha |a(r*)| (|a| from an FP64 direct sum, r* by
    bisection), and opened inside it. A wrong power of w or r, or the wrong mass, moves r*."""
    rng = np.random.default_rng(5)
    cl = np.zeros((40, 4), np.float32)
    cl[:, :3] = rng.uniform(-0.05, 0.05, (40, 3)).astype(np.float32)
    cl[:, 3] = np.float32(1.0 / 40)
    t = orc.Tree(cl)
    root = t.nodes()[0]
    M = float(np.float32(root["mass64"]))
    w = float(root["bmax"][0] - root["bmin"][0])
    c = root["com64"].astype(np.float32).astype(np.float64)
    u = np.array([0.8, 0.36, 0.48])
    alpha = 1e-3

    def at(r):
        return (c + r * u).astype(np.float32)

    def amag(r):
        p = np.zeros((1, 4), np.float32)
        p[0, :3] = at(r)
        return float(np.linalg.norm(numpy_direct_f64(p, cl)[0]))

    def g(r):
        x = at(r).astype(np.float64)
        rr = float(np.linalg.norm(x - c))
        return M * w ** order / rr ** (order + 2) - alpha * amag(r)

    lo, hi = 2 * w, 100 * w  # g(lo) > 0 > g(hi)
    assert g(lo) > 0 > g(hi)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if g(mid) > 0 else (lo, mid)
    rstar = 0.5 * (lo + hi)
    for f, opened in ((0.98, True), (1.02, False)):
        r = f * rstar
        p = np.zeros((1, 4), np.float32)
        p[0, :3] = at(r)
        est = np.array([alpha * amag(r)], np.float32)
        _, v = t.barnes_hut_rel(p, 1e6, est, order=order, return_visits=True)
        assert (v[0] > 1) == opened, (f, v[0], rstar / w)


@pytest.mark.parametrize("order", [2, 3])
def test_relative_buckets(orc, order):
    """Opened nodes of at most `bucket` particles summed particle by particle: with est = 0 (every
    massive internal node opened) the terms are exactly the theta = 0 traversal's leaves (distinct
    positions: one particle per leaf), summed in another order; a bucket holding the whole set
    opens the root into the direct sum; bucket = 0 is the plain relative traversal."""
    s = workloads.direct_sum_case(1500, seed=35).particles
    t = orc.Tree(s)
    all_, va = t.barnes_hut(s, 0.0, com64=True, accum64=True, return_visits=True)
    for bucket in (1, 4, 16, 31):
        b, vb = t.barnes_hut_rel(s, 0.6, np.zeros(len(s), np.float32), order=order,
                                 bucket=bucket, accum64=True, return_visits=True)
        np.testing.assert_array_equal(vb, va)
        np.testing.assert_allclose(b, all_, rtol=1e-6, atol=1e-12)
    small = s[:20]
    ts = orc.Tree(small)
    got, v = ts.barnes_hut_rel(small, 0.6, np.zeros(20, np.float32), order=order, bucket=20,
                               accum64=True, return_visits=True)
    assert (v == 19).all()  # the root is one bucket: every other particle once
    np.testing.assert_allclose(got[:, :3], numpy_direct_f64(small, small), rtol=1e-6)
    est = np.full(len(s), 1e-4, np.float32)
    np.testing.assert_array_equal(t.barnes_hut_rel(s, 0.6, est, order=order),
                                  t.barnes_hut_rel(s, 0.6, est, order=order, bucket=0))


@pytest.mark.parametrize("order", [2, 3])
def test_relative_criterion_error_follows_alpha(orc, order):
    """Against the FP64 direct sum the error falls with alpha. For order 2 every accepted node's
    neglected quadrupole is bounded by alpha |a|, so the rms error stays within a few alpha; for
    order 3 the plain monopole walk keeps its unbounded quadrupole error (the learned correction
    supplies that term), so only the monotone decrease is asserted."""
    s = workloads.direct_sum_case(2048, seed=33).particles
    t = orc.Tree(s)
    ref = orc.direct_f64(s, s)
    amag = np.linalg.norm(ref, axis=1).astype(np.float32)
    errs = []
    for alpha in (1e-2, 1e-3, 1e-4):
        got = t.barnes_hut_rel(s, 1e6, np.float32(alpha) * amag, order=order, com64=True,
                               accum64=True)
        errs.append(metrics.rms_rel(got, ref))
        if order == 2:
            assert errs[-1] < 5 * alpha, (alpha, errs[-1])
    assert errs[0] > errs[1] > errs[2], errs


@pytest.mark.parametrize("theta,tol", [(None, 1e-2), (0.0, 1e-2), (0.5, 1e-1)])
def test_circular_orbit_oracle(orc, theta, tol):
    def accel(p):
        massive = p[p[:, 3] != 0]
        if theta is None:
            return orc.brute_force(p, massive, threads=1)
        return orc.barnes_hut(p, massive, theta, threads=1)

    ed, ee = circular_orbit(accel, 20)
    assert ed < tol and ee < tol
