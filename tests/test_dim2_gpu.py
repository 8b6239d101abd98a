"""Two-dimensional particles (glam::Vec2 / f32, particular/src/gravity/impls/mod.rs:381-418) on
the GPU: records {x, y, mu}, accelerations {ax, ay}, through the 2D C ABI entry points."""
import numpy as np
import pytest

import particular_amd as pa

pytestmark = pytest.mark.gpu


def disc(n, seed, massless=0):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0, 1, n)) * 2.0
    a = rng.uniform(0, 2 * np.pi, n)
    p = np.zeros((n, 3), np.float32)
    p[:, 0] = r * np.cos(a)
    p[:, 1] = r * np.sin(a)
    p[:, 2] = 1.0 / n
    if massless:
        p[-massless:, 2] = 0.0
    return p


def direct_2d_f64(aff, afg, eps=0.0):
    d = afg[None, :, :2].astype(np.float64) - aff[:, None, :2].astype(np.float64)
    r2 = (d * d).sum(-1)
    rs = r2 + eps * eps
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(r2 > 0, afg[None, :, 2] / (rs * np.sqrt(rs)), 0.0)
    return (d * s[..., None]).sum(1)


def rel_rms(a, ref):
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.mean(e * e)))


def as3d(p2):
    p = np.zeros((p2.shape[0], 4), np.float32)
    p[:, :2] = p2[:, :2]
    p[:, 3] = p2[:, 2]
    return p


def test_brute_force_2d_matches_3d_path_and_fp64(res):
    p = disc(3000, 1)
    for inter, eps in ((pa.Acceleration.checked(), 0.0), (pa.AccelerationSoftened.checked(1e-2), 1e-2)):
        a2 = pa.BruteForce(res, inter).compute(p)
        a3 = pa.BruteForce(res, inter).compute(as3d(p))
        assert a2.shape == (3000, 2)
        np.testing.assert_array_equal(a2, a3[:, :2])
        assert not a3[:, 2].any()
        assert rel_rms(a2, direct_2d_f64(p, p, eps)) < 1e-6


def test_barnes_hut_2d_matches_3d_path_and_oracle(res, orc):
    """The quadtree's decisions: the oracle's traversal (same decisions, FP64 sums) on the z = 0
    embedding; theta = 0 is the 2D direct sum."""
    p = disc(4000, 2)
    errs = []
    for theta in (0.0, 0.5, 0.9):
        b2 = pa.BarnesHut(res, theta, pa.Acceleration.checked()).compute(p)
        b3 = pa.BarnesHut(res, theta, pa.Acceleration.checked()).compute(as3d(p))
        np.testing.assert_array_equal(b2, b3[:, :2])
        tree = orc.Tree(as3d(p))
        ref = tree.barnes_hut(as3d(p), theta, com64=True, accum64=True)[:, :2]
        assert rel_rms(b2, ref) < 5e-7
        errs.append(rel_rms(b2, direct_2d_f64(p, p)))
        print("2D Barnes-Hut theta", theta, "rms vs FP64", errs[-1])
    assert errs[0] < 1e-6 and errs[0] < errs[1] < errs[2]


def test_barnes_hut_2d_relative_criterion(res):
    """The relative opening criterion through the 2D entry points: the same bits as the 3D path
    on the z = 0 embedding (plain and with per-particle estimates), and a smaller error against the
    2D direct sum as alpha falls."""
    p = disc(4000, 3)
    exact = direct_2d_f64(p, p)
    errs = []
    for alpha in (1e-2, 1e-3, 1e-4):
        for pp in (False, True):
            alg = pa.BarnesHut(res, 1.0, pa.Acceleration.checked(), relative=alpha,
                               estimate_per_particle=pp)
            b2 = alg.compute(p)
            np.testing.assert_array_equal(b2, alg.compute(as3d(p))[:, :2])
        errs.append(rel_rms(b2, exact))
    print("2D relative criterion rms vs FP64", errs)
    assert errs[0] > errs[1] > errs[2]


def test_reordered_2d_split(res):
    p = disc(2000, 3, massless=1500)
    bf = pa.BruteForce(res, pa.AccelerationSoftened.checked(1e-3))
    r = bf.compute(pa.Reordered.new(p))
    b = bf.compute(pa.Between(p, p[p[:, 2] != 0]))
    np.testing.assert_array_equal(r, b)
    assert pa.Reordered.new(p).affecting().shape[0] == 500


def test_device_tensors_2d(res):
    import torch

    p = disc(2048, 4)
    dev = torch.from_numpy(p).cuda()
    a = pa.BruteForce(res, pa.Acceleration.checked()).compute(dev)
    h = pa.BruteForce(res, pa.Acceleration.checked()).compute(p)
    np.testing.assert_array_equal(a.cpu().numpy(), h)
    bh = pa.BarnesHut(res, 0.5, pa.Acceleration.checked()).compute(pa.Between(dev, dev))
    np.testing.assert_array_equal(bh.cpu().numpy(),
                                  pa.BarnesHut(res, 0.5, pa.Acceleration.checked()).compute(p))


def test_empty_and_massless_2d(res):
    p = disc(100, 5)
    z = np.zeros((0, 3), np.float32)
    assert pa.BruteForce(res, pa.Acceleration.checked()).compute(pa.Between(z, p)).shape == (0, 2)
    out = pa.BruteForce(res, pa.Acceleration.checked()).compute(pa.Between(p, z))
    assert out.shape == (100, 2) and not out.any()
    out = pa.BarnesHut(res, 0.5, pa.Acceleration.checked()).compute(pa.Between(p, z))
    assert not out.any()
