"""Checks shared by the Barnes-Hut GPU tests: octree equality with the oracle and accuracy of the
walk against the oracle's traversals (test infrastructure)."""
import numpy as np

import particular_amd as pa
from particular_amd import metrics


def assert_same_tree(res, orc, particles, tree=None):
    """The last octree built on `res` equals the oracle's octree of `particles` node for node:
    count, particle order, level, particle range, leaf/internal, bit-identical f32 widths, FP64
    masses, centres of mass and skip pointers (pre-order subtree ends)."""
    t = tree if tree is not None else orc.Tree(particles)
    on = t.nodes()
    gn, gorder = pa.tree_export(res, particles.shape[0])
    assert gn.shape[0] == on.shape[0], (gn.shape[0], on.shape[0])
    np.testing.assert_array_equal(gorder, t.order())
    np.testing.assert_array_equal(gn["level"], on["level"])
    np.testing.assert_array_equal(gn["begin"], on["begin"])
    np.testing.assert_array_equal(gn["count"], on["count"])
    np.testing.assert_array_equal(gn["internal"], on["internal"])
    ow = on["bmax"][:, 0] - on["bmin"][:, 0]
    internal = on["internal"] == 1
    np.testing.assert_array_equal(gn["width"][internal], ow[internal])
    np.testing.assert_allclose(gn["mass64"], on["mass64"], rtol=1e-12)
    np.testing.assert_allclose(gn["com"], on["com64"].astype(np.float32), rtol=3e-7, atol=1e-30)
    np.testing.assert_array_equal(gn["skip"], t.skips())
    return t


def assert_reference_com(res, particles, tree):
    """A build with reference-exact centres of mass (PNB_BH_REF_COM) gives every node the f32
    mass and centre of mass of the reference's folds (impls/mod.rs:120-134 over each node's slice
    in the reference's order, tree/mod.rs:108-125), as the oracle restates them, bit for bit."""
    bh(res, pa.Between(particles[:1], particles), 0.5, reference_com=True)
    gn, _ = pa.tree_export(res, particles.shape[0])
    on = tree.nodes()
    assert gn.shape[0] == on.shape[0]
    np.testing.assert_array_equal(gn["mass"], on["mass"])
    np.testing.assert_array_equal(gn["com"], on["com"])


def bh(res, storage, theta, interaction=None, **kw):
    return pa.BarnesHut(res, theta, interaction or pa.Acceleration.checked(), **kw).compute(storage)


def check_against_oracle(res, tree, particles, theta, idx=None, label=""):
    """Three comparisons on the same inputs and theta; returns the measured distances.
    * the default walk (FP64 centres of mass) against the oracle with the same node data whose
      accepted terms are summed in FP64 -- isolates the traversal and summation: rms <= 5e-7;
    * the walk with reference-exact f32 centres of mass (PNB_BH_REF_COM) against the
      reference-exact restatement of parallel::BarnesHut -- the parity bar: rms <= 1e-5;
    * the default walk against the reference: at most the parity bar or, where the reference's
      own f32 centre-of-mass rounding (reference vs exact-COM oracle) exceeds it, that rounding
      plus 2e-6."""
    sel = slice(None) if idx is None else idx
    aff = particles[sel]
    got = bh(res, particles, theta)[sel]
    got_ref = bh(res, particles, theta, reference_com=True)[sel]
    exact = tree.barnes_hut(aff, theta, com64=True, accum64=True)
    ref = tree.barnes_hut(aff, theta, com64=False)
    e_alg = metrics.summary(got, exact)
    e_par = metrics.summary(got_ref, ref)
    e_dev = metrics.summary(got, ref)
    e_refcom = metrics.rms_rel(ref, exact)
    print(f"{label} theta {theta}: default vs exact-COM oracle {e_alg}; reference-COM mode vs "
          f"reference {e_par}; default vs reference {e_dev}; reference's own COM rounding "
          f"{e_refcom:.3e}", flush=True)
    assert e_alg["rms_rel"] < 5e-7 and e_alg["max_rel"] < 1e-4
    assert e_par["rms_rel"] < 1e-5
    assert e_dev["rms_rel"] < max(1e-5, e_refcom + 2e-6)
    return {"theta": theta, "default_vs_exact": e_alg, "refcom_vs_reference": e_par,
            "default_vs_reference": e_dev, "reference_com_rounding": e_refcom}
