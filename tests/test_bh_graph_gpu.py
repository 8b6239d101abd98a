"""Capturable Barnes-Hut (pnb_bh_set_capturable): once sized by an ordinary call, the build and
walk make no host round trip, so they can be captured in a hipGraph. The tree, its node data and
the accelerations (plain and with the learned correction) are those of the ordinary build, bit
for bit; graph replays with new positions equal eager calls on the same positions; a tree that
outgrows the sizes gives NaN and an error from pnb_bh_check (never a hang)."""
import contextlib
import gc

import numpy as np
import pytest

import particular_amd as pa
from particular_amd import _lib, metrics, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture
def cres():
    """A context of its own: the capturable mode and its sizes belong to one context."""
    r = pa.HipResources(0)
    yield r
    r.close()


@contextlib.contextmanager
def capture(g):
    """torch.cuda.graph(g) with no garbage collection during the capture: a dead reference cycle
    collected then (an earlier test's graph and its memory pool, held by a traceback) would free
    device memory while the stream is capturing, which aborts the process."""
    import torch

    gc.collect()
    gc.disable()
    try:
        with torch.cuda.graph(g):
            yield
    finally:
        gc.enable()


def _bh(res, dev, theta=0.5):
    return pa.BarnesHut(res, theta, pa.Acceleration.checked()).compute(dev)


@pytest.mark.parametrize("n", [64, 4096, 100_000])
def test_capturable_equals_ordinary(cres, orc, n):
    import torch

    s = workloads.direct_sum_case(n, seed=500 + n)
    dev = torch.from_numpy(s.particles).cuda()
    ref = _bh(cres, dev).cpu().numpy()
    nodes_ref, order_ref = pa.tree_export(cres, n)
    cres.set_bh_capturable(True)
    sizing = _bh(cres, dev).cpu().numpy()  # the ordinary build that sizes the buffers
    got = _bh(cres, dev).cpu().numpy()     # no host round trip
    cres.bh_check()
    np.testing.assert_array_equal(sizing, ref)
    np.testing.assert_array_equal(got, ref)
    nodes, order = pa.tree_export(cres, n)
    np.testing.assert_array_equal(order, order_ref)
    for f in nodes_ref.dtype.names:
        np.testing.assert_array_equal(nodes[f], nodes_ref[f], err_msg=f)
    exact = orc.Tree(s.particles).barnes_hut(s.particles, 0.5, com64=True, accum64=True)
    assert metrics.rms_rel(got[:, :3], exact) < 5e-7


def test_between_distinct_sets(cres):
    import torch

    s = workloads.direct_sum_case(6000, seed=61)
    aff = torch.from_numpy(s.particles[:2500].copy()).cuda()
    afg = torch.from_numpy(s.particles[1000:].copy()).cuda()
    alg = pa.BarnesHut(cres, 0.6, pa.Acceleration.checked())
    ref = alg.compute(pa.Between(aff, afg)).cpu().numpy()
    cres.set_bh_capturable(True)
    alg.compute(pa.Between(aff, afg))
    got = alg.compute(pa.Between(aff, afg)).cpu().numpy()
    cres.bh_check()
    np.testing.assert_array_equal(got, ref)


def test_graph_replays_with_new_positions(cres):
    """A hipGraph of the whole evaluation (build + walk), replayed after the positions change in
    place; each replay is checked against an eager call on the same positions, with the
    program's own eager work and synchronisations between replays."""
    import torch

    n = 20_000
    s = workloads.direct_sum_case(n, seed=77)
    dev = torch.from_numpy(s.particles.copy()).cuda()
    both = pa.Between(dev, dev)  # one device pointer twice: pnb_bh_accel_f32_dev
    cres.set_bh_capturable(True)
    alg = pa.BarnesHut(cres, 0.5, pa.Acceleration.checked())
    alg.compute(both)  # sizes
    cres.bh_check()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with capture(g):
        static_out = alg.compute(both)
    cres.reset_stream()
    rng = np.random.default_rng(1)
    for step in range(4):
        p = s.particles.copy()
        p[:, :3] += rng.normal(scale=2e-3, size=(n, 3)).astype(np.float32)
        dev.copy_(torch.from_numpy(p))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = static_out.cpu().numpy()
        cres.bh_check()
        eager = alg.compute(both).cpu().numpy()
        cres.bh_check()
        np.testing.assert_array_equal(got, eager)
        assert np.all(np.isfinite(got))
        print(f"replay {step}: bit-identical to the eager call", flush=True)


def test_outgrown_sizes_give_nan_and_an_error(cres, orc):
    import torch

    n = 2000
    spread = workloads.direct_sum_case(n, seed=3).particles
    dev = torch.from_numpy(spread.copy()).cuda()
    cres.set_bh_capturable(True)
    alg = pa.BarnesHut(cres, 0.5, pa.Acceleration.checked())
    alg.compute(dev)
    cres.bh_check()
    deep = spread.copy()
    deep[1::2, :3] = deep[0::2, :3] + np.float32(1e-6)  # 1,000 close pairs: deep chains
    dev.copy_(torch.from_numpy(deep))
    out = alg.compute(dev).cpu().numpy()
    with pytest.raises(pa.PnbError) as e:
        cres.bh_check()
    assert e.value.status == _lib.PNB_E_STATE
    assert np.all(np.isnan(out[:, :3]))
    # sized anew by an ordinary call, the deep tree then runs without round trips
    cres.set_bh_capturable(True)
    alg.compute(dev)
    got = alg.compute(dev).cpu().numpy()
    cres.bh_check()
    exact = orc.Tree(deep).barnes_hut(deep, 0.5, com64=True, accum64=True)
    assert metrics.rms_rel(got[:, :3], exact) < 5e-7


def test_refused_flags(cres):
    import torch

    s = workloads.direct_sum_case(500, seed=9)
    dev = torch.from_numpy(s.particles).cuda()
    cres.set_bh_capturable(True)
    with pytest.raises(pa.PnbError) as e:
        pa.BarnesHut(cres, 0.5, pa.Acceleration.checked(), reference_com=True).compute(dev)
    assert e.value.status == _lib.PNB_E_ARG
    with pytest.raises(pa.PnbError) as e:
        pa.BarnesHut(cres, 1.0, pa.Acceleration.checked(), relative=1e-4).compute(dev)
    assert e.value.status == _lib.PNB_E_ARG
    cres.set_bh_capturable(False)
    _bh(cres, dev)  # the ordinary mode again


def test_reordered(cres):
    """Reordered (the mu != 0 split done on the device) in the capturable mode: eager calls equal
    the ordinary mode's; inside a capture it is refused before any work is recorded (the split
    reads its count on the host), and the capture itself then ends normally."""
    import torch

    s = workloads.direct_sum_case(3000, seed=10)
    p = s.particles.copy()
    p[::3, 3] = 0.0  # a third massless: affecting set smaller than the affected one
    dev = torch.from_numpy(p).cuda()
    alg = pa.BarnesHut(cres, 0.5, pa.Acceleration.checked())
    ref = alg.compute(pa.Reordered(dev)).cpu().numpy()
    cres.set_bh_capturable(True)
    alg.compute(pa.Reordered(dev))  # sizes
    got = alg.compute(pa.Reordered(dev)).cpu().numpy()
    cres.bh_check()
    np.testing.assert_array_equal(got, ref)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with capture(g):
        with pytest.raises(pa.PnbError) as e:
            alg.compute(pa.Reordered(dev))
    cres.reset_stream()
    assert e.value.status == _lib.PNB_E_STATE


@pytest.mark.parametrize("network", ["quadrupole", "octupole"])
def test_corrected_walk(cres, network):
    """The learned correction in the capturable mode: the node moments summed on the sizing
    build's level plan, the network over the internal nodes counted on the device, the corrected
    walk; eager calls equal the ordinary mode's bit for bit, graph replays with new positions
    equal eager calls, and a call whose flags differ from the sizing call's is refused inside a
    capture."""
    import torch

    from particular_amd import learned

    learned.init(cres, 0)
    learned.set_params(cres, learned.quadrupole_params() if network == "quadrupole"
                       else learned.octupole_params())
    n = 20_000
    s = workloads.two_plummer_case(n, seed=81)
    dev = torch.from_numpy(s.particles.copy()).cuda()
    both = pa.Between(dev, dev)
    alg = pa.BarnesHut(cres, 0.5, pa.Acceleration.checked(), mlp=True)
    ref = alg.compute(both).cpu().numpy()
    plain = _bh(cres, dev).cpu().numpy()
    assert np.abs(ref - plain).max() > 0  # the correction is applied
    cres.set_bh_capturable(True)
    np.testing.assert_array_equal(alg.compute(both).cpu().numpy(), ref)  # sizes
    np.testing.assert_array_equal(alg.compute(both).cpu().numpy(), ref)  # no round trip
    cres.bh_check()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with capture(g):
        static_out = alg.compute(both)
    cres.reset_stream()
    rng = np.random.default_rng(2)
    for step in range(3):
        p = s.particles.copy()
        p[:, :3] += rng.normal(scale=2e-3, size=(n, 3)).astype(np.float32)
        dev.copy_(torch.from_numpy(p))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = static_out.cpu().numpy()
        cres.bh_check()
        eager = alg.compute(both).cpu().numpy()
        cres.bh_check()
        np.testing.assert_array_equal(got, eager)
        assert np.all(np.isfinite(got))
        print(f"replay {step}: bit-identical to the eager call", flush=True)
    # the plain walk was not sized with these buffers: refused inside a capture
    g2 = torch.cuda.CUDAGraph()
    with capture(g2):
        with pytest.raises(pa.PnbError) as e:
            _bh(cres, dev)
    cres.reset_stream()
    assert e.value.status == _lib.PNB_E_STATE
