"""The allocation-free brute force (pnb_bf_accel_f32_dev_noalloc) against pnb_bf_accel_f32_dev on
the same device inputs: the same kernels and shapes, so bit-identical results. Two calls in a row,
the second smaller, so that what the first left in the shared static scratch cannot reach the
second; a shape whose partial sums exceed the scratch is refused with an argument error."""
import numpy as np
import pytest

from particular_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _call(res, fn, a, b, eps, checked):
    import torch

    out = torch.full((a.shape[0], 4), float("nan"), dtype=torch.float32, device="cuda")
    rc = fn(res.handle, a.data_ptr(), a.shape[0], b.data_ptr() if b is not None else None,
            0 if b is None else b.shape[0], eps, checked, out.data_ptr())
    res.synchronize()
    return rc, out.cpu().numpy()


def _inputs(n, m, seed):
    import torch

    s = workloads.direct_sum_case(max(n, m), seed=seed)
    a = torch.from_numpy(np.ascontiguousarray(s.particles[:n])).cuda()
    b = torch.from_numpy(np.ascontiguousarray(s.particles[-m:])).cuda()
    torch.cuda.synchronize()
    return a, b


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_equals_dev_entry_point(res, order):
    import torch

    L = _lib.lib()
    res.use_torch_stream()
    res.clear_bf_tuning()
    cases = [(300_000, 65_536, 1e-3, 1), (5_000, 3_000, 0.0, 1)]  # 32 and 2 affecting ranges
    if order == "small_then_large":
        cases = cases[::-1]
    for n, m, eps, chk in cases:
        a, b = _inputs(n, m, seed=n + m)
        rc_ref, ref = _call(res, L.pnb_bf_accel_f32_dev, a, b, eps, chk)
        rc, got = _call(res, L.pnb_bf_accel_f32_dev_noalloc, a, b, eps, chk)
        assert rc_ref == _lib.PNB_OK and rc == _lib.PNB_OK, res.last_error()
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, ref)
        print(f"n={n} m={m}: bit-identical to pnb_bf_accel_f32_dev", flush=True)
    # no affecting particles: zeros; no affected: nothing to do
    a, _ = _inputs(100, 1, seed=7)
    rc, got = _call(res, L.pnb_bf_accel_f32_dev_noalloc, a, None, 0.0, 1)
    assert rc == _lib.PNB_OK and np.all(got == 0.0)
    assert L.pnb_bf_accel_f32_dev_noalloc(res.handle, None, 0, None, 0, 0.0, 1, None) == _lib.PNB_OK
    torch.cuda.synchronize()


def test_noalloc_refuses_oversized_split(res):
    L = _lib.lib()
    res.use_torch_stream()
    a, b = _inputs(300_000, 65_536, seed=11)
    try:
        res.set_bf_tuning(0, 1, 4, 64)  # 64 ranges x 300,000 bodies > 2^24 partial sums
        rc, _ = _call(res, L.pnb_bf_accel_f32_dev_noalloc, a, b, 0.0, 1)
        assert rc == _lib.PNB_E_ARG
        assert "allocation-free" in res.last_error()
        rc, ref = _call(res, L.pnb_bf_accel_f32_dev, a, b, 0.0, 1)  # the allocating form runs it
        assert rc == _lib.PNB_OK and np.all(np.isfinite(ref))
    finally:
        res.clear_bf_tuning()
