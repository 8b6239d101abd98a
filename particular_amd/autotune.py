"""Autotuner for the brute-force and batched-ensemble kernel shapes.

For every size bucket it times each candidate shape with the library's own event profiling (device
time of the force kernel, median of a few launches) and keeps the fastest:
  * brute force: (variant, affected per lane, affecting splits) per affected-count bucket, at a
    fixed affecting count;
  * batched ensemble: affected per lane per system-size bucket, on uniform ensembles.
The table is written as JSON and applied with :func:`apply`. Run on the target GPU:
``python -m particular_amd.autotune --out particular_amd/tuning/mi355x.json``.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys

import numpy as np

from . import _lib, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "mi355x.json")

BF_BUCKETS = [1024, 4096, 16384, 65536, 262144, 1048576]
BF_M = 65536
BATCH_BUCKETS = [64, 256, 1024, 4096, 16384, 65536]


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _time_bf(res, dev_a, n, dev_b, m, reps):
    import torch

    out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    L = _lib.lib()
    res.use_torch_stream()
    res.check(L.pnb_bf_accel_f32_dev(res.handle, dev_a, n, dev_b, m, 1e-3, 1, out.data_ptr()))
    times = []
    for _ in range(reps):
        res.profile(True)
        res.check(L.pnb_bf_accel_f32_dev(res.handle, dev_a, n, dev_b, m, 1e-3, 1, out.data_ptr()))
        ms, cnt = res.profile_read(_lib.PROF_BF)
        times.append(ms / max(cnt, 1))
    res.profile(False)
    return statistics.median(times)


def tune_bf(res, reps=3):
    import torch

    cus = 256
    table = []
    src = workloads.massive_massless_case(n_massive=BF_M, n_total=max(BF_BUCKETS) + BF_M)
    massive = torch.from_numpy(src.particles[src.particles[:, 3] != 0].copy()).cuda()
    allp = torch.from_numpy(src.particles.copy()).cuda()
    for n in BF_BUCKETS:
        best = None
        for variant in (0, 1):
            for per_lane in (1, 2, 4, 8):
                blocks = -(-n // (256 * per_lane))
                for j_split in (1, 2, 4, 8, 16, 32, 64):
                    if j_split > 1 and blocks * j_split > 64 * cus:
                        continue
                    if BF_M // j_split < 1024:
                        continue
                    res.set_bf_tuning(0, variant, per_lane, j_split)
                    t = _time_bf(res, allp.data_ptr(), n, massive.data_ptr(), BF_M, reps)
                    if best is None or t < best[0]:
                        best = (t, variant, per_lane, j_split)
        res.clear_bf_tuning()
        t, variant, per_lane, j_split = best
        rate = n * BF_M / (t * 1e-3) / 1e9
        _log(f"bf n={n}: variant={variant} per_lane={per_lane} j_split={j_split} "
             f"{t:.3f} ms ({rate:.0f} G interactions/s)")
        table.append({"n_min": n, "variant": variant, "per_lane": per_lane, "j_split": j_split,
                      "ms": t, "g_interactions_per_s": rate})
    return table


def tune_batch(res, total=2_000_000, reps=3):
    L = _lib.lib()
    table = []
    for size in BATCH_BUCKETS:
        num = max(1, total // size)
        offsets = np.arange(num + 1, dtype=np.int64) * size
        pos, vel = workloads.plummer(num * size, seed=size)
        mu = np.full(num * size, 1.0 / size, np.float32)
        best = None
        for per_lane in (1, 2, 4, 8):
            res.reset_stream()
            res.check(L.pnb_batch_set_tuning(res.handle, 0, per_lane))
            res.check(L.pnb_batch_init(res.handle, pos.ctypes.data, vel.ctypes.data,
                                       mu.ctypes.data, offsets.ctypes.data, num, 1e-3, 1))
            times = []
            for _ in range(reps):
                res.profile(True)
                res.check(L.pnb_batch_run(res.handle, 1, 1e-3, 0))
                ms, cnt = res.profile_read(_lib.PROF_BATCH)
                times.append(ms / max(cnt, 1))
            res.profile(False)
            t = statistics.median(times)
            if best is None or t < best[0]:
                best = (t, per_lane)
        t, per_lane = best
        rate = num * size * size / (t * 1e-3) / 1e9
        _log(f"batch size={size}: per_lane={per_lane} {t:.3f} ms ({rate:.0f} G interactions/s)")
        table.append({"n_min": size, "per_lane": per_lane, "ms": t, "g_interactions_per_s": rate})
    return table


def apply(res, path: str = DEFAULT_TABLE) -> bool:
    """Installs a tuning table on a HipResources context; False when the file does not exist."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        t = json.load(f)
    res.clear_bf_tuning()
    for e in t.get("brute_force", []):
        # a bucket's shape applies from the geometric midpoint below it
        res.set_bf_tuning(int(e["n_min"] * 0.7) if e["n_min"] > BF_BUCKETS[0] else 0,
                          e["variant"], e["per_lane"], e["j_split"])
    L = _lib.lib()
    for e in t.get("batch", []):
        n_min = int(e["n_min"] * 0.7) if e["n_min"] > BATCH_BUCKETS[0] else 0
        res.check(L.pnb_batch_set_tuning(res.handle, n_min, e["per_lane"]))
    return True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--skip-batch", action="store_true")
    a = ap.parse_args()
    import particular_amd as pa

    res = pa.HipResources(0)
    table = {"device": "MI355X (gfx950)", "brute_force": tune_bf(res)}
    if not a.skip_batch:
        table["batch"] = tune_batch(res)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(a.out)


if __name__ == "__main__":
    main()
