"""NN-corrected Barnes-Hut versus plain Barnes-Hut at equal error (two-Plummer workload).

1. Training set: `--snapshots` two-Plummer systems (seeded; each seed also seen under a few random
   rotations, which changes how the octree cuts it), `--waves` wavefronts each, rows and
   brute-force-minus-Barnes-Hut targets generated on the GPU at `--theta`.
2. Full-batch Adam on the GPU (`--iters`).
3. A held-out snapshot: plain Barnes-Hut at several theta and NN-corrected Barnes-Hut at
   `--theta`, each timed (tree build + walk, device time from HIP events) and checked against
   the FP64 direct sum on a seeded sample. Reports the particles/s of each, the error of each,
   the plain-BH theta with the same error as the corrected walk (log-linear interpolation) and
   the resulting speed ratio.
Prints one JSON document.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib, learned, metrics, workloads


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rotation(seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.Philox(key=[seed, 99]))
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def snapshot(n: int, seed: int, rot: int):
    s = workloads.two_plummer_case(n, seed=seed)
    p = s.particles.copy()
    if rot:
        p[:, :3] = (s.pos @ rotation(seed * 1000 + rot).T).astype(np.float32)
    return p


def timed_bh(res, dev, theta, mlp, reps=2):
    import torch

    import particular_amd as pa

    alg = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=mlp)
    out = alg.compute(pa.Between(dev, dev))
    times = []
    for _ in range(reps):
        res.profile(True)
        out = alg.compute(pa.Between(dev, dev))
        torch.cuda.synchronize()
        b, _ = res.profile_read(_lib.PROF_BH_BUILD)
        w, _ = res.profile_read(_lib.PROF_BH_WALK)
        times.append(b + w)
    res.profile(False)
    return out, min(times)


def fp64_reference(res, dev, idx_dev, sample):
    import torch

    out = torch.empty((sample, 3), dtype=torch.float64, device="cuda")
    pts = dev[idx_dev].contiguous()
    res.use_torch_stream()
    res.check(_lib.lib().pnb_direct_f64_dev(res.handle, pts.data_ptr(), sample, dev.data_ptr(),
                                            dev.shape[0], 0.0, out.data_ptr()))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--snapshots", type=int, default=64)
    ap.add_argument("--rotations", type=int, default=8, help="rotations per seed")
    ap.add_argument("--waves", type=int, default=128)
    ap.add_argument("--theta", type=float, default=0.8)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--sample", type=int, default=16384)
    ap.add_argument("--thetas", default="0.3,0.4,0.5,0.6,0.7,0.8")
    ap.add_argument("--params-out", default="")
    a = ap.parse_args()
    import torch

    import particular_amd as pa

    res = pa.HipResources(0)
    learned.init(res, 1808)
    learned.clear(res)
    t0 = time.time()
    seeds = max(1, a.snapshots // a.rotations)
    k = 0
    for sd in range(seeds):
        for rot in range(a.rotations):
            if k >= a.snapshots:
                break
            p = snapshot(a.n, 7000 + sd, rot)
            dev = torch.from_numpy(p).cuda()
            learned.collect(res, dev, a.theta, a.waves, seed=k)
            k += 1
            del dev
        log(f"collected {k} snapshots, rows/waves/dropped = {learned.info(res)}, {time.time() - t0:.0f} s")
    rows, waves, dropped = learned.info(res)
    res.profile(True)
    hist = learned.train(res, a.iters, a.lr)
    train_ms, _ = res.profile_read(_lib.PROF_MLP)
    res.profile(False)
    log(f"trained {a.iters} iterations: loss {hist[0]:.4e} -> {hist[-1]:.4e} ({train_ms:.0f} ms)")
    if a.params_out:
        np.save(a.params_out, learned.get_params(res))

    # held-out evaluation
    p = snapshot(a.n, 9999, 0)
    dev = torch.from_numpy(p).cuda()
    idx = np.random.default_rng(5).choice(a.n, a.sample, replace=False)
    ref = fp64_reference(res, dev, torch.from_numpy(idx).cuda(), a.sample)
    plain = []
    for th in [float(x) for x in a.thetas.split(",")]:
        out, ms = timed_bh(res, dev, th, False)
        err = metrics.summary(out.cpu().numpy()[idx], ref)
        plain.append({"theta": th, "ms": ms, "mparticles_per_s": a.n / ms / 1e3, **err})
        log(f"plain BH theta={th}: {ms:.2f} ms, rms {err['rms_rel']:.3e}")
    out, ms_nn = timed_bh(res, dev, a.theta, True)
    err_nn = metrics.summary(out.cpu().numpy()[idx], ref)
    log(f"NN BH theta={a.theta}: {ms_nn:.2f} ms, rms {err_nn['rms_rel']:.3e}")
    # plain theta with the NN error (interpolate log(rms) and log(ms) linearly in theta)
    th = np.array([q["theta"] for q in plain])
    lr = np.log([q["rms_rel"] for q in plain])
    lt = np.log([q["ms"] for q in plain])
    o = np.argsort(lr)
    target = np.log(err_nn["rms_rel"])
    if lr.min() <= target <= lr.max():
        th_eq = float(np.interp(target, lr[o], th[o]))
        ms_eq = float(np.exp(np.interp(th_eq, th, lt)))
    else:
        th_eq, ms_eq = None, None
    at04 = next((q for q in plain if abs(q["theta"] - 0.4) < 1e-9), None)
    result = {
        "workload": f"two Plummer spheres, {a.n} particles, theta={a.theta}, held-out seed",
        "training": {"snapshots": k, "waves_per_snapshot": a.waves, "rows": rows,
                     "lane_targets": waves * 64, "dropped_rows": dropped, "iterations": a.iters,
                     "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                     "train_ms": train_ms},
        "plain": plain,
        "nn": {"theta": a.theta, "ms": ms_nn, "mparticles_per_s": a.n / ms_nn / 1e3, **err_nn},
        "plain_theta_at_nn_error": th_eq,
        "plain_ms_at_nn_error": ms_eq,
        "speedup_vs_plain_at_equal_error": (ms_eq / ms_nn) if ms_eq else None,
        "nn_error_vs_plain_same_theta": err_nn["rms_rel"] / next(
            q["rms_rel"] for q in plain if abs(q["theta"] - a.theta) < 1e-9)
        if any(abs(q["theta"] - a.theta) < 1e-9 for q in plain) else None,
        "plain_theta_0.4": at04,
    }
    print(json.dumps(result, indent=1))


if __name__ == "__main__":
    main()
