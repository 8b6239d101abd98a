"""Throughput of the acceleration path on one MI355X.

Workload (direct sum at scale): 1,048,576 particles, 65,536 massive (mu = 1/65,536, Plummer
sphere, scale 1) and 983,040 massless uniform in a cube of side 4, softening 1e-3,
AccelerationSoftened::checked, kick-drift-kick leapfrog at dt = 1e-3 with FP64 positions. One step
is one full acceleration evaluation (1,048,576 x 65,536 = 6.87e10 pair interactions) plus the
FP64 kick/drift. Inputs are seeded synthetic data, resident in HBM before the timed region.

Prints ONE JSON line. `value` is G pair interactions per second over the timed steps. The roofline
entry uses the brute-force kernel's average launch time from HIP events recorded on its stream in
the timed region, 20 flop per interaction, against the 157.3 TFLOP/s FP32 peak. The CPU baseline
times the oracle's restatement of parallel::BruteForce (same per-particle fold, all host threads
up to 16) on a bounded sample of affected particles against the full massive set.

Further entries of the same line: `barnes_hut` (two Plummer spheres of 4,194,304 bodies: plain
walks and corrected walks with three networks -- the quadrupole term, the trained network and the
analytic octupole term -- across angles and the relative criterion, the equal-error comparison,
the walk's roofline against the brute-force pair rate with counters from profiles/, the CPU
baseline of parallel::BarnesHut), `batch` (2,048 systems, 100 KDK steps in one hipGraph),
`all_massive` (1,048,576 bodies, one-sided brute force against BruteForcePairs) and
`small_barnes_hut` (per-call latency on 64 and 4,096 bodies, ordinary, capturable and replayed
from a hipGraph).
"""
from __future__ import annotations

import argparse
import ctypes
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOP_PER_INTERACTION = 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# A one-GPU box grants this process a share of 16 host threads, whatever nproc reports: the CPU
# baselines use that share and record both counts.
HOST_THREAD_SHARE = 16


def host_cores() -> dict:
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity": affinity,
            "threads_used": max(1, min(affinity, HOST_THREAD_SHARE)),
            "cpu_model": cpu_model()}


def cpu_threads() -> int:
    return host_cores()["threads_used"]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic(name: str):
    return load_profile(f"{name}_traffic")


def load_profile(name: str):
    """A machine-readable profile summary from profiles/ (counters measured with rocprofv3)."""
    path = os.path.join(ROOT, "profiles", f"{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def bh_roofline(rows, bf_result):
    """The plain walk at the angle where its error is closest to 1e-4 rms: accepted pair terms
    per second (live) against the brute-force kernel's pair rate on the same chip (live), with the
    walk's counters (VALU per wavefront step, cache hit rates, HBM traffic) from the rocprofv3
    summary of the same build in profiles/."""
    plain = [r for r in rows if r["criterion"] == "geometric" and not r["learned_correction"]]
    if not plain:
        return None
    r = min(plain, key=lambda r: abs(np.log(r["rms_rel"] / 1e-4)))
    bf_rate = bf_result["config"]["interactions_per_step"] / (
        bf_result["roofline"]["avg_kernel_ms"] * 1e-3) / 1e9
    out = {
        "bound": "issue",
        "kernel": "bh_walk_kernel<0> (plain walk)",
        "theta": r["theta"],
        "rms_rel": r["rms_rel"],
        "achieved": r["accepted_G_terms_per_s"],
        "peak": bf_rate,
        "unit": "G pair terms/s (peak: the brute-force kernel's pair rate, live)",
        "frac": r["accepted_G_terms_per_s"] / bf_rate,
        "walk_ms": r["walk_ms"],
        "wave_steps_per_wave": r["wave_steps_per_wave"],
        "accepted_terms_per_particle": r["accepted_terms_per_particle"],
    }
    prof = load_profile("bh_walk_two_plummer_4M")
    if prof:
        key = min(prof["theta"], key=lambda k: abs(float(k) - r["theta"]))
        c = prof["theta"][key]
        out.update({
            "counters_theta": float(key),
            "valu_per_wave_step": c["valu_per_wave_step"],
            "salu_per_wave_step": c["salu_per_wave_step"],
            "valu_lane_ops_per_accepted_term": c["valu"] * 64 / (
                r["accepted_terms_per_particle"] * 4194304) if float(key) == r["theta"] else None,
            "valu_per_cu_cycle": c["valu_per_cu_cycle"],
            "sqc_hit_rate": c["sqc_hit_rate"],
            "l2_hit_rate": c["l2_hit_rate"],
            "traffic": c["hbm_read_bytes"] + c["hbm_write_bytes"],
            "hbm_GBps": c["hbm_GBps"],
            "hbm_frac": c["hbm_GBps"] / HBM_PEAK_GBS,
            "profile": prof["source"],
        })
    return out


def bench_bf(args):
    import particular_amd as pa
    from particular_amd import _lib, workloads

    L = _lib.lib()
    log("generating workload")
    s = workloads.massive_massless_case()
    n = s.n
    mu = np.ascontiguousarray(s.particles[:, 3])
    pos = np.ascontiguousarray(s.pos)
    vel = np.ascontiguousarray(s.vel)
    dt = 1e-3
    res = pa.HipResources(0)
    log("init")
    res.check(L.pnb_leapfrog_init(res.handle, pos.ctypes.data, vel.ctypes.data, mu.ctypes.data, n,
                                  s.softening, 1))
    views = [ctypes.c_void_p() for _ in range(3)]
    nn, mm = ctypes.c_size_t(), ctypes.c_size_t()
    res.check(L.pnb_leapfrog_views(res.handle, *[ctypes.byref(v) for v in views], ctypes.byref(nn),
                                   ctypes.byref(mm)))
    m = mm.value
    log(f"n={n} m={m}; warmup {args.warmup}")
    res.check(L.pnb_leapfrog_run(res.handle, args.warmup, dt, 0))
    res.synchronize()

    res.profile(True)
    t0 = time.perf_counter()
    res.check(L.pnb_leapfrog_run(res.handle, args.steps, dt, 0))
    res.synchronize()
    t1 = time.perf_counter()
    bf_ms, bf_launches = res.profile_read(_lib.PROF_BF)
    drift_ms, drift_launches = res.profile_read(_lib.PROF_DRIFT)
    res.profile(False)
    elapsed = t1 - t0
    interactions = n * m * args.steps
    log(f"timed {args.steps} steps in {elapsed:.4f} s; bf kernel {bf_ms / max(bf_launches, 1):.3f} ms avg")

    graph_ms = None
    if args.graph:
        res.check(L.pnb_leapfrog_run(res.handle, args.steps, dt, 1))  # capture + first replay
        res.synchronize()
        g0 = time.perf_counter()
        res.check(L.pnb_leapfrog_run(res.handle, args.steps, dt, 1))
        res.synchronize()
        graph_ms = (time.perf_counter() - g0) * 1e3 / args.steps

    # accuracy of the last accelerations against the FP64 direct sum on a sample of affected particles
    import torch

    sample = 16384
    shadow, afg, acc = (v.value for v in views)
    out64 = torch.empty((sample, 3), dtype=torch.float64, device="cuda")
    res.check(L.pnb_direct_f64_dev(res.handle, shadow, sample, afg, m, float(s.softening),
                                   out64.data_ptr()))
    acc_h = np.zeros((n, 4), np.float32)
    res.check(L.pnb_leapfrog_read(res.handle, None, None, acc_h.ctypes.data))
    from particular_amd import metrics

    acc_err = metrics.summary(acc_h[:sample], out64.cpu().numpy())
    log(f"accuracy vs FP64 on {sample} particles: {acc_err}")

    avg_kernel_s = bf_ms / 1e3 / max(bf_launches, 1)
    achieved = FLOP_PER_INTERACTION * n * m / avg_kernel_s / 1e12
    traffic = load_traffic("bf")
    result = {
        "metric": "G interactions/s (BF); Mparticles/s BH+NN vs plain BH @1e-4 rms err, 1 GPU",
        "value": interactions / elapsed / 1e9,
        "unit": "G interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Plummer + uniform cube)",
        "config": {
            "workload": "direct sum at scale: 1,048,576 particles (65,536 massive Plummer, "
                        "983,040 massless cube side 4), eps=1e-3, KDK leapfrog dt=1e-3, FP64 x/v",
            "n_affected": n,
            "n_affecting": m,
            "interactions_per_step": n * m,
        },
        "roofline": {
            "bound": "valu",
            "kernel": "bf_kernel",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "avg_kernel_ms": avg_kernel_s * 1e3,
            "launches": bf_launches,
        },
        "accuracy_vs_fp64": acc_err,
        "drift_kernel_avg_ms": drift_ms / max(drift_launches, 1),
        "graph_ms_per_step": graph_ms,
    }
    res.close()
    return result, s


def cpu_baseline_bf(s, sample: int):
    from oracle import oracle

    threads = cpu_threads()
    massive = s.particles[s.particles[:, 3] != 0]
    aff = s.particles[:sample]
    oracle.brute_force(aff[:256], massive, s.softening, True, threads)  # warm up
    t0 = time.perf_counter()
    oracle.brute_force(aff, massive, s.softening, True, threads)
    dt = time.perf_counter() - t0
    return {
        "value": aff.shape[0] * massive.shape[0] / dt / 1e9,
        "unit": "G interactions/s",
        "cores": threads,
        "kind": "port",
        "host": host_cores(),
        "sample": f"{aff.shape[0]} affected x {massive.shape[0]} massive, one evaluation "
                  f"({dt:.2f} s), C restatement of parallel::BruteForce; CPU: {cpu_model()}",
    }


def time_at_error(rows, rms):
    """Time per evaluation at a given rms error on a (rms, ms) curve, log-log interpolation;
    None outside the measured range."""
    e = np.log([r["rms_rel"] for r in rows])
    t = np.log([r["ms_per_eval"] for r in rows])
    o = np.argsort(e)
    x = np.log(rms)
    if not (e.min() <= x <= e.max()):
        return None
    return float(np.exp(np.interp(x, e[o], t[o])))


def tree_size(res) -> int:
    from particular_amd import _lib

    nodes, level = ctypes.c_size_t(), ctypes.c_int()
    res.check(_lib.lib().pnb_bh_tree_size(res.handle, ctypes.byref(nodes), ctypes.byref(level)))
    return nodes.value


def bench_bh(args):
    """Barnes-Hut on two Plummer spheres (4,194,304 bodies, mu = 1/N): plain at theta 0.2 ... 0.8
    and with the learned correction under three networks -- the quadrupole term (the zero-output
    start of training, learned.quadrupole_params), the trained network (learned.trained_params,
    from particular_amd.nn_bh) and the analytic octupole term (learned.octupole_params) -- with
    the reference's geometric criterion; then each with the relative criterion (a node also opens
    when its first neglected term exceeds alpha times the particle's estimated acceleration;
    theta = 1 as geometric floor) at several alpha. Tree build + walk (+ node network) device time from HIP events (the relative walk's
    time includes its estimate pass), accepted pair terms and wavefront node steps from the walk's
    counters, accuracy against the FP64 direct sum on a seeded sample, and the particles/s ratios
    at equal rms error (1e-4): each curve is also run at the knob (theta or alpha) where its error
    meets 1e-4 (secant steps on the measured curve), and the times are interpolated between the
    rows closest to the target."""
    import torch

    import particular_amd as pa
    from particular_amd import _lib, learned, metrics, workloads

    log(f"generating two-Plummer workload ({args.bh_n})")
    s = workloads.two_plummer_case(args.bh_n)
    dev = torch.from_numpy(s.particles).cuda()
    res = pa.HipResources(0)
    res.use_torch_stream()
    sample = min(16384, s.n)
    idx = np.random.default_rng(5).choice(s.n, sample, replace=False)
    pts = dev[torch.from_numpy(idx).cuda()].contiguous()
    ref = torch.empty((sample, 3), dtype=torch.float64, device="cuda")
    res.check(_lib.lib().pnb_direct_f64_dev(res.handle, pts.data_ptr(), sample, dev.data_ptr(),
                                            s.n, 0.0, ref.data_ptr()))
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    learned.init(res, 0)
    trained = learned.trained_params()
    networks = {"quadrupole": learned.quadrupole_params(), "octupole_reference": learned.octupole_params()}
    if trained is not None:
        networks["trained"] = trained
    current = {"name": None}

    def use(name):
        if current["name"] != name:
            learned.set_params(res, networks[name])
            current["name"] = name

    def run(knob, mlp, rel=False):
        theta, alpha = (1.0, knob) if rel else (knob, None)
        alg = pa.BarnesHut(res, theta, pa.Acceleration.checked(), mlp=mlp, relative=alpha)
        acc = alg.compute(pa.Between(dev, dev))
        torch.cuda.synchronize()
        reps = 3
        res.profile(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            acc = alg.compute(pa.Between(dev, dev))
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        build_ms, nb = res.profile_read(_lib.PROF_BH_BUILD)
        walk_ms, nw = res.profile_read(_lib.PROF_BH_WALK)
        sort_ms, ns = res.profile_read(_lib.PROF_SORT)
        net_ms, nn = res.profile_read(_lib.PROF_MLP)
        res.profile(False)
        steps, accepted = pa.bh_stats(res)
        build_ms, walk_ms, sort_ms = build_ms / nb, walk_ms / nw, sort_ms / max(ns, 1)
        net_ms = net_ms / max(nn, 1)
        nodes = tree_size(res)
        err = metrics.summary(acc.cpu().numpy()[idx], ref)
        # the walk's counters accumulate over the reps since the last build: one build per call
        row = {
            "theta": theta,
            "criterion": "relative" if rel else "geometric",
            "alpha": alpha,
            "knob": knob,
            "learned_correction": mlp,
            "network": current["name"] if mlp else None,
            "mparticles_per_s": s.n / wall / 1e6,
            "ms_per_eval": wall * 1e3,
            "build_ms": build_ms,
            "sort_ms": sort_ms,
            "walk_ms": walk_ms,
            "node_network_ms": net_ms if mlp else 0.0,
            "nodes": int(nodes),
            "accepted_terms_per_particle": accepted / s.n,
            "wave_steps_per_wave": steps / (s.n / 64),
            "node_record_GBps": steps * 32 / (walk_ms * 1e-3) / 1e9,
            "accepted_G_terms_per_s": accepted / (walk_ms * 1e-3) / 1e9,
            **err,
        }
        log(f"BH{(' + ' + current['name']) if mlp else ''} theta={theta}"
            f"{f' alpha={alpha}' if rel else ''}: {row['ms_per_eval']:.2f} ms "
            f"({row['mparticles_per_s']:.1f} Mp/s), build {build_ms:.2f} ms, walk {walk_ms:.2f} ms, "
            f"rms {err['rms_rel']:.2e}")
        return row

    target = 1e-4

    def refine(rows, mlp, rel=False, net=None):
        """Adds runs at the knob (theta, or alpha for the relative criterion) where the rms
        error meets the target: secant steps of log(rms) against log(knob) between the closest
        bracketing rows (at most two)."""
        if net:
            use(net)
        for _ in range(2):
            lo = [r for r in rows if r["rms_rel"] <= target]
            hi = [r for r in rows if r["rms_rel"] > target]
            if not lo or not hi:
                return
            a = max(lo, key=lambda r: r["rms_rel"])
            b = min(hi, key=lambda r: r["rms_rel"])
            if abs(np.log(a["rms_rel"] / target)) < 0.02:
                return
            f = np.log(target / a["rms_rel"]) / np.log(b["rms_rel"] / a["rms_rel"])
            k = float(np.exp(np.log(a["knob"]) + f * (np.log(b["knob"]) - np.log(a["knob"]))))
            rows.append(run(float(f"{k:.4g}"), mlp, rel))

    def knobs(text):
        return [float(v) for v in text.split(",") if v]

    plain = [run(th, False) for th in knobs(args.bh_thetas)]
    plain_rel = [run(a, False, True) for a in knobs(args.bh_rel_alphas)]
    refine(plain, False)
    refine(plain_rel, False, True)
    curves = {}
    for net, th_arg, al_arg in (("quadrupole", args.bh_quad_thetas, args.bh_quad_rel_alphas),
                                ("trained", args.bh_nn_thetas, args.bh_nn_rel_alphas),
                                ("octupole_reference", args.bh_oct_thetas, args.bh_oct_rel_alphas)):
        if net not in networks:
            continue
        use(net)
        geo = [run(th, True) for th in knobs(th_arg)]
        rel = [run(a, True, True) for a in knobs(al_arg)]
        refine(geo, True, net=net)
        refine(rel, True, True, net=net)
        curves[net] = (geo, rel)
    res.close()

    def at(rows):
        return time_at_error(rows, target) if rows else None

    def ratio(a, b):
        return (a / b) if (a and b) else None

    t_plain, t_prel = at(plain), at(plain_rel)
    per_net = {}
    for net, (geo, rel) in curves.items():
        t_g, t_r = at(geo), at(rel)
        per_net[net] = {
            "ms_at_target": t_g,
            "mparticles_per_s_at_target": s.n / t_g / 1e3 if t_g else None,
            "speedup_vs_plain": ratio(t_plain, t_g),
            "relative_criterion_ms_at_target": t_r,
            "relative_criterion_speedup_vs_plain_relative": ratio(t_prel, t_r),
            "relative_criterion_speedup_vs_plain_geometric": ratio(t_plain, t_r),
        }
    q, t = per_net.get("quadrupole", {}), per_net.get("trained", {})
    comparison = {
        "target_rms": target,
        "plain_ms_at_target": t_plain,
        "plain_mparticles_per_s_at_target": s.n / t_plain / 1e3 if t_plain else None,
        "plain_relative_ms_at_target": t_prel,
        "plain_relative_vs_plain_geometric": ratio(t_plain, t_prel),
        "networks": per_net,
        # the learned network against its own starting point, each at the same error and with the
        # same criterion
        "trained_vs_quadrupole": ratio(q.get("ms_at_target"), t.get("ms_at_target")),
        "trained_vs_quadrupole_relative_criterion": ratio(q.get("relative_criterion_ms_at_target"),
                                                          t.get("relative_criterion_ms_at_target")),
        "speedup_at_equal_error": t.get("speedup_vs_plain"),
        "network": ("trained: particular_amd/params/two_plummer_4M.npy (particular_amd.nn_bh, 64 "
                    "snapshots; record beside it); quadrupole: its zero-output start "
                    "(learned.quadrupole_params); octupole_reference: the analytic octupole term "
                    "through the same network (learned.octupole_params)"),
    }
    log(f"equal error {target}: plain {t_plain} ms; " + "; ".join(
        f"{k} {v['ms_at_target']} ms (relative {v['relative_criterion_ms_at_target']} ms)"
        for k, v in per_net.items()))
    rows = plain + plain_rel
    for geo, rel in curves.values():
        rows += geo + rel
    return rows, comparison, s


def cpu_baseline_bh(s, theta, sample):
    """The oracle's restatement of parallel::BarnesHut: sequential tree build (timed whole) and
    the threaded traversal timed on a sample of affected particles, scaled to all N."""
    from oracle import oracle

    threads = cpu_threads()
    t0 = time.perf_counter()
    tree = oracle.Tree(s.particles)
    build = time.perf_counter() - t0
    aff = s.particles[np.random.default_rng(3).choice(s.n, sample, replace=False)]
    t0 = time.perf_counter()
    tree.barnes_hut(aff, theta, threads=threads)
    walk = (time.perf_counter() - t0) * s.n / sample
    return {
        "value": s.n / (build + walk) / 1e6,
        "unit": "M particles/s",
        "cores": threads,
        "kind": "port",
        "host": host_cores(),
        "sample": f"theta={theta}: sequential build of all {s.n} bodies ({build:.2f} s) + threaded "
                  f"traversal of {sample} affected scaled to {s.n} ({walk:.2f} s)",
    }


def bench_batch(args):
    """Batched ensemble: 2048 Plummer systems (64..65,536 bodies, log-uniform), eps = 1e-3,
    `--batch-steps` KDK steps captured in one hipGraph."""
    import particular_amd as pa
    from particular_amd import _lib, autotune, workloads

    L = _lib.lib()
    log("generating ensemble")
    e = workloads.ensemble_case()
    res = pa.HipResources(0)
    tuned = autotune.apply(res)
    pos, vel = np.ascontiguousarray(e.pos), np.ascontiguousarray(e.vel)
    mu = np.ascontiguousarray(e.particles[:, 3])
    off = np.ascontiguousarray(e.offsets, np.int64)
    res.check(L.pnb_batch_init(res.handle, pos.ctypes.data, vel.ctypes.data, mu.ctypes.data,
                               off.ctypes.data, e.num_systems, e.softening, 1))
    n, items, inter = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_ulonglong()
    res.check(L.pnb_batch_info(res.handle, ctypes.byref(n), ctypes.byref(items), ctypes.byref(inter)))
    steps = args.batch_steps
    res.check(L.pnb_batch_run(res.handle, 1, 1e-3, 0))
    res.synchronize()
    t0 = time.perf_counter()
    res.check(L.pnb_batch_run(res.handle, steps, 1e-3, 1))
    res.synchronize()
    wall = time.perf_counter() - t0
    res.close()
    return {
        "systems": e.num_systems, "particles": n.value, "work_items": items.value,
        "interactions_per_step": inter.value, "steps": steps, "graph": True, "tuned": tuned,
        "seconds": wall, "g_interactions_per_s": inter.value * steps / wall / 1e9,
    }


def bench_small_bh(args):
    """Latency of one Barnes-Hut evaluation (theta = 0.5, tree rebuilt every call) on small
    Plummer spheres, device tensors, median of `--small-calls` calls each synchronised: the
    ordinary build (two host round trips), the capturable build (none) called eagerly, and the
    capturable build replayed from a hipGraph."""
    import torch

    import particular_amd as pa
    from particular_amd import workloads

    out = {}
    for n in (int(v) for v in args.small_sizes.split(",")):
        s = workloads.direct_sum_case(n, seed=11)
        dev = torch.from_numpy(s.particles).cuda()
        both = pa.Between(dev, dev)  # one device pointer twice: pnb_bh_accel_f32_dev
        row = {}
        for mode in ("ordinary", "capturable", "graph"):
            res = pa.HipResources(0)
            alg = pa.BarnesHut(res, 0.5, pa.Acceleration.checked())
            if mode != "ordinary":
                res.set_bh_capturable(True)
            alg.compute(both)
            alg.compute(both)
            torch.cuda.synchronize()
            call = lambda: alg.compute(both)
            if mode == "graph":
                g = torch.cuda.CUDAGraph()
                # no garbage collection while capturing: a dead cycle freed then would release
                # device memory on the capturing stream
                gc.collect()
                gc.disable()
                try:
                    with torch.cuda.graph(g):
                        alg.compute(both)
                finally:
                    gc.enable()
                res.reset_stream()
                call = g.replay
            times = []
            for _ in range(args.small_calls):
                t0 = time.perf_counter()
                call()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
            if mode != "ordinary":
                res.bh_check()
            res.close()
            row[f"{mode}_us"] = float(np.median(times) * 1e6)
        out[str(n)] = row
        log(f"Barnes-Hut per call at {n} bodies: " +
            ", ".join(f"{k} {v:.0f}" for k, v in row.items()))
    return out


def bench_all_massive(args):
    """All-massive brute force at N = args.massive_n (one Plummer sphere, mu = 1/N,
    AccelerationSoftened::checked(1e-3), every body affecting every other): the one-sided kernel
    (N^2 pair terms, sequential::BruteForce's sums) and the pair kernel (N (N - 1) / 2 pair terms,
    both accelerations from each, sequential::BruteForcePairs' sums). Device time per evaluation
    from HIP events on the library stream; interactions counted as N (N - 1) for both, so the
    rates compare directly; each against the FP64 direct sum on 4,096 sampled bodies."""
    import torch

    import particular_amd as pa
    from particular_amd import _lib, metrics, workloads

    n = args.massive_n
    log(f"generating all-massive Plummer workload ({n})")
    s = workloads.direct_sum_case(n, seed=2024)
    dev = torch.from_numpy(s.particles).cuda()
    res = pa.HipResources(0)
    res.use_torch_stream()
    eps = 1e-3
    sample = min(4096, n)
    idx = np.random.default_rng(9).choice(n, sample, replace=False)
    pts = dev[torch.from_numpy(idx).cuda()].contiguous()
    ref = torch.empty((sample, 3), dtype=torch.float64, device="cuda")
    res.check(_lib.lib().pnb_direct_f64_dev(res.handle, pts.data_ptr(), sample, dev.data_ptr(), n,
                                            eps, ref.data_ptr()))
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    out = {"workload": f"{n} bodies, one Plummer sphere (scale 1), mu = 1/N, "
                       "AccelerationSoftened::checked(1e-3), all massive",
           "interactions_per_eval": n * (n - 1)}
    inter = n * (n - 1)
    for name, alg in (("one_sided", pa.BruteForce(res, pa.AccelerationSoftened.checked(eps))),
                      ("pairs", pa.BruteForcePairs(res, pa.AccelerationSoftened.checked(eps)))):
        acc = alg.compute(dev)
        torch.cuda.synchronize()
        res.profile(True)
        for _ in range(args.massive_reps):
            acc = alg.compute(dev)
        torch.cuda.synchronize()
        ms, launches = res.profile_read(_lib.PROF_BF)
        res.profile(False)
        ms /= max(launches, 1)
        err = metrics.summary(acc.cpu().numpy()[idx], ref)
        out[name] = {"ms_per_eval": ms, "g_interactions_per_s": inter / (ms * 1e-3) / 1e9,
                     "pair_terms_per_eval": n * n if name == "one_sided" else n * (n - 1) // 2,
                     **err}
        log(f"all-massive {name}: {ms:.2f} ms per evaluation "
            f"({out[name]['g_interactions_per_s']:.0f} G interactions/s), rms {err['rms_rel']:.2e}")
    out["pairs_speedup"] = out["one_sided"]["ms_per_eval"] / out["pairs"]["ms_per_eval"]
    prof = load_profile("pairs_kernel_all_massive_1M")
    if prof and n == 1 << 20:
        out["counters"] = {k: prof[k] for k in (
            "valu_lane_ops_per_interaction", "valu_per_cu_cycle",
            "one_sided_valu_lane_ops_per_interaction", "one_sided_valu_per_cu_cycle", "source")}
    res.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", default=True)
    ap.add_argument("--no-graph", dest="graph", action="store_false")
    ap.add_argument("--bh", action="store_true", default=True)
    ap.add_argument("--no-bh", dest="bh", action="store_false")
    ap.add_argument("--bh-n", type=int, default=4194304)
    ap.add_argument("--bh-thetas", default="0.2,0.25,0.3,0.5,0.8", help="plain Barnes-Hut angles")
    ap.add_argument("--bh-quad-thetas", default="0.3,0.4",
                    help="angles with the quadrupole network (empty: none)")
    ap.add_argument("--bh-quad-rel-alphas", default="2e-4,4e-4",
                    help="relative criterion with the quadrupole network (empty: none)")
    ap.add_argument("--bh-oct-thetas", default="0.4,0.5",
                    help="angles with the analytic octupole network (empty: none)")
    ap.add_argument("--bh-oct-rel-alphas", default="3e-4,6e-4",
                    help="relative criterion with the analytic octupole network (empty: none)")
    ap.add_argument("--bh-nn-thetas", default="0.4,0.5",
                    help="angles with the trained network (empty: none)")
    ap.add_argument("--bh-rel-alphas", default="3e-5,1e-4,2e-4",
                    help="plain Barnes-Hut with the relative criterion (empty: none)")
    ap.add_argument("--bh-nn-rel-alphas", default="2e-4,4e-4",
                    help="relative criterion with the trained network (empty: none)")
    ap.add_argument("--bh-cpu-sample", type=int, default=16384)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy a buffer of this size with torch (HBM counter calibration)")
    ap.add_argument("--batch", action="store_true", default=True)
    ap.add_argument("--no-batch", dest="batch", action="store_false")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--massive", action="store_true", default=True,
                    help="all-massive brute force, one-sided and over pairs")
    ap.add_argument("--no-massive", dest="massive", action="store_false")
    ap.add_argument("--massive-n", type=int, default=1048576)
    ap.add_argument("--massive-reps", type=int, default=2)
    ap.add_argument("--small", action="store_true", default=True,
                    help="per-call latency of Barnes-Hut on 64 and 4,096 bodies")
    ap.add_argument("--no-small", dest="small", action="store_false")
    ap.add_argument("--small-calls", type=int, default=200)
    ap.add_argument("--small-sizes", default="64,4096", help="body counts of the latency leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the acceleration path is built and measured on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        import torch

        x = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + written")
        del x, y
    result, s = bench_bf(args)
    result["cpu_baseline"] = None if args.no_cpu else cpu_baseline_bf(s, args.cpu_sample)
    if args.bh:
        bh_rows, comparison, s3 = bench_bh(args)
        result["barnes_hut"] = {
            "workload": f"two Plummer spheres, {s3.n} bodies (scale 1, centres x = -5/+5, "
                        "bulk v_x = +-0.3), mu = 1/N, Acceleration::checked, FP64 centres of mass",
            "rows": bh_rows,
            "corrected_vs_plain": comparison,
            "cpu_baseline": None if args.no_cpu else cpu_baseline_bh(s3, 0.5, args.bh_cpu_sample),
        }
        result["barnes_hut"]["roofline"] = bh_roofline(bh_rows, result)
    if args.batch:
        result["batch"] = bench_batch(args)
    if args.massive:
        result["all_massive"] = bench_all_massive(args)
    if args.small:
        result["small_barnes_hut"] = bench_small_bh(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
