#!/usr/bin/env python3
"""Benchmark of the MI355X registration hot path.

One step = one batched affine cost evaluation of 64 candidate transforms
(one optimiser line-search grid) for the 182x218x182 mutual-information
pair (64x64-bin joint histograms), inputs resident in HBM.  The JSON line
also carries the 256-bin (reference-semantics) and correlation rates, the
single-candidate rate, CDM iterations/s at 182x218x182, the roofline of
the dominant kernel and the CPU reference timed on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "affine cost GVox/s; CDM iter/s @182x218x182; served regs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
CANDIDATES = 64


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_candidates(rng, K, shape_from, shape_to):
    import tipl_amd
    from tipl_amd import phantoms
    P = phantoms.candidate_params(rng, K)
    Ts = np.stack([tipl_amd.affine_matrix(p, shape_from, (1, 1, 1), shape_to, (1, 1, 1)) for p in P])
    return P, Ts


def time_steps(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.synchronize()
    return time.perf_counter() - t0


# Each GPU comes with a share of 16 host cores; the CPU legs use at most that
# many threads (fewer if the process may run on fewer cores).
CPU_SHARE = 16


def cpu_info():
    """The host's CPU model, core counts and the thread count the CPU legs use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "threads": max(1, min(CPU_SHARE, affinity))}


def ref_cost_rate(kind, F, M, Ts, threads, seconds_target):
    """The reference's CPU cost (oracle/_ref: "mi" 256 bins or "corr") on a
    bounded sample of the candidate set, candidates spread over `threads` host
    threads as the reference's optimisers do (each evaluation on one thread);
    the set is repeated until about `seconds_target` of work.  None without
    the reference build."""
    from oracle.bindings import load_reference
    ref = load_reference()
    if ref is None:
        return None
    ref.set_threads(threads)
    t0 = time.perf_counter()
    ref.cost_batch(kind, F, M, Ts[:threads])
    # wall time per evaluation with every thread busy (one evaluation each)
    one = (time.perf_counter() - t0) / max(1, min(threads, len(Ts)))
    n = max(threads, int(round(seconds_target / max(one, 1e-4))))
    Ta = np.asarray(Ts, np.float32).reshape(-1, 12)
    T = np.tile(Ta, (-(-n // len(Ta)), 1))[:n]
    t0 = time.perf_counter()
    ref.cost_batch(kind, F, M, T)
    dt = time.perf_counter() - t0
    evals = len(T) / dt
    name = "MI (256-bin)" if kind == "mi" else "correlation"
    return {"value": round(F.size * evals / 1e9, 4), "unit": "GVox/s", "cores": threads,
            "kind": "reference", "evals_per_s": round(evals, 3),
            "sample": f"{len(T)} {name} evaluations of the {'x'.join(map(str, F.shape[::-1]))} "
                      f"pair (the {len(Ta)} candidates in turn), {threads} threads, {dt:.1f} s"}


def cpu_baseline(F, M, Ts, threads, seconds_target=15.0):
    """The reference's CPU MI cost, or the oracle's (one thread) without the
    reference build."""
    r = ref_cost_rate("mi", F, M, Ts, threads, seconds_target)
    if r is not None:
        return r
    from oracle.bindings import load_oracle
    orc = load_oracle()
    t0 = time.perf_counter()
    K = 0
    while time.perf_counter() - t0 < seconds_target:
        orc.mi_cost(F, M, Ts[K % len(Ts)])
        K += 1
    dt = time.perf_counter() - t0
    evals = K / dt
    return {"value": round(F.size * evals / 1e9, 4), "unit": "GVox/s", "cores": 1, "kind": "port",
            "evals_per_s": round(evals, 3),
            "sample": f"{K} MI (256-bin) oracle evaluations, 1 thread, {dt:.1f} s"}


# Issue cost of one wave64 VALU instruction on this GPU, measured with a
# timing probe of single instruction kinds (8 waves per SIMD, 8 independent
# chains each): FP32 add/mul/sub and bitwise ops, and conversions, v_fract,
# 24-bit multiplies, v_perm, v_cndmask.
VALU_CYCLES_FAST, VALU_CYCLES_SLOW = 2.5, 4.3
SIMDS, CLOCK_GHZ = 1024, 2.4


def valu_issue(prof, kernel_ms):
    """The binding resource of the affine kernels: VALU issue.  From the
    committed counter pass (SQ_INSTS_VALU per launch and the kernel's mix of
    fast and slow instructions): the time the SIMDs need just to issue the
    launch's VALU instructions, against the measured kernel time."""
    if not prof or "valu_insts" not in prof:
        return None
    per_simd = prof["valu_insts"] / SIMDS
    cyc = prof["slow_frac"] * VALU_CYCLES_SLOW + (1.0 - prof["slow_frac"]) * VALU_CYCLES_FAST
    issue_ms = per_simd * cyc / (CLOCK_GHZ * 1e9) * 1e3
    return {"resource": "VALU issue", "valu_instructions_per_launch": prof["valu_insts"],
            "simds": SIMDS, "clock_ghz": CLOCK_GHZ,
            "cycles_per_instruction": round(cyc, 2),
            "slow_instruction_fraction": prof["slow_frac"],
            "issue_ms": round(issue_ms, 4), "frac": round(issue_ms / kernel_ms, 3),
            "source": prof["source"]}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` as the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes measured them (calibrated against the copy kernel), or
    None when no such profile is committed."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None


def cpu_cdm_baseline(It, Is, threads, seconds_target=12.0):
    """The reference's cdm() loop body at one 182x218x182 level (its own step
    functions and par_for threads), timed over a bounded number of iterations."""
    from oracle.bindings import load_reference
    ref = load_reference()
    if ref is None:
        return None
    ref.set_threads(threads)
    one = ref.cdm_iterations(It, Is, 1)
    iters = max(1, min(20, int(seconds_target / max(one, 1e-3))))
    # one iteration already about the sample's length: it is the sample
    dt = one if iters <= 2 else ref.cdm_iterations(It, Is, iters)
    iters = 1 if iters <= 2 else iters
    return {"value": round(iters / dt, 4), "unit": "iter/s", "cores": threads, "kind": "reference",
            "sample": f"{iters} iterations of the reference cdm() loop body at 182x218x182, "
                      f"{threads} threads, {dt:.1f} s"}


SERVE_SIZES = [(96, 96, 96), (128, 128, 128), (182, 218, 182), (256, 256, 256)]


def serving_request(i):
    """Request i of the serving workload: size from SERVE_SIZES, 70% affine /
    30% affine+CDM, int16 or float32, pair generated from seed 4000 + i."""
    from tipl_amd import phantoms
    rng = np.random.default_rng(4000 + i)
    shape = SERVE_SIZES[int(rng.integers(0, len(SERVE_SIZES)))]
    nonlinear = bool(rng.random() < 0.3)
    dtype = np.int16 if rng.random() < 0.5 else np.float32
    fixed, moving, _ = phantoms.serving_pair(4000 + i, shape, nonlinear, dtype)
    return fixed, moving, nonlinear


def batch_pair(i):
    """Pair i of the multi-pair workload: a cube with side uniform in [64, 192],
    the tissue phantom and its remapped, affinely moved copy (seed 3000 + i)."""
    from tipl_amd import phantoms
    rng = np.random.default_rng(3000 + i)
    side = int(rng.integers(64, 193))
    fixed, moving, _ = phantoms.serving_pair(3000 + i, (side, side, side), False, np.float32)
    return fixed, moving


def batch_reg_pair(i):
    """Pair i of the batched-registration workload: a cube with side uniform
    in [64, 192], the tissue phantom and its remapped copy under a random
    affine and a smooth nonlinear warp, float32 (seed 3200 + i)."""
    from tipl_amd import phantoms
    rng = np.random.default_rng(3200 + i)
    side = int(rng.integers(64, 193))
    fixed, moving, _ = phantoms.serving_pair(3200 + i, (side, side, side), True, np.float32)
    return fixed, moving


def make_batch_reg_pairs(n):
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(8, n)) as pool:
        return pool.map(batch_reg_pair, range(n))


def batch_registration_bench(pairs, slots=16):
    """The batch of uneven affine+CDM pairs registered one after another on
    the plain path (tiplhip_register) against all of them submitted to the
    server at once (their candidate sets merged into shared multi-pair
    launches, their CDM iterations into batched steps); the served run is
    repeated with CDM batching off to measure what it adds.  Every served
    result is compared with the plain one."""
    import tipl_amd
    ctx = tipl_amd.Context(0)
    tipl_amd.register(ctx, pairs[0][0], pairs[0][1], nonlinear=True)  # warm-up
    t0 = time.perf_counter()
    plain = [tipl_amd.register(ctx, f, m, nonlinear=True, want_field=False) for f, m in pairs]
    plain_dt = time.perf_counter() - t0
    ctx.close()

    def served_run(cdm_batch):
        os.environ["TIPLHIP_SERVER_CDM_BATCH"] = "1" if cdm_batch else "0"
        try:
            with tipl_amd.Server(0, slots=slots, hbm_budget=192 << 30) as srv:
                t0 = time.perf_counter()
                tickets = [srv.submit(f, m, nonlinear=True, want_field=False) for f, m in pairs]
                served = [srv.wait(t) for t in tickets]
                dt = time.perf_counter() - t0
                st = srv.stats()
        finally:
            os.environ.pop("TIPLHIP_SERVER_CDM_BATCH", None)
        same = all(np.array_equal(a["T"].view(np.uint32), b["T"].view(np.uint32)) and
                   a["evaluations"] == b["evaluations"] and a["cdm_iterations"] == b["cdm_iterations"]
                   and a["r_nonlinear"] == b["r_nonlinear"] for a, b in zip(served, plain))
        return dt, st, same

    alone_dt, _, same_alone = served_run(False)
    served_dt, st, same = served_run(True)
    return {"pairs": len(pairs), "sides": [int(f.shape[0]) for f, _ in pairs], "slots": slots,
            "plain_pairs_per_s": round(len(pairs) / plain_dt, 3),
            "batched_pairs_per_s": round(len(pairs) / served_dt, 3),
            "batched_over_plain": round(plain_dt / served_dt, 2),
            "cdm_iterations": int(sum(r["cdm_iterations"] for r in plain)),
            "cdm_batched_steps": int(st["cdm_steps"]), "cdm_batched_fields": int(st["cdm_fields"]),
            "served_cdm_unbatched_pairs_per_s": round(len(pairs) / alone_dt, 3),
            "cdm_batching_gain": round(alone_dt / served_dt, 3),
            "results_equal_plain": bool(same and same_alone)}


def make_batch_pairs(n):
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(8, n)) as pool:
        return pool.map(batch_pair, range(n))


def multi_pair_bench(ctx, pairs, K=64, reps=5):
    """32 pairs of uneven cubes, K candidates each: one multi-pair launch per
    step (tiplhip_cost_batch_pairs) against one launch per pair, MI 64 bins."""
    import tipl_amd
    from tipl_amd import phantoms
    rng = np.random.default_rng(3100)
    vols, voxels = [], 0
    for f, m in pairs:
        vf, vm = ctx.volume(f, 255.99), ctx.volume(m, 255.99)
        P = phantoms.candidate_params(rng, K)
        Ts = np.stack([tipl_amd.affine_matrix(p, f.shape, (1, 1, 1), m.shape, (1, 1, 1)) for p in P])
        vols.append((vf, vm, Ts))
        voxels += f.size * K
    batched = ctx.cost_batch_pairs(vols, "mi", 64)
    separate = [ctx.cost_batch(vf, vm, Ts, "mi", 64) for vf, vm, Ts in vols]
    equal = all(np.array_equal(a, b) for a, b in zip(batched, separate))
    t_b = time_steps(ctx, lambda: ctx.cost_batch_pairs(vols, "mi", 64), reps, 1) / reps
    t_s = time_steps(ctx, lambda: [ctx.cost_batch(vf, vm, Ts, "mi", 64) for vf, vm, Ts in vols],
                     reps, 1) / reps
    for vf, vm, _ in vols:
        vf.free()
        vm.free()
    return {"pairs": len(pairs), "candidates_per_pair": K, "sides": [int(f.shape[0]) for f, _ in pairs],
            "batched_ms": round(t_b * 1e3, 3), "separate_ms": round(t_s * 1e3, 3),
            "batched_gvox_s": round(voxels / t_b / 1e9, 2),
            "separate_gvox_s": round(voxels / t_s / 1e9, 2),
            "batched_over_separate": round(t_s / t_b, 2), "results_equal": bool(equal)}


def make_serving_requests(n):
    """Generated in worker processes before the GPU is touched (fork start)."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(8, n)) as pool:
        return pool.map(serving_request, range(n))


def serving_bench(reqs, slots, rate_factor, seed=4000):
    """The plain single-pair path (one registration at a time, tiplhip_register)
    against the serving loop on the same requests: (1) saturation -- every
    request submitted at once, throughput = completed / (last completion -
    first submit); (2) a Poisson stream offered at rate_factor x the plain
    path's rate, with latency percentiles (submit to result).  Every served
    result is compared with the plain one, bit for bit."""
    import tipl_amd
    ctx = tipl_amd.Context(0)
    tipl_amd.register(ctx, reqs[0][0], reqs[0][1])  # warm-up (code objects, pools)
    plain = []
    t0 = time.perf_counter()
    for f, m, nl in reqs:
        plain.append(tipl_amd.register(ctx, f, m, nonlinear=nl, want_field=False))
    plain_dt = time.perf_counter() - t0
    plain_rate = len(reqs) / plain_dt
    ctx.close()

    def run(arrivals):
        with tipl_amd.Server(0, slots=slots, hbm_budget=192 << 30) as srv:
            start = time.perf_counter()
            tickets = []
            for (f, m, nl), at in zip(reqs, arrivals):
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit(f, m, nonlinear=nl, want_field=False))
            served = [srv.wait(t) for t in tickets]
            st = srv.stats()
        same = all(np.array_equal(a["T"].view(np.uint32), b["T"].view(np.uint32)) and
                   a["evaluations"] == b["evaluations"] and a["cdm_iterations"] == b["cdm_iterations"]
                   for a, b in zip(served, plain))
        return st, same

    sat, same1 = run(np.zeros(len(reqs)))
    rate = rate_factor * plain_rate
    rng = np.random.default_rng(seed)
    stream, same2 = run(np.cumsum(rng.exponential(1.0 / rate, len(reqs))))
    split = {k: round(float(np.mean([p[k] for p in plain])), 2) for k in ("prep_ms", "affine_ms", "cdm_ms")}
    return {
        "served_regs_per_s": round(sat["regs_per_s"], 3),
        "plain_regs_per_s": round(plain_rate, 3),
        "served_over_plain": round(sat["regs_per_s"] / plain_rate, 2),
        "saturation_latency_p50_ms": round(sat["latency_p50_ms"], 1),
        "saturation_latency_p99_ms": round(sat["latency_p99_ms"], 1),
        "offered_rate_per_s": round(rate, 3),
        "latency_p50_ms": round(stream["latency_p50_ms"], 1),
        "latency_p95_ms": round(stream["latency_p95_ms"], 1),
        "latency_p99_ms": round(stream["latency_p99_ms"], 1),
        "stream_regs_per_s": round(stream["regs_per_s"], 3),
        "requests": len(reqs), "slots": slots,
        "nonlinear_requests": int(sum(nl for _, _, nl in reqs)),
        "batch_launches": int(sat["batch_launches"]), "batched_pairs": int(sat["batched_pairs"]),
        "cdm_batched_steps": int(sat["cdm_steps"]), "cdm_batched_fields": int(sat["cdm_fields"]),
        "hbm_peak_gib": round(sat["hbm_peak"] / 2 ** 30, 2),
        "results_equal_plain": bool(same1 and same2),
        "plain_mean_split_ms": split,
    }


def field_errors(fixed, d, f, step=2):
    """Errors of a CDM field f against the generating displacement d
    (moving(x) = fixed(x + d(x)); the CDM field should satisfy
    f(x) + d(x + f(x)) = 0), on every `step`-th voxel:
      * rms over the phantom (fixed > 150), and the truth's rms there;
      * at edge voxels (the top 10% of |grad fixed|, sigma 1): the rms of the
        residual along the image gradient -- the component an image
        registration can observe -- and the truth's rms along it.
    Inside the piecewise-constant phantom the tangential and interior parts
    of d leave the images unchanged, so no image cost can recover them."""
    from scipy import ndimage
    sl = tuple(slice(0, s, step) for s in fixed.shape)
    zz, yy, xx = np.meshgrid(*[np.arange(0, s, step, dtype=np.float64) for s in fixed.shape],
                             indexing="ij")
    fs = f[sl].astype(np.float64)
    coords = [zz + fs[..., 2], yy + fs[..., 1], xx + fs[..., 0]]
    dd = np.stack([ndimage.map_coordinates(d[..., k], coords, order=1, mode="nearest")
                   for k in range(3)], -1)
    res = fs + dd
    inside = fixed[sl] > 150.0
    fx = fixed.astype(np.float64)
    g = np.stack([ndimage.gaussian_filter(fx, 1.0, order=o)[sl]
                  for o in ((0, 0, 1), (0, 1, 0), (1, 0, 0))], -1)   # (gx, gy, gz)
    gm = np.linalg.norm(g, axis=-1)
    edge = gm > np.percentile(gm, 90)
    gn = g / (gm[..., None] + 1e-12)
    rms = lambda v: round(float(np.sqrt((v ** 2).mean())), 4)
    return {"field_rms_error_voxels": rms(np.linalg.norm(res, axis=-1)[inside]),
            "true_displacement_rms_voxels": rms(np.linalg.norm(d[sl], axis=-1)[inside]),
            "edge_normal_rms_error_voxels": rms((res * gn).sum(-1)[edge]),
            "edge_normal_true_rms_voxels": rms((d[sl] * gn).sum(-1)[edge])}


def cdm_config2(ctx):
    """CDM at 160x192x160: the tissue phantom warped by 30 RBF displacements
    (max 6 voxels, Jacobian > 0), registered two ways:
      (i)  the reference's cdm(): cdm_param defaults, its pyramid, cost-driven
           iteration control and cdm_smooth annealing (reg/cdm.hpp:36-44, 373-474);
      (ii) the fixed schedule: 3 levels x 48 iterations, the field smoothed by
           a sigma-2 binomial gaussian each iteration.
    Reports iterations/s over each whole run and the recovered field's errors."""
    from tipl_amd import phantoms
    fixed, moving, d = phantoms.warped_pair(2000, (160, 192, 160))
    vt, vs = ctx.volume(fixed, 255.99), ctx.volume(moving, 255.99)
    It8, Is8 = vt.read().astype(np.float64), vs.read()
    corr_before = float(np.corrcoef(It8.reshape(-1), Is8.astype(np.float64).reshape(-1))[0, 1])
    out = {"shape": [160, 192, 160], "image_correlation_before": round(corr_before, 4)}
    modes = {"reference_cdm": ctx.cdm_param(),
             "fixed_schedule": ctx.cdm_param(levels=3, level_iterations=48, field_bandwidth=8)}
    for name, p in modes.items():
        ctx.cdm(vt, vs, p)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        rc, f, n = ctx.cdm(vt, vs, p)
        dt = time.perf_counter() - t0
        Js8 = ctx.compose(Is8, f).astype(np.float64)
        r = {"status": int(rc), "iterations": int(n), "seconds": round(dt, 4),
             "iter_per_s": round(n / dt, 2),
             "image_correlation_after": round(float(np.corrcoef(It8.reshape(-1), Js8.reshape(-1))[0, 1]), 4)}
        r.update(field_errors(fixed, d, f))
        if name == "fixed_schedule":
            r.update({"levels": 3, "iterations_per_level": 48, "field_sigma_voxels": 2.0})
        out[name] = r
    vt.free()
    vs.free()
    return out


def cdm2d_bench(ctx, threads, with_cpu, reps=5):
    """The 2-D CDM (cdm<uint8,2>, the reference's cdm_cuda<out,uint8,2>):
    whole multi-level cdm() calls on a smooth slice and its smoothly warped
    copy, on the GPU, and -- with the CPU legs -- the reference's own CPU
    cdm() on the same pair with the host's threads (one call)."""
    from scipy.ndimage import gaussian_filter, map_coordinates
    out = {}
    for shape in ((218, 182), (512, 512)):
        rng = np.random.default_rng(5000 + shape[0])
        a = gaussian_filter(rng.random(shape), 3.0)
        a = ((a - a.min()) / (a.max() - a.min()) * 255.99).astype(np.uint8)
        yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
        w = [gaussian_filter(rng.standard_normal(shape), 12.0) for _ in range(2)]
        w = [2.0 * v / np.abs(v).max() for v in w]
        b = map_coordinates(a.astype(np.float32), [yy + w[0], xx + w[1]], order=1)
        b = np.clip(b, 0, 255).astype(np.uint8)
        ctx.cdm2d(a, b)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            rc, f, iters = ctx.cdm2d(a, b)
        dt = (time.perf_counter() - t0) / reps
        e = {"status": rc, "iterations": iters, "ms_per_call": round(dt * 1e3, 3),
             "iter_per_s": round(iters / dt, 1)}
        if with_cpu:
            from oracle.bindings import load_reference
            ref = load_reference()
            if ref is not None:
                ref.set_threads(threads)
                t0 = time.perf_counter()
                ok, fr, _ = ref.cdm2(a, b)
                dtr = time.perf_counter() - t0
                e["cpu_reference_ms_per_call"] = round(dtr * 1e3, 1)
                e["cpu_threads"] = threads
                e["speedup_vs_cpu"] = round(dtr / dt, 1)
                e["field_rms_vs_reference"] = float(np.sqrt(((f - fr) ** 2).sum(-1).mean()))
        out[f"{shape[1]}x{shape[0]}"] = e
    return out


def mm_reg_pipeline(ctx):
    """mm_reg's nonlinear half on the 160x192x160 warped pair with a given
    (identity) affine: J and r_linear, It_match_fov, cdm in both directions
    (fixed schedule, 3 levels x 48 iterations, sigma-2 field smoothing), both
    mappings, the warped subject and r_nonlinear.  Timed with the reverse
    direction off and on (it runs on a second stream at the same time); the
    host copies of the requested outputs are inside the timing."""
    import tipl_amd
    from tipl_amd import phantoms
    fixed, moving, _ = phantoms.warped_pair(2000, (160, 192, 160))
    ident = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
    kw = dict(nonlinear=True, arg=ident, levels=3, level_iterations=48, field_bandwidth=8)
    out = {"shape": [160, 192, 160], "levels": 3, "iterations_per_level": 48}
    for reverse in (False, True):
        outs = ("to2from", "warped") + (("f2t", "from2to") if reverse else ())
        tipl_amd.register(ctx, fixed, moving, reverse=reverse, outputs=outs, **kw)  # warm-up
        t0 = time.perf_counter()
        r = tipl_amd.register(ctx, fixed, moving, reverse=reverse, outputs=outs, **kw)
        dt = time.perf_counter() - t0
        key = "bidirectional" if reverse else "forward_only"
        out[key + "_s"] = round(dt, 4)
        out[key + "_nonlinear_ms"] = round(r["cdm_ms"], 2)
        out["r_linear"] = round(r["r_linear"], 5)
        out["r_nonlinear"] = round(r["r_nonlinear"], 5)
    out["reverse_overlap"] = round(2 * out["forward_only_nonlinear_ms"] / out["bidirectional_nonlinear_ms"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cdm-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-config2", action="store_true",
                    help="skip the 160x192x160 CDM and mm_reg legs (for short profiling runs)")
    ap.add_argument("--serve-requests", type=int, default=96,
                    help="requests in the serving leg (0 skips it)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--batch-pairs", type=int, default=32,
                    help="pairs in the multi-pair launch and batched-registration legs (0 skips them)")
    ap.add_argument("--serve-rate", type=float, default=1.0,
                    help="Poisson arrival rate of the latency run, as a multiple of the plain "
                         "path's rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each pair or "
              f"candidate set fits one MI355X; registrations shard across GPUs with no exchange, "
              f"one server per GPU)", file=sys.stderr)
        return 2

    reqs = []
    if args.serve_requests > 0:
        log(f"generating {args.serve_requests} serving requests")
        reqs = make_serving_requests(args.serve_requests)
    bpairs, rpairs = [], []
    if args.batch_pairs > 0:
        log(f"generating {args.batch_pairs} uneven pairs (affine, and affine+CDM)")
        bpairs = make_batch_pairs(args.batch_pairs)
        rpairs = make_batch_reg_pairs(args.batch_pairs)

    import tipl_amd
    from tipl_amd import phantoms
    log("generating the 182x218x182 tissue pair")
    fixed, moving, truth = phantoms.tissue_pair(1000)
    ctx = tipl_amd.Context(0)
    vf = ctx.volume(fixed, 255.99)   # quantised on the device (normalize_upper_lower2)
    vm = ctx.volume(moving, 255.99)
    shape = fixed.shape
    N = int(np.prod(shape))
    rng = np.random.default_rng(1000)
    P, Ts = make_candidates(rng, CANDIDATES, shape, shape)

    res = {}
    # --- launch shapes from the autotuner (results do not depend on them) -----
    shapes = {}
    for cost, bins in (("mi", 64), ("mi", 256), ("corr", 256)):
        for K in (64, 1):
            zc, pc, _ = ctx.autotune(vf, vm, Ts[:K], cost, bins, reps=3)
            shapes[f"{cost}{bins if cost == 'mi' else ''}_k{K}"] = [zc, pc]
    log(f"autotuned launch shapes (planes per item, workgroups per CU): {shapes}")

    # --- headline: K=64 MI candidates per launch, 64x64 bins ------------------
    headline = lambda: ctx.cost_batch(vf, vm, Ts, "mi", 64)  # noqa: E731
    time_steps(ctx, headline, 0, args.warmup)
    # kernel timing records event pairs without waiting on them, so the
    # timed region keeps its launch rate; the average covers exactly the
    # timed launches
    ctx.set_kernel_timing(True)
    dt = time_steps(ctx, headline, args.steps, 0)
    kms, kn = ctx.kernel_time()
    ctx.set_kernel_timing(False)
    assert kn == args.steps, (kn, args.steps)
    k_avg_ms = kms / kn
    value = N * CANDIDATES * args.steps / dt / 1e9
    ms_per_step = dt / args.steps * 1e3
    log(f"mi64 K=64: {value:.1f} GVox/s, {ms_per_step:.3f} ms/step, kernel {k_avg_ms:.3f} ms")

    # --- the other affine configurations (same candidate set) -------------
    def rate(cost, bins, K):
        T = Ts[:K]
        n = max(5, args.steps // (2 if K > 1 else 1))
        t = time_steps(ctx, lambda: ctx.cost_batch(vf, vm, T, cost, bins), n, 2)
        return N * K * n / t / 1e9, t / n * 1e3

    res["mi256_k64_gvox_s"], res["mi256_k64_ms"] = rate("mi", 256, 64)
    res["corr_k64_gvox_s"], res["corr_k64_ms"] = rate("corr", 256, 64)
    res["mi64_k1_gvox_s"], res["mi64_k1_ms"] = rate("mi", 64, 1)
    res["mi256_k1_gvox_s"], res["mi256_k1_ms"] = rate("mi", 256, 1)
    res["corr_k1_gvox_s"], res["corr_k1_ms"] = rate("corr", 256, 1)
    for k, v in res.items():
        log(f"{k} = {v:.3f}")

    # --- CDM iterations/s at one 182x218x182 level -------------------------
    log("CDM single level at 182x218x182")
    vt = ctx.volume(fixed, 255.99)
    vs = ctx.volume(phantoms.warp_field(fixed, phantoms.rbf_field(shape, 30, 6.0, 2000)), 255.99)
    copy_gbps = ctx.stream_copy(1 << 30, 10)
    log(f"device copy {copy_gbps:.0f} GB/s")
    ctx.cdm_level_iterations(vt, vs, 2)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.cdm_level_iterations(vt, vs, args.cdm_iters)
    ctx.synchronize()
    cdm_dt = time.perf_counter() - t0
    cdm_ips = args.cdm_iters / cdm_dt
    log(f"CDM {cdm_ips:.1f} iter/s")
    # the CDM iteration against HBM: 520 B of algorithmic traffic per voxel,
    # the reference's step sequence at its fused minimum (compose 14, gradient
    # 14, twelve Jacobi sweeps 12 x 36, accumulate 36, smooth 24); the
    # two-sweep Jacobi kernel and the fused accumulate+smooth move less than
    # that; whole iterations timed, host control included
    cdm_alg = 520.0 * N
    cdm_roofline = {"bound": "hbm", "achieved": round(cdm_alg * cdm_ips / 1e9, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(cdm_alg * cdm_ips / 1e9 / HBM_PEAK_GBS, 4),
                    "algorithmic_bytes": cdm_alg,
                    "traffic": (profiled_traffic("cdm_iteration") or {}).get("bytes"),
                    "what": "one cdm() loop iteration at 182x218x182 (compose, gradient, 12 "
                            "Jacobi sweeps, accumulate + smooth), host control included"}

    # --- roofline of the dominant kernel (MI 64-bin joint histogram; a launch
    # this large takes the z-product-table instantiation) --------------------
    alg_bytes = 2.0 * N * CANDIDATES            # 1 B from + 1 B to footprint per voxel-candidate
    achieved = alg_bytes / (k_avg_ms * 1e-3) / 1e9
    prof = profiled_traffic("mi_hist_kernel<64, true>")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": prof["bytes"] if prof else None,
                "kernel": "mi_hist_kernel<64, true>", "kernel_ms": round(k_avg_ms, 4),
                "algorithmic_bytes": alg_bytes,
                "traffic_source": prof["source"] if prof else None,
                "limiter": valu_issue(prof, k_avg_ms)}

    mp_res = None
    if bpairs:
        log("multi-pair launches")
        mp_res = multi_pair_bench(ctx, bpairs)
        log(f"multi-pair {mp_res}")

    cdm2 = mmreg = None
    if not args.no_config2:
        log("CDM at 160x192x160: the reference cdm() and the fixed schedule")
        cdm2 = cdm_config2(ctx)
        log(f"cdm config 2 {cdm2}")
        log("mm_reg nonlinear half at 160x192x160")
        mmreg = mm_reg_pipeline(ctx)
        log(f"mm_reg pipeline {mmreg}")

    serve = None
    if reqs:
        log("serving leg")
        serve = serving_bench(reqs, args.serve_slots, args.serve_rate)
        log(f"serving {serve}")
    breg = None
    if rpairs:
        log("batched affine+CDM registrations of uneven pairs")
        breg = batch_registration_bench(rpairs)
        log(f"batched registrations {breg}")

    host = cpu_info()
    log("2-D cdm()")
    cdm2d = cdm2d_bench(ctx, host["threads"], not args.no_cpu)
    log(f"cdm 2-D {cdm2d}")

    cpu = cpu_corr = cpu_cdm = None
    cpu_1core = {}
    if not args.no_cpu:
        log(f"CPU reference baselines on {host}")
        F8 = vf.read()
        M8 = vm.read()
        nt = host["threads"]
        cpu = cpu_baseline(F8, M8, Ts, nt)
        log(f"cpu mi {cpu}")
        cpu_corr = ref_cost_rate("corr", F8, M8, Ts, nt, 12.0)
        log(f"cpu corr {cpu_corr}")
        cpu_cdm = cpu_cdm_baseline(vt.read(), vs.read(), nt)
        log(f"cpu cdm {cpu_cdm}")
        # the same legs on one core (shorter samples)
        cpu_1core = {}
        for k, leg in (("mi256", lambda: ref_cost_rate("mi", F8, M8, Ts, 1, 5.0)),
                       ("corr", lambda: ref_cost_rate("corr", F8, M8, Ts, 1, 5.0)),
                       ("cdm", lambda: cpu_cdm_baseline(vt.read(), vs.read(), 1, 6.0))):
            cpu_1core[k] = leg()
            log(f"cpu one core {k} {cpu_1core[k]}")
    # GPU over CPU on the same computation (the CPU legs are the reference's
    # 256-bin MI, correlation and cdm() loop body)
    speedup = {}
    if cpu:
        speedup["mi256_k64_vs_cpu_all_cores"] = round(res["mi256_k64_gvox_s"] / cpu["value"], 1)
    if cpu_corr:
        speedup["corr_k64_vs_cpu_all_cores"] = round(res["corr_k64_gvox_s"] / cpu_corr["value"], 1)
    if cpu_cdm:
        speedup["cdm_vs_cpu_all_cores"] = round(cdm_ips / cpu_cdm["value"], 1)
    for k, r in cpu_1core.items():
        if r:
            gpu = {"mi256": res["mi256_k64_gvox_s"], "corr": res["corr_k64_gvox_s"], "cdm": cdm_ips}[k]
            speedup[f"{k}_vs_cpu_one_core"] = round(gpu / r["value"], 1)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GVox/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded 4-class tissue phantom, monotone remap + random affine)",
        "config": {"workload": "affine MI cost, 64 candidates per launch, 64x64-bin joint "
                               "histogram, 182x218x182 uint8", "shape": list(shape),
                   "candidates_per_step": CANDIDATES, "bins": 64, "launch_shapes": shapes},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cdm_iter_per_s": round(cdm_ips, 3),
        "cdm_roofline": cdm_roofline,
        "cdm_shape": list(shape),
        "cdm_cpu_baseline": cpu_cdm,
        "corr_cpu_baseline": cpu_corr,
        "cpu_one_core": cpu_1core or None,
        "cpu_host": host,
        "speedup_same_computation": speedup or None,
        "hbm_copy_gbps": round(copy_gbps, 1),
        "cdm_config2": cdm2,
        "cdm_2d": cdm2d,
        "mm_reg_pipeline": mmreg,
        "multi_pair": mp_res,
        "serving": serve,
        "batch_registration": breg,
        **{k: round(v, 4) for k, v in res.items()},
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
