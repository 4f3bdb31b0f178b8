#!/usr/bin/env python
"""Particle-steps/s of the Brownian/LLG step on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 100,000
magnetite particles, lognormal core radius (median 10 nm, sigma_ln = 0.25)
plus a 5 nm shell, volume fraction 10 % in a periodic cube, 50 mT along z,
Gilbert damping 0.05, T = 300 K, water (1e-3 Pa s), Ewald tuned to 1e-6 with
the reference's error model, fixed dt = 1e-12 s Heun steps, in-kernel Philox
thermal noise. Inputs are resident in HBM before the timed region.

One JSON line on stdout; progress on stderr.
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

# algorithmic flop counts per unit of work (exp, erfc, sqrt, div count 1)
FLOP_PAIR_FORCE = 94     # real-space pair, force + field + pressure
FLOP_PAIR_FIELD = 62     # real-space pair, field only
FLOP_K_BUILD = 21        # S(k) build per (k, particle)
FLOP_K_SCATTER_F = 38    # force + field scatter per (k, particle)
FLOP_K_SCATTER_B = 22    # field-only scatter per (k, particle)
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (packed)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_context(mds, workload, precision, rc_frac, seed, info=None):
    if info is None:
        cfg, r, rh, L, pos, mm, ea = workload.system("c3", seed=seed)
        mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
        alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, rc_frac)
    else:
        cfg, r, rh, L, pos, mm, ea = (info[k] for k in ("cfg", "r", "rh", "L", "pos", "mm", "ea"))
        mu, alpha, rc, nc = info["mu"], info["alpha"], info["rc"], info["nc"]
    cfg = mds.config_from_string(workload.config_text("c3", seed=seed, precision=precision))
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(alpha, rc, nc)
    ctx.set_state(pos, mm, ea)
    ctx.use_philox(seed, 0)
    sysinfo = {"n": int(cfg.numb_part), "L": L, "alpha": alpha, "rc": rc, "nc": nc, "cfg": cfg, "r": r, "rh": rh,
               "pos": pos, "mm": mm, "ea": ea, "mu": mu}
    return ctx, sysinfo


def time_path(ctx, steps, warmup):
    """Wall time of `steps` graph-replayed steps (the fixed-dt run path, the
    k-space build beside the pair kernels), then the per-kernel breakdown
    from a second run of `steps` eager steps with HIP-event timing on the
    context's stream (kernels in order, one stream)."""
    t, dt = 0.0, 1e-12
    ctx.set_graph(True)
    t, dt, _ = ctx.run(warmup, t, dt)
    ctx.synchronize()
    t0 = time.perf_counter()
    t, dt, _ = ctx.run(steps, t, dt)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.set_graph(False)
    ctx.enable_timing(True)
    ctx.synchronize()
    t, dt, _ = ctx.run(steps, t, dt)
    ctx.synchronize()
    stats = ctx.kernel_stats()
    counts = ctx.work_counts()
    ctx.enable_timing(False)
    return t1 - t0, stats, counts


def tune(ctx, info, args):
    """The library's autotuner: r_c (alpha, n_c from the reference's error
    model at errTolEwald) and the cell reach, timed on the device."""
    ctx.set_state(info["pos"], info["mm"], info["ea"])
    if args.reach > 0 and args.rc_frac > 0:
        ctx.set_tiling(args.reach)
        res = None
    else:
        rcs = (args.rc_frac,) if args.rc_frac > 0 else (0.15, 0.175, 0.2, 0.2125, 0.225, 0.2375, 0.25, 0.275, 0.3,
                                                         0.35)
        reaches = (args.reach,) if args.reach > 0 else (2, 3, 4)
        res = ctx.autotune(rcs, reaches, reps=3)
        log(f"  autotune: {res} -> {ctx.params()} {ctx.tiling()}")
    ctx.set_state(info["pos"], info["mm"], info["ea"])
    ctx.use_philox(args.seed, 0)
    return res


# kernel class -> kernel symbol (short form) per path, for the committed
# rocprofv3 FETCH_SIZE / WRITE_SIZE summary
SYMBOL = {
    ("pairs_force", "fp64"): "k_pairs_list<true,false>", ("pairs_force", "mixed"): "k_pairs_list<true,true>",
    ("pairs_field", "fp64"): "k_pairs_list<false,false>", ("pairs_field", "mixed"): "k_pairs_list<false,true>",
    ("kbuild", "fp64"): "k_kbuild<8>", ("kbuild", "mixed"): "k_kbuild_mfma",
    ("kscatter_force", "fp64"): "k_kscatter<true>", ("kscatter_force", "mixed"): "k_kscatter_mfma<true>",
    ("kscatter_field", "fp64"): "k_kscatter<false>", ("kscatter_field", "mixed"): "k_kscatter<false>",
}
PROFILE = os.path.join(ROOT, "profiles", "c3_kernels_mi355x.md")


def profiled_traffic(cls, prec):
    """HBM bytes per dispatch (read + written) of a kernel class from the
    committed profiler summary (counters scaled by a copy-kernel reading), or
    None when the summary does not list it."""
    try:
        with open(PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == SYMBOL.get((cls, prec)):
                    return float(parts[2])
    except OSError:
        pass
    return None


def roofline(stats, counts, n, nk, steps, prec):
    """Dominant kernel class, its algorithmic flops per launch over its
    event-timed average duration."""
    lf, _ = stats["pairs_force"]
    lb, _ = stats["pairs_field"]
    # each unique pair is evaluated once from each of its particles
    pairs_per_launch = counts["pair_evals"] / 2 / max(1, lf + lb)
    per_launch = {
        "pairs_force": FLOP_PAIR_FORCE * pairs_per_launch,
        "pairs_field": FLOP_PAIR_FIELD * pairs_per_launch,
        "kbuild": FLOP_K_BUILD * nk * n,
        "kscatter_force": FLOP_K_SCATTER_F * nk * n,
        "kscatter_field": FLOP_K_SCATTER_B * nk * n,
    }
    dom = max(per_launch, key=lambda k: stats[k][1])
    launches, total_ms = stats[dom]
    avg_s = total_ms / 1e3 / max(1, launches)
    achieved = per_launch[dom] / avg_s / 1e12
    breakdown = {k: {"launches": v[0], "ms_per_step": v[1] / steps} for k, v in stats.items() if v[0]}
    peak = PEAK_FP64_TFLOPS if prec == "fp64" else PEAK_FP32_TFLOPS
    unit = ("FP64 vector ALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)" if prec == "fp64" else
            "FP32 vector ALU (packed FP32 disabled; 157.3 TFLOP/s is the packed peak)")
    note = ("; the pair kernels are vector code (no dense contraction)" if prec == "fp64" else
            "; the pair kernels are vector code (no dense contraction), the k-space kernels FP32 matrix-core code")
    # the dominant kernels are vector-ALU code: "mfma" is the schema's word for
    # a compute bound (as against "hbm"), with the unit named beside it
    bound = "mfma" if dom.startswith("k") and prec == "mixed" else "vector-alu"
    return {"kernel": dom, "bound": bound, "compute_unit": unit + note,
            "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": profiled_traffic(dom, prec), "traffic_source": os.path.relpath(PROFILE, ROOT),
            "avg_launch_ms": avg_s * 1e3, "flop_per_launch": per_launch[dom],
            "unique_pairs_per_launch": pairs_per_launch, "kvecs": nk}, breakdown


def mix64(z):
    """splitmix64 finaliser: replica seeds from (master seed, replica)."""
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def ensemble(mds, workload, precision, replicas, steps, warmup, master=20261203):
    """Batched ensemble: `replicas` independent suspensions in one context
    (every kernel launch covers all of them), N uniform in [500, 20000],
    volume fraction uniform in [1 %, 10 %], applied field on a 0-100 mT grid,
    Philox seed hash(master, replica); monodisperse 10 nm cores + 5 nm shell,
    alpha_G 0.1, fixed dt 1e-12 s; r_c / L and the cell reach autotuned on the
    whole batch; steps replayed from a HIP graph."""
    rng = np.random.default_rng(master)
    ns = rng.integers(500, 20001, replicas)
    phis = rng.uniform(0.01, 0.10, replicas)
    fields = np.linspace(0.0, 0.1, replicas)
    cfg = mds.config_from_string(workload.config_text("c1", seed=master, precision=precision))
    ctx = mds.Context(cfg)
    t0 = time.perf_counter()
    systems = []
    for k in range(replicas):
        n = int(ns[k])
        r = np.full(n, cfg.r_mag_mean)
        rh = r + cfg.d_shell
        L = (np.sum(4.0 / 3.0 * cfg.my_pi * rh ** 3) / phis[k]) ** (1.0 / 3.0)
        sd = mix64(master * 1000003 + k)
        pos = mds.init_positions(cfg, rh, L, sd, periodic=True)
        mm, ea = mds.init_orientations(cfg, n, sd ^ 0x5bd1e995)
        systems.append(dict(r_mag=r, r_hydr=rh, l_box=L, field_z=float(fields[k]), seed=sd, pos=pos, mm=mm, ea=ea))
    ctx.add_replicas(systems)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    tuned = ctx.autotune((0.2, 0.25, 0.3, 0.35), (0, 2, 3), reps=1)
    t_tune = time.perf_counter() - t0
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    ctx.run(warmup, 0.0, 1e-12)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(steps, 0.0, 1e-12)
    ctx.synchronize()
    el = time.perf_counter() - t0
    mz = ctx.replica_mz()
    total = int(ns.sum())
    out = {"value": total * steps / el, "unit": "particle-steps/s", "replicas": replicas, "particles": total,
           "steps": steps, "warmup": warmup, "ms_per_step": el / steps * 1e3, "precision": precision,
           "autotune": tuned, "init_s": t_init, "tune_s": t_tune, "device_gib": ctx.device_bytes() / 2 ** 30,
           "mz_first_last": [float(mz[0]), float(mz[-1])], "graph": True}
    ctx.close()
    return out


def serving_leg(precision, n_req, steps_scale, budget_gib, loads=(0.8, 0.5), seed=20261203, cu_fractions=None,
                cu_spans=None, **lane_kw):
    """Request stream through the continuous-batching loop: N log-uniform in
    [1e3, 2e5], steps log-uniform in [1e3, 1e5] x steps_scale, field uniform
    in [0, 100] mT, volume fraction uniform in [1 %, 10 %], own seeds. First
    all requests at t = 0 (saturated throughput), then the same requests with
    Poisson arrivals at each `loads` x the saturated request rate; latency of
    each request against its isolated-solve time (the same request served
    alone on an idle server)."""
    import copy
    from md_sim_wolfschwenger_amd import serving
    rng = np.random.default_rng(seed)
    base = serving.poisson_stream(n_req, 0.0, rng, steps_range=(1e3 * steps_scale, 1e5 * steps_scale))
    kw = dict(precision=precision, hbm_budget=budget_gib * 2 ** 30, **lane_kw)
    if cu_fractions is not None:
        kw["cu_fractions"] = cu_fractions
    if cu_spans is not None:
        kw["cu_spans"] = cu_spans
    last = [time.perf_counter()]

    def progress(msg):
        # at most one line per 20 s: a long leg still shows that it runs
        now = time.perf_counter()
        if now - last[0] >= 20.0:
            last[0] = now
            log(f"    {msg}")

    srv = serving.LaneServer(**kw)
    sat = srv.serve(copy.deepcopy(base), log=progress)
    srv.close()
    work = sum(q.n * q.steps for q in base)
    iso = np.array(serving.isolated_latencies(copy.deepcopy(base), precision, progress=progress))
    out = {"requests": n_req, "steps_range": [1e3 * steps_scale, 1e5 * steps_scale], "n_range": [1e3, 2e5],
           "hbm_budget_gib": budget_gib, "precision": precision,
           "lane_policy": {k: list(v) if isinstance(v, tuple) else v for k, v in lane_kw.items()},
           "saturated": {"throughput": sat["throughput"], "wall_s": sat["wall_s"], "p50_latency_s":
                         sat["p50_latency_s"], "p99_latency_s": sat["p99_latency_s"]},
           "isolated_p50_s": float(np.percentile(iso, 50)), "isolated_p99_s": float(np.percentile(iso, 99)),
           "by_load": []}
    for load in loads:
        rate = load * sat["throughput"] / (work / n_req)
        reqs = copy.deepcopy(base)
        arr = np.random.default_rng(seed + 1)
        t = 0.0
        for q in reqs:
            t += arr.exponential(1.0 / rate)
            q.arrival = t
        srv = serving.LaneServer(**kw)
        ld = srv.serve(reqs, log=progress)
        srv.close()
        lat = np.array([q.latency for q in reqs])
        ratio = lat / iso
        out["by_load"].append({
            "load": load, "arrival_rate_per_s": rate, "offered_particle_steps_per_s": load * sat["throughput"],
            "throughput": ld["throughput"], "wall_s": ld["wall_s"],
            "p50_latency_s": ld["p50_latency_s"], "p99_latency_s": ld["p99_latency_s"],
            "latency_over_isolated_p50": float(np.percentile(ratio, 50)),
            "latency_over_isolated_p99": float(np.percentile(ratio, 99)),
            "lanes": {"bounds_n": list(srv.bounds), "requests": [l["requests"] for l in ld["lanes"]],
                      "cu_ranges": srv.cu_ranges, "seconds": [l["seconds"] for l in ld["lanes"]],
                      "stats": [l["stats"] for l in ld["lanes"]]},
            "worst": [{"n": q.n, "steps": q.steps, "lane": srv.lane_of(q), "latency_s": float(q.latency),
                       "wait_s": float(q.admitted - q.arrival), "isolated_s": float(i), "chunks": q.chunks,
                       "chunks_s": float(q.run_s)}
                      for q, i in sorted(zip(reqs, iso), key=lambda x: -x[0].latency / x[1])[:3]]})
    # the headline figures at the first load (80 %)
    first = out["by_load"][0]
    out.update({k: first[k] for k in ("load", "throughput", "p50_latency_s", "p99_latency_s",
                                      "latency_over_isolated_p50", "latency_over_isolated_p99")})
    return out


def cpu_baseline(sysinfo, threads, name="c3"):
    """One Heun step of a benchmark system (the bench's N = 1e5 one, or C2)
    with the CPU oracle (the reference's algorithm restated: O(N^2) pair
    scan, k-space two-pass sums) on `threads` OpenMP threads."""
    import oracle
    from tests.refcfg import derived, parse_ini
    from md_sim_wolfschwenger_amd import workload
    cfg = sysinfo["cfg"]
    n = sysinfo["n"]
    c = derived(parse_ini(workload.config_text(name)))
    c["fixed_dt"] = 1
    r, rh = sysinfo["r"], sysinfo["rh"]
    props = {"r_mag": r, "r_hydr": rh, "vol_mag": 4 / 3 * cfg.my_pi * r ** 3, "mag_mom": sysinfo["mu"],
             "zeta_rot": 8 * cfg.my_pi * cfg.vis * rh ** 3, "zeta_trans": 6 * cfg.my_pi * cfg.vis * rh}
    s = oracle.System(n, sysinfo["L"], sysinfo["alpha"], sysinfo["rc"], 2 * r.max(), sysinfo["nc"], props, c,
                      threads=threads)
    s.p.fixed_dt = 1
    rng = np.random.default_rng(0)
    z = rng.standard_normal((n, 3)) * 1e-12
    t0 = time.perf_counter()
    out = s.step(sysinfo["pos"], sysinfo["mm"], sysinfo["ea"], 0.0, 1e-12, z, z * 1e-10, z)
    t1 = time.perf_counter()
    assert out["status"] == 0
    return n / (t1 - t0), t1 - t0


def c2_system(mds, workload, seed):
    cfg, r, rh, L, pos, mm, ea = workload.system("c2", seed=seed)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.3)
    return {"n": len(r), "L": L, "alpha": alpha, "rc": rc, "nc": nc, "cfg": cfg, "r": r, "rh": rh, "pos": pos,
            "mm": mm, "ea": ea, "mu": mu}


def adaptive_leg(mds, workload, precision, steps, seed):
    """The reference's default loop on C2 (N = 1e4): adaptive Dormand-Prince
    (SOLVER = 2, src/configuration.ini:22; src/simulation.cpp:796-862 with
    AdjustTimeStep 912-951), Philox noise, r_c = 0.3 L; one host read of the
    error norm per attempt. Particle-steps/s over accepted steps."""
    s = c2_system(mds, workload, seed)
    text = workload.config_text("c2", seed=seed, fixed=False, precision=precision).replace("SOLVER = 0", "SOLVER = 2")
    cfg = mds.config_from_string(text)
    ctx = mds.Context(cfg)
    ctx.set_particles(s["r"], s["rh"], s["L"])
    ctx.set_ewald(s["alpha"], s["rc"], s["nc"])
    ctx.set_state(s["pos"], s["mm"], s["ea"])
    ctx.use_philox(seed, 0)
    t, dt = 0.0, cfg.delta_t_init
    for _ in range(3):
        t, dt, _, _ = ctx.step(t, dt)
    ctx.synchronize()
    t_start, t0 = t, time.perf_counter()
    for _ in range(steps):
        t, dt, _, _ = ctx.step(t, dt)
    el = time.perf_counter() - t0
    ctx.close()
    return {"value": s["n"] * steps / el, "unit": "particle-steps/s", "precision": precision, "n": s["n"],
            "solver": "Dormand-Prince (SOLVER=2), adaptive", "steps": steps, "ms_per_step": el / steps * 1e3,
            "mean_dt_s": (t - t_start) / steps, "absTol": cfg.abs_tol}


def mz_run(args):
    """The C3 system stepped `--mz-run` times (fixed dt 1e-12 s, Philox noise
    of the bench's seed, so both precisions draw the same noise), in chunks
    of 1000 graph-replayed steps with a progress line each; writes the
    per-step <m_z>, the final state and the step counter, from which
    `--mz-resume` continues (same noise stream)."""
    import md_sim_wolfschwenger_amd as mds
    from md_sim_wolfschwenger_amd import workload
    ctx, info = make_context(mds, workload, args.mz_precision, 0.225, args.seed)
    ctx.set_tiling(3)
    t, dt, done, series = 0.0, 1e-12, 0, []
    if args.mz_resume:
        z = np.load(args.mz_resume)
        ctx.set_state(z["pos"], z["mm"], z["ea"])
        t, done, series = float(z["t"]), int(z["steps"]), [z["mz"]]
        ctx.use_philox(args.seed, done)
    ctx.set_graph(True)
    os.makedirs(os.path.dirname(args.mz_out) or ".", exist_ok=True)

    def save():
        pos, mm, ea = ctx.get_state()
        np.savez(args.mz_out, mz=np.concatenate(series), pos=pos, mm=mm, ea=ea, t=t, steps=done)

    t0 = time.perf_counter()
    left = args.mz_run
    while left > 0:
        k = min(1000, left)
        t, dt, mz = ctx.run(k, t, dt, want_mz=True)
        series.append(np.asarray(mz))
        done += k
        left -= k
        log(f"{args.mz_precision} step {done}: <m_z> {series[-1][-1]:.6f} ({time.perf_counter() - t0:.0f} s)")
        if done % 10000 == 0:
            save()  # a resumable point
    save()
    ctx.close()
    print(json.dumps({"mz_run": args.mz_precision, "steps": done, "mz_last": float(np.concatenate(series)[-1])}))


def copy_calibration(mib):
    """Device-to-device copies of `mib` MiB (the HIP runtime's blit kernel):
    the reference reading for scaling the FETCH_SIZE / WRITE_SIZE counters of
    a profiler run to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    size = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(size)):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(size))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(size), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": size, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed steps (default: one chunk of the run loop, whose first step is checked eagerly)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rc-frac", type=float, default=0.0, help="r_c / L; 0 = autotune")
    ap.add_argument("--reach", type=int, default=0, help="cell reach of the pair search; 0 = autotune")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-overlap", action="store_true",
                    help="k-space chain in order on one stream also in the timed run (for profiles whose kernel "
                         "durations match the per-kernel breakdown)")
    # the GPU host's CPU share is 16 cores per GPU; more threads would only
    # oversubscribe it
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--adaptive-steps", type=int, default=50,
                    help="accepted steps of the adaptive Dormand-Prince leg at C2 (0 = skip)")
    ap.add_argument("--replicas", type=int, default=256, help="replicas of the batched-ensemble leg (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--requests", type=int, default=100, help="requests of the serving leg (0 = skip)")
    ap.add_argument("--steps-scale", type=float, default=0.005,
                    help="serving steps are log-uniform in [1e3, 1e5] times this (the configured stream's mean "
                         "request is 8e8 particle-steps, about a minute of the whole device: 100 of them take hours)")
    ap.add_argument("--serving-precisions", default="mixed", help="comma-separated: mixed, fp64")
    ap.add_argument("--serving-loads", default="0.8,0.5",
                    help="comma-separated offered loads of the serving leg, as fractions of its saturated rate")
    ap.add_argument("--serving-max-active", default="0,0,3",
                    help="per serving lane, at most this many requests at once (0: no limit)")
    ap.add_argument("--serving-order", default="fifo,fifo,sjf",
                    help="per serving lane, admission order: fifo (arrival) or sjf (least N x steps first)")
    ap.add_argument("--serving-overlap-lanes", default="0,0,0",
                    help="per serving lane, 1: the k-space chain and contact search on helper streams")
    ap.add_argument("--serving-cu-spans", default="",
                    help="per serving lane first:last fraction of the CU ids, e.g. 0:0.125,0.125:0.375,0.375:1 "
                         "(empty: the server's default spans)")
    ap.add_argument("--hbm-budget-gib", type=float, default=32.0)
    ap.add_argument("--mz-run", type=int, default=0,
                    help="only run the C3 system for this many steps at r_c = 0.225 L, reach 3, and write its "
                         "<m_z> series (the long-run mixed-vs-FP64 magnetisation check)")
    ap.add_argument("--mz-precision", default="mixed", choices=("fp64", "mixed"))
    ap.add_argument("--mz-resume", default="", help="state file of an earlier --mz-run to continue from")
    ap.add_argument("--mz-out", default="mz_run.npz")
    ap.add_argument("--peak", action="store_true",
                    help="measure the device's FP64 / FP32 vector FMA rate and copy bandwidth, print one JSON line")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run device-to-device copies of this size (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
        return
    if args.mz_run > 0:
        mz_run(args)
        return
    if args.peak:
        from md_sim_wolfschwenger_amd import capi
        res = capi.peak_probe(0)
        res.update(cus=capi.lib().mds_device_cus(0), datasheet={"fp64_vector_tflops": PEAK_FP64_TFLOPS,
                                                               "fp32_packed_tflops": PEAK_FP32_TFLOPS,
                                                               "hbm_gbs": 8000.0})
        print(json.dumps(res))
        return
    import md_sim_wolfschwenger_amd as mds
    from md_sim_wolfschwenger_amd import workload

    log("building the N=1e5 workload")
    ctx, info = make_context(mds, workload, "fp64", args.rc_frac if args.rc_frac > 0 else 0.3, args.seed)
    n = info["n"]
    nk = ctx.work_counts()["kvecs"]
    log(f"L={info['L']:.4e} alpha*L={info['alpha'] * info['L']:.2f} rc/L={args.rc_frac} nc={info['nc']} nk={nk}")
    results = {}
    for prec in ("fp64", "mixed"):
        if prec == "mixed":
            ctx.close()
            ctx, _ = make_context(mds, workload, "mixed", args.rc_frac, args.seed, info)
        ctx.set_kspace_overlap(not args.no_overlap)
        tuned = tune(ctx, info, args)
        tiling = dict(ctx.tiling(), **ctx.params())
        nk = ctx.work_counts()["kvecs"]
        el, stats, counts = time_path(ctx, args.steps, args.warmup)
        value = n * args.steps / el
        roof, breakdown = roofline(stats, counts, n, nk, args.steps, prec)
        log(f"{prec}: {value:.4e} particle-steps/s, {el / args.steps * 1e3:.3f} ms/step, tiling {tiling}")
        for k, v in breakdown.items():
            log(f"  {k:16s} {v['ms_per_step']:8.3f} ms/step ({v['launches']} launches)")
        # one CompInteractions of the bench's initial state, for the error of
        # the mixed path against the FP64 path (itself pinned to the oracle)
        ctx.set_state(info["pos"], info["mm"], info["ea"])
        ctx.set_noise(None, None, None)
        F, B, _, _ = ctx.comp_interactions()
        results[prec] = {"value": value, "el": el, "roof": roof, "breakdown": breakdown, "tiling": tiling,
                         "F": F, "B": B, "nk": nk}
    ctx.close()
    # the CPU baseline runs the FP64 path's tuned Ewald parameters
    p64 = results["fp64"]["tiling"]
    info.update(alpha=p64["alpha"], rc=p64["r_cut"], nc=p64["nc"])
    cfg = info["cfg"]
    m_min, rh_min = float(info["mu"].min()), float(info["rh"].min())
    f_al = 3 * cfg.mu0 * m_min ** 2 / (2 * cfg.my_pi * (2 * rh_min) ** 4)
    b_al = cfg.mu0 * m_min / (2 * cfg.my_pi * (2 * rh_min) ** 3)
    mom = info["mm"] * info["mu"][:, None]
    rel = lambda a, b, fl: float(np.abs(a - b).max() / max(np.abs(b).max(), fl))
    Bf, Bm, Ff, Fm = results["fp64"]["B"], results["mixed"]["B"], results["fp64"]["F"], results["mixed"]["F"]
    mixed_err = {"field": rel(Bm, Bf, b_al), "force": rel(Fm, Ff, f_al),
                 "torque": rel(np.cross(mom, Bm), np.cross(mom, Bf), 0.0),
                 "norm": "max|mixed - fp64| / max(max|fp64|, contact value), over all particles and components"}
    log(f"mixed vs fp64 error: {mixed_err}")
    from md_sim_wolfschwenger_amd import capi
    pred = {}
    for prec in ("fp64", "mixed"):
        tl = results[prec]["tiling"]
        pf, pt = capi.ewald_predicted_errors(cfg, info["mu"], info["rh"], info["L"], tl["alpha"], tl["r_cut"],
                                             tl["nc"])
        pred[prec] = {"force_over_contact": pf, "torque_over_contact": pt}
    log(f"predicted Ewald rms errors: {pred}")
    adapt = {}
    if args.adaptive_steps > 0:
        for prec in ("fp64", "mixed"):
            log(f"adaptive Dormand-Prince at C2, {prec}")
            adapt[prec] = adaptive_leg(mds, workload, prec, args.adaptive_steps, args.seed)
            log(f"  {adapt[prec]}")
    value = results["mixed"]["value"]
    el = results["mixed"]["el"]
    roof, breakdown = results["mixed"]["roof"], results["mixed"]["breakdown"]
    ens = {}
    if args.replicas > 0:
        for prec in ("fp64", "mixed"):
            log(f"batched ensemble: {args.replicas} replicas, {prec}")
            ens[prec] = ensemble(mds, workload, prec, args.replicas, args.ensemble_steps, 2)
            log(f"  {ens[prec]}")
    serv = {}
    if args.requests > 0:
        for prec in args.serving_precisions.split(","):
            log(f"serving: {args.requests} requests, {prec}")
            serv[prec] = serving_leg(prec, args.requests, args.steps_scale, args.hbm_budget_gib,
                                     loads=tuple(float(x) for x in args.serving_loads.split(",")),
                                     max_active=tuple(int(x) for x in args.serving_max_active.split(",")),
                                     sjf=tuple(x == "sjf" for x in args.serving_order.split(",")),
                                     overlap_lanes=tuple(x == "1" for x in args.serving_overlap_lanes.split(",")),
                                     cu_spans=(tuple(tuple(float(v) for v in span.split(":"))
                                                     for span in args.serving_cu_spans.split(","))
                                               if args.serving_cu_spans else None))
            log(f"  {serv[prec]}")
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: one oracle Heun step at N={n} on {args.cpu_threads} threads")
        cv, cs = cpu_baseline(info, args.cpu_threads)
        nproc = os.cpu_count() or 0
        cpu = {"value": cv, "unit": "particle-steps/s", "cores": args.cpu_threads, "kind": "port",
               "host_cpus": nproc, "affinity_cpus": len(os.sched_getaffinity(0)),
               "sample": f"one fixed-dt Heun step of the same N={n} system ({cs:.1f} s), CPU oracle "
                         f"(the reference's O(N^2) pair scan and two-pass k sums restated in C++, OpenMP on "
                         f"{args.cpu_threads} threads: the GPU host's CPU share is 16 cores per GPU)"}
        # the reference is single-threaded: 1 thread against all cores on C2
        # (N = 1e4, one Heun step), where one thread finishes in seconds
        c2 = c2_system(mds, workload, args.seed)
        legs = {}
        for th in (1, args.cpu_threads):
            log(f"CPU baseline: one oracle Heun step at C2 (N=1e4) on {th} threads")
            v, secs = cpu_baseline(c2, th, "c2")
            legs[str(th)] = {"value": v, "seconds": secs}
        cpu["c2_threads"] = {"unit": "particle-steps/s", "sample": "one fixed-dt Heun step of C2 (N=1e4, "
                             "lognormal sigma_ln 0.2, phi=0.05, 10 mT, r_c=0.3 L)", "by_threads": legs,
                             "speedup_all_cores": legs[str(args.cpu_threads)]["value"] / legs["1"]["value"]}
    line = {
        "metric": "particle-steps/s at N=1e5 mixed vs FP64; p50/p99 ensemble-request latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "mixed (f32 real-space pair terms and k-space matrix-core products, f64 geometry and accumulation)", "data": "synthetic",
        "config": {"workload": "C3: N=100000 lognormal core (median 10 nm radius, sigma_ln 0.25) + 5 nm shell, "
                               "phi=0.10, B=50 mT z, alpha_G=0.05, T=300 K, eta=1e-3 Pa s, Ewald 1e-6, "
                               "fixed dt=1e-12 s Heun, Philox noise",
                   "n": n, "precision": "mixed", "ewald_and_tiling": results["mixed"]["tiling"],
                   "kvecs": results["mixed"]["nk"], "fp64_ewald_and_tiling": results["fp64"]["tiling"]},
        "fp64_value": results["fp64"]["value"], "mixed_value": results["mixed"]["value"],
        "mixed_over_fp64": results["mixed"]["value"] / results["fp64"]["value"],
        "mixed_err_vs_fp64": mixed_err,
        "ewald_predicted_rms_error": pred,
        "adaptive_dormand_prince_c2": adapt or None,
        "roofline": roof, "roofline_fp64": results["fp64"]["roof"], "kernel_ms_per_step": breakdown,
        "kernel_ms_per_step_fp64": results["fp64"]["breakdown"], "cpu_baseline": cpu,
        "ensemble": ens or None,
        "serving": serv or None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
