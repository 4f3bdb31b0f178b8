"""Continuous-batching request loop for ensemble solves on one device.

Each request is an independent suspension (N particles, volume fraction,
applied field, Philox seed, number of fixed-dt steps). Active requests are
replicas of one device context (libmdsim.so), so every kernel launch
advances all of them; their state stays resident in HBM from admission to
completion. The loop runs in chunks of steps replayed from a HIP graph; at
each chunk boundary finished requests are retired (their state is copied
out) and arrived requests are admitted while the estimated device footprint
of the active set stays within the HBM budget. A request's result does not
depend on what else runs beside it: its Ewald parameters, cell grid, clock
and Philox counter are its own.

The reference has no counterpart (one system per process,
src/simulation.cpp:12-16, 440-520).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi, workload


@dataclass
class Request:
    rid: int
    n: int
    steps: int
    field_z: float          # T, along z
    phi: float              # volume fraction
    seed: int
    arrival: float = 0.0    # s after the loop starts
    # filled by the loop
    admitted: float = math.nan
    finished: float = math.nan
    mz: float = math.nan
    state: tuple | None = None
    ewald: dict = field(default_factory=dict)
    chunks: int = 0         # chunks of its lane it was part of
    run_s: float = 0.0      # wall time of those chunks

    @property
    def latency(self):
        return self.finished - self.arrival


def rc_frac_for(n: int, scale: float) -> float:
    """r_c / L of a system of n particles: `scale` at n = 1e5 (the autotuned
    optimum there), times (1e5 / n)^(1/6) (the real-space/k-space cost
    balance at fixed accuracy), kept in [0.1, 0.5]."""
    return min(0.5, max(0.1, scale * (1e5 / n) ** (1.0 / 6.0)))


def footprint_bytes(n: int, nc: int, l_box: float = 0.0, r_cut: float = 0.0, skin: float = 0.005) -> int:
    """Device bytes a replica of n particles with k-space cut-off nc adds to
    a context (per-particle arrays, sin/cos tables, S(k) partials, cells),
    with 25 % headroom, plus its pair-list region as the planner sizes it
    (given the box, the cut-off and the list skin)."""
    from .capi import lib
    per_particle = 1250 + 48 * (nc + 1)
    nk = int(2.2 * nc ** 3) + 1
    nseg = max(1, nk // 8)
    nsegw = (nseg + 63) // 64
    chunk = min(1024, max(32, (n * nsegw + 8191) // 8192))  # the planner's chunking
    nchunks = (n + chunk - 1) // chunk
    kspace = nk * 8 * (3 + 2 * nchunks)
    cells = 64 * n
    words = lib().mds_pair_list_words(n, l_box, r_cut, skin) if (l_box > 0 and r_cut > 0 and skin > 0) else 0
    return int(1.25 * (n * per_particle + kspace + cells)) + 4 * int(words)


class Arbiter:
    """Chunks of smaller-request lanes before those of larger ones, without
    stream priorities (which have hung on this hardware): before each chunk
    a lane waits while a lane of smaller requests has a chunk on the device,
    for at most `max_hold` seconds, so a large request's chunks run between
    the small ones' instead of beside them."""

    def __init__(self, lanes: int, max_hold: float = 0.05):
        import threading
        self.busy = [0] * lanes
        self.cv = threading.Condition()
        self.max_hold = max_hold
        self.held = [0.0] * lanes

    def begin(self, k: int):
        t0 = time.perf_counter()
        with self.cv:
            self.cv.wait_for(lambda: not any(self.busy[:k]), timeout=self.max_hold)
            self.busy[k] += 1
        self.held[k] += time.perf_counter() - t0

    def end(self, k: int):
        with self.cv:
            self.busy[k] -= 1
            self.cv.notify_all()


class Budget:
    """Device bytes shared by the lanes of a server (thread-safe)."""

    def __init__(self, total: float):
        import threading
        self.total = total
        self.used = 0
        self.lock = threading.Lock()

    def reserve(self, nbytes: int, force: bool = False) -> bool:
        with self.lock:
            if force or self.used + nbytes <= self.total:
                self.used += nbytes
                return True
            return False

    def release(self, nbytes: int):
        with self.lock:
            self.used -= nbytes


class Server:
    """Serve `Request`s on one device with continuous batching (one lane:
    one device context, one HIP stream)."""

    def __init__(self, precision: str = "fp64", hbm_budget: float = 64 * 2 ** 30, chunk_seconds: float = 0.25,
                 rc_scale: float = 0.225, dt: float = 1e-12, graph: bool = True, base: str = "c1", device: int = 0,
                 budget: Budget | None = None, reserve_particles: int = 0, kspace_overlap: bool = True,
                 max_active: int = 0, sjf: bool = False):
        self.cfg = capi.config_from_string(workload.config_text(base, precision=precision))
        self.precision = precision
        self.hbm_budget = hbm_budget
        self.budget = budget
        self.chunk_seconds = chunk_seconds
        self.max_active = max_active  # 0: no limit
        self.sjf = sjf
        self.rc_scale = rc_scale
        self.dt = dt
        self.device = device
        self.ctx = capi.Context(self.cfg, device)
        self.ctx.set_tiling(0)
        self.ctx.set_kspace_overlap(kspace_overlap)
        self.graph = graph
        self.ctx.set_graph(graph)
        if reserve_particles > 0:
            self.ctx.reserve(reserve_particles)
        self.active: list[Request] = []
        self.done_steps: dict[int, int] = {}
        self.step_time = None  # s per step of the active set, per particle
        self.min_steps = 4     # steps per chunk at least (a chunk has a fixed cost of a few syncs)
        self.times = {"admit": 0.0, "run": 0.0, "retire": 0.0, "idle": 0.0}
        self.trace = None  # a list: admissions, chunks and retirements with their times are appended
        # chunks the context rolled back and redid (flags), and the slowest chunk
        self.stats = {"chunks_redone": 0, "slowest_chunk_s": 0.0, "slowest_chunk_steps": 0,
                      "slowest_chunk_particles": 0}
        self.philox_on = False

    def close(self):
        self.ctx.close()

    # ------------------------------------------------------------ requests
    def params_of(self, req: Request) -> dict:
        """Radii (monodisperse cores + shell of the base configuration), box
        from the volume fraction and Ewald parameters from the reference's
        error model at rc_frac_for(n)."""
        if not req.ewald:
            cfg = self.cfg
            r = np.full(req.n, cfg.r_mag_mean)
            rh = r + cfg.d_shell
            L = (np.sum(4.0 / 3.0 * cfg.my_pi * rh ** 3) / req.phi) ** (1.0 / 3.0)
            mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
            alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, rc_frac_for(req.n, self.rc_scale))
            req.ewald = {"alpha": alpha, "r_cut": rc, "nc": nc, "l_box": L}
        return req.ewald

    def system_of(self, req: Request) -> dict:
        """Initial suspension of a request, drawn on the device: positions by
        parallel random sequential addition, orientations uniform on the
        sphere, seeded by the request."""
        cfg = self.cfg
        p = self.params_of(req)
        r, L, pos, mm, ea = capi.init_suspension_device(cfg, req.n, req.seed, periodic=True, l_box=p["l_box"],
                                                        device=self.device)
        return dict(r_mag=r, r_hydr=r + cfg.d_shell, l_box=L, field_z=req.field_z, seed=req.seed, pos=pos, mm=mm,
                    ea=ea)

    def footprint(self, req: Request) -> int:
        p = self.params_of(req)
        return footprint_bytes(req.n, p["nc"], p["l_box"], p["r_cut"])

    def admit(self, reqs: list[Request], now: float):
        if not reqs:
            return
        systems = [self.system_of(q) for q in reqs]
        first = self.ctx.add_replicas(systems)
        for k, q in enumerate(reqs):
            p = self.params_of(q)
            self.ctx.replica_set_ewald(first + k, p["alpha"], p["r_cut"], p["nc"])
            # the replica stops at its own step count, so chunks need not end
            # when the first of the active requests is done
            self.ctx.replica_stop_at(first + k, q.steps)
            q.admitted = now
            self.done_steps[q.rid] = 0
            self.active.append(q)
        if not self.philox_on:
            # switch to in-kernel draws once: later replicas join with their
            # own seed and a step counter of 0, without resetting the others
            self.ctx.use_philox(reqs[0].seed, 0)
            self.philox_on = True

    def retire(self, now: float):
        keep = []
        for idx in reversed(range(len(self.active))):
            q = self.active[idx]
            if self.done_steps[q.rid] >= q.steps:
                q.state = self.ctx.replica_state(idx)
                q.mz = float(np.mean(q.state[1][:, 2]))
                q.finished = now
                self.ctx.remove_replica(idx)
                if self.budget is not None:
                    self.budget.release(self.footprint(q))
            else:
                keep.append(q)
        self.active = list(reversed(keep))

    # ------------------------------------------------------------ loop
    def serve(self, requests: list[Request], clock=time.perf_counter, log=None, t0: float | None = None,
              arbiter: Arbiter | None = None, lane: int = 0) -> dict:
        """Run until every request is finished. Arrivals are wall-clock
        offsets from t0 (default: now). Returns throughput and latency
        statistics. With an arbiter, each chunk is bracketed by its
        begin/end for lane `lane`."""
        pending = sorted(requests, key=lambda q: q.arrival)
        t0 = clock() if t0 is None else t0
        total_ps = 0
        chunks = 0
        used = 0
        while pending or self.active:
            now = clock() - t0
            # admit arrived requests while they fit (and, with max_active, while
            # fewer than that many run), in arrival order or the least work first
            batch = []
            used = sum(self.footprint(q) for q in self.active)
            arrived = [q for q in pending if q.arrival <= now]
            if self.sjf:
                arrived.sort(key=lambda q: (q.n * q.steps, q.arrival))
            for q in arrived:
                if self.max_active and len(self.active) + len(batch) >= self.max_active:
                    break
                f = self.footprint(q)
                alone = not (self.active or batch)
                if self.budget is not None:
                    # a lane with nothing active may always take one request
                    if not self.budget.reserve(f, force=alone):
                        break
                elif used + f > self.hbm_budget and not alone:
                    break
                used += f
                batch.append(q)
            for q in batch:
                pending.remove(q)
            if batch:
                a0 = clock()
                self.admit(batch, now)
                self.times["admit"] += clock() - a0
                if self.trace is not None:
                    self.trace.append(("admit", a0 - t0, clock() - t0, [q.n for q in batch]))
            if not self.active:
                # idle until the next arrival (or until another lane frees memory)
                wait = pending[0].arrival - (clock() - t0)
                time.sleep(wait if wait > 0 else 0.002)
                self.times["idle"] += max(wait, 0.002)
                continue
            # a chunk of about chunk_seconds (requests that finish inside it stop
            # at their step count on the device), at most until the last is done
            remaining = max(q.steps - self.done_steps[q.rid] for q in self.active)
            n_act = sum(q.n for q in self.active)
            if self.step_time:
                s = max(self.min_steps, int(self.chunk_seconds / (self.step_time * n_act)))
            else:
                s = self.min_steps
            s = min(s, remaining)
            if arbiter is not None:
                arbiter.begin(lane)
            c0 = clock()
            redone = self.ctx.chunks_redone()
            try:
                # the clock argument only matters for a single replica (several keep theirs on the device)
                self.ctx.run(s, self.done_steps[self.active[0].rid] * self.dt, self.dt)
                self.ctx.synchronize()
            finally:
                if arbiter is not None:
                    arbiter.end(lane)
            el = clock() - c0
            self.times["run"] += el
            self.stats["chunks_redone"] += self.ctx.chunks_redone() - redone
            if self.trace is not None:
                self.trace.append(("chunk", c0 - t0, c0 - t0 + el, s, n_act, len(self.active)))
            if el > self.stats["slowest_chunk_s"]:
                self.stats.update(slowest_chunk_s=el, slowest_chunk_steps=s, slowest_chunk_particles=n_act)
            self.step_time = el / (s * n_act)
            # particle-steps actually taken (finished requests stop on the device)
            total_ps += sum(q.n * min(s, q.steps - self.done_steps[q.rid]) for q in self.active)
            chunks += 1
            for q in self.active:
                self.done_steps[q.rid] = min(q.steps, self.done_steps[q.rid] + s)
                q.chunks += 1
                q.run_s += el
            r0 = clock()
            self.retire(clock() - t0)
            self.times["retire"] += clock() - r0
            if self.trace is not None and clock() - r0 > 0.001:
                self.trace.append(("retire", r0 - t0, clock() - t0))
            if log:
                log(f"t={clock() - t0:8.3f}s chunk {chunks}: {s} steps x {len(self.active)} left, "
                    f"{n_act} particles, {el / s * 1e3:.2f} ms/step, queue {len(pending)}")
        wall = clock() - t0
        out = summary(requests, wall, total_ps, chunks)
        out["seconds"] = dict(self.times)
        out["stats"] = dict(self.stats)
        return out


def summary(requests: list[Request], wall: float, particle_steps: int, chunks: int) -> dict:
    lat = np.array([q.latency for q in requests]) if requests else np.zeros(1)
    return {"requests": len(requests), "wall_s": wall, "particle_steps": particle_steps,
            "throughput": particle_steps / wall if wall > 0 else 0.0, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "chunks": chunks}


class LaneServer:
    """Requests routed by size to lanes (one device context, HIP stream and
    host thread each), all sharing one HBM budget. Within a lane requests are
    batched continuously; lanes run concurrently on the device, so a small
    request does not advance at the step rate of the largest ones. Chunks are
    shorter in lanes of smaller requests (admission happens at chunk
    boundaries)."""

    def __init__(self, precision: str = "fp64", hbm_budget: float = 64 * 2 ** 30,
                 bounds: tuple = (6e3, 4e4), chunk_seconds: tuple = (0.01, 0.04, 0.2),
                 cu_fractions: tuple | None = None,
                 cu_spans: tuple | None = ((0.0, 0.125), (0.125, 0.375), (0.375, 1.0)), device: int = 0,
                 arbiter_hold: float = 0.0, max_active: tuple = (0, 0, 3), sjf: tuple = (False, False, True),
                 overlap_lanes: tuple | None = None, **kw):
        """cu_fractions: share of the device's compute units given to each
        lane (disjoint ranges of CU ids); cu_spans: per lane a (first, last)
        fraction of the CU id range, which may overlap (a small-request lane
        can keep a slice to itself while a larger lane shares the rest);
        None lets every lane use all. The default spans give the three lanes
        1/8, 1/4 and 5/8 of the CUs to themselves: a small system's ~45
        dependent launches per step otherwise each wait for CU slots behind
        the long-running blocks of the large systems' pair kernels (measured
        7-50x slower steps), at the cost of the saturated throughput.
        max_active, sjf: per lane, at most that many requests at once (0: no
        limit) and admission of the least work (N x steps) first instead of
        in arrival order. By default the large-request lane runs at most
        three, the least work first: its requests each fill the device, so
        batching more of them adds little throughput and slows every one of
        them (measured on the C5 stream at 80 % load: p99 latency over
        isolated 16.5x -> 10.1x, p50 3.4x -> 2.3x, saturated throughput
        2.53e7 -> 2.28e7 particle-steps/s)."""
        assert len(chunk_seconds) == len(bounds) + 1
        self.bounds = bounds
        # arbiter_hold > 0: a lane holds its chunks while a lane of smaller
        # requests runs one (at most that many seconds per chunk)
        self.arbiter = Arbiter(len(bounds) + 1, arbiter_hold) if arbiter_hold > 0 else None
        self.budget = Budget(hbm_budget)
        # each lane reserves room for a few of its largest requests, so
        # admissions and retirements reuse device memory
        tops = list(bounds) + [2e5]
        # one stream per lane by default (overlap_lanes): the lanes already
        # run side by side, and helper streams in every lane would exceed the
        # four hardware queues a process gets by default, so that streams of
        # different lanes would share a queue and wait on each other
        overlap = kw.pop("kspace_overlap", None)
        if overlap_lanes is None:
            overlap_lanes = (bool(overlap),) * (len(bounds) + 1)
        assert len(max_active) == len(self.bounds) + 1 and len(sjf) == len(self.bounds) + 1
        assert len(overlap_lanes) == len(self.bounds) + 1
        self.lanes = [Server(precision=precision, hbm_budget=hbm_budget, chunk_seconds=cs, budget=self.budget,
                             reserve_particles=int(4 * top), device=device, max_active=int(ma), sjf=bool(sj),
                             kspace_overlap=bool(ov), **kw)
                      for cs, top, ma, sj, ov in zip(chunk_seconds, tops, max_active, sjf, overlap_lanes)]
        self.cu_ranges = None
        if cu_fractions:
            assert len(cu_fractions) == len(self.lanes)
            total = capi.lib().mds_device_cus(device)
            edges = np.round(np.cumsum([0.0] + list(cu_fractions)) / sum(cu_fractions) * total).astype(int)
            self.cu_ranges = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:])]
        elif cu_spans:
            assert len(cu_spans) == len(self.lanes)
            total = capi.lib().mds_device_cus(device)
            self.cu_ranges = [(int(round(a * total)), int(round(b * total))) for a, b in cu_spans]
        if self.cu_ranges:
            for lane, (a, b) in zip(self.lanes, self.cu_ranges):
                lane.ctx.set_cu_mask(list(range(a, max(b, a + 1))))

    def lane_of(self, req: Request) -> int:
        k = 0
        while k < len(self.bounds) and req.n >= self.bounds[k]:
            k += 1
        return k

    def close(self):
        for s in self.lanes:
            s.close()

    def serve(self, requests: list[Request], clock=time.perf_counter, log=None) -> dict:
        import threading
        parts = [[] for _ in self.lanes]
        for q in requests:
            parts[self.lane_of(q)].append(q)
        out = [None] * len(self.lanes)
        errs = []
        t0 = clock()

        def work(k):
            try:
                out[k] = (self.lanes[k].serve(parts[k], clock, log, t0, self.arbiter, k) if parts[k]
                          else summary([], 0.0, 0, 0))
            except Exception as e:  # surfaced below
                errs.append(e)

        th = [threading.Thread(target=work, args=(k,)) for k in range(len(self.lanes))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errs:
            raise errs[0]
        wall = clock() - t0
        res = summary(requests, wall, sum(o["particle_steps"] for o in out), sum(o["chunks"] for o in out))
        res["lanes"] = [{"requests": len(p), "chunks": o["chunks"], "seconds": o.get("seconds"), "stats": o.get("stats"),
                         "held_s": self.arbiter.held[k] if self.arbiter else 0.0}
                        for k, (p, o) in enumerate(zip(parts, out))]
        return res


def isolated_latencies(requests: list[Request], precision: str = "fp64", progress=None, **kw) -> list[float]:
    """Latency of each request served alone: arriving at an idle server
    (context already created) and leaving when its state is copied out, so
    admission (initial suspension, planning, graph capture) and retirement
    count as they do in the loop. The requests are solved one after another
    on copies; the originals are not modified. progress(msg), if given, is
    called after each request."""
    import copy
    s = Server(precision=precision, **kw)
    out = []
    try:
        for k, q in enumerate(requests):
            a = copy.deepcopy(q)
            a.arrival, a.state, a.ewald = 0.0, None, {}
            s.serve([a])
            out.append(a.latency)
            if progress:
                progress(f"isolated {k + 1}/{len(requests)}: N={q.n} steps={q.steps} {a.latency:.3f} s")
    finally:
        s.close()
    return out


def poisson_stream(n_requests: int, rate: float, rng: np.random.Generator, n_range=(1e3, 2e5),
                   steps_range=(1e3, 1e5), field_range=(0.0, 0.1), phi_range=(0.01, 0.10)) -> list[Request]:
    """Requests with Poisson arrivals at `rate` per second, N and steps
    log-uniform, field and volume fraction uniform, per-request seeds."""
    t = 0.0
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        out.append(Request(k, n, steps, float(rng.uniform(*field_range)), float(rng.uniform(*phi_range)),
                           int(rng.integers(1, 2 ** 62)), t))
    return out
