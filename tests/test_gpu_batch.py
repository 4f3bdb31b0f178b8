"""Replicas in one context (batched ensembles): every result of a replica in
a batch equals the same system run alone in its own context, including
Philox noise, graph-replayed steps, and replicas admitted and retired
between runs. The real-space sums are fixed-point integers added by atomics
and every other sum has a fixed order, so states agree bit for bit
(`test_batch_state_bitwise_equal_alone`); the other checks keep a 1e-12
relative tolerance."""
import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import workload
from refcfg import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12

# (N, volume fraction, r_c / L, overlapping pairs planted, field T, seed)
SYSTEMS = [(600, 0.02, 0.3, 0, 0.02, 101), (1500, 0.08, 0.22, 3, 0.0, 202), (900, 0.05, 0.25, 1, 0.08, 303)]


def make_system(n, phi, rc_frac, overlaps, seed, precision="fp64"):
    text = workload.config_text("c2", seed=seed, noise=True, fixed=True, precision=precision, n=n).replace(
        "volFrac = 0.05", f"volFrac = {phi}")
    cfg = mds.config_from_string(text)
    rng = np.random.default_rng(seed)
    r, rh = workload.radii(cfg, n, rng)
    L = workload.box_length(cfg, rh)
    pos = workload.place(rh, L, rng)
    for k in range(overlaps):
        a, b = 2 * k, 2 * k + 1
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        pos[b] = pos[a] + u * (r[a] + r[b]) * 0.9
    pos = np.where(pos > L / 2, pos - L, pos)
    pos = np.where(pos < -L / 2, pos + L, pos)
    ea = workload.orientations(n, rng, cfg.my_pi)
    mu = cfg.sat_mag * 4 / 3 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, rc_frac)
    return dict(cfg=cfg, r=r, rh=rh, L=L, pos=pos, ea=ea, alpha=alpha, rc=rc, nc=nc)


def alone(s, field, seed=None):
    cfg = s["cfg"]
    cfg.mag_flux_dens = field
    ctx = mds.Context(cfg)
    ctx.set_particles(s["r"], s["rh"], s["L"])
    ctx.set_ewald(s["alpha"], s["rc"], s["nc"])
    ctx.set_state(s["pos"], s["ea"], s["ea"])
    if seed is None:
        ctx.set_noise(None, None, None)
    else:
        ctx.use_philox(seed, 0)
    return ctx


def batch(systems, fields, seeds, precision="fp64"):
    ctx = mds.Context(systems[0]["cfg"])
    for s, f, sd in zip(systems, fields, seeds):
        i = ctx.add_replica(s["r"], s["rh"], s["L"], f, sd if sd is not None else 0, s["pos"], s["ea"], s["ea"])
        ctx.replica_set_ewald(i, s["alpha"], s["rc"], s["nc"])
    return ctx


def test_comp_interactions_batch_equals_alone():
    print("batch comp_interactions", flush=True)
    sys_ = [make_system(n, phi, rcf, ov, sd) for n, phi, rcf, ov, _, sd in SYSTEMS]
    fields = [f for *_, f, _ in SYSTEMS]
    ctx = batch(sys_, fields, [None] * 3)
    ctx.set_noise(None, None, None)
    F, B, P, X = ctx.comp_interactions()
    off = 0
    Psum = np.zeros(3)
    for s, f in zip(sys_, fields):
        n = len(s["r"])
        Fa, Ba, Pa, Xa = alone(s, f).comp_interactions()
        assert rel_err(F[off:off + n], Fa) < TOL
        assert rel_err(B[off:off + n], Ba) < TOL
        assert rel_err(X[off:off + n], Xa) < 1e-15
        Psum += Pa
        off += n
    assert rel_err(P, Psum) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_philox_graph_run_batch_equals_alone(precision):
    """Three replicas with their own seeds, fields and Ewald parameters, six
    graph-replayed steps, against each replica run alone without a graph."""
    print(f"batch run {precision}", flush=True)
    sys_ = [make_system(n, phi, rcf, ov, sd, precision) for n, phi, rcf, ov, _, sd in SYSTEMS]
    fields = [f for *_, f, _ in SYSTEMS]
    seeds = [7 + 1000 * k for k in range(3)]
    ctx = batch(sys_, fields, seeds)
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    steps = 6
    t, dt, mz = ctx.run(steps, 0.0, 1e-12, want_mz=True)
    assert mz.shape == (steps, 3)
    assert t == pytest.approx(steps * 1e-12)
    for k, (s, f, sd) in enumerate(zip(sys_, fields, seeds)):
        a = alone(s, f, sd)
        ta, _, mza = a.run(steps, 0.0, 1e-12, want_mz=True)
        pa, ma, ea = a.get_state()
        pb, mb, eb = ctx.replica_state(k)
        assert rel_err(mb, ma) < TOL, k
        assert rel_err(eb, ea) < TOL, k
        assert rel_err(pb, pa) < 1e-14, k
        assert np.allclose(mz[:, k], mza, rtol=1e-11, atol=1e-14)
        info = ctx.replica_info(k, clock=True)
        assert info["steps"] == steps and info["t"] == pytest.approx(ta)


def test_admit_and_retire_between_runs():
    """Continuous batching: A and B run 4 steps, C joins and all run 4, A
    leaves and B, C run 4 more. B must equal 12 steps alone, C 8 steps
    alone (each replica keeps its own clock and Philox counter)."""
    print("batch admit/retire", flush=True)
    sA, sB, sC = [make_system(n, phi, rcf, ov, sd) for n, phi, rcf, ov, _, sd in SYSTEMS]
    fA, fB, fC = [f for *_, f, _ in SYSTEMS]
    ctx = batch([sA, sB], [fA, fB], [11, 22])
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    ctx.run(4, 0.0, 1e-12)
    i = ctx.add_replica(sC["r"], sC["rh"], sC["L"], fC, 33, sC["pos"], sC["ea"], sC["ea"])
    ctx.replica_set_ewald(i, sC["alpha"], sC["rc"], sC["nc"])
    assert i == 2
    ctx.run(4, 0.0, 1e-12)
    ctx.remove_replica(0)
    assert ctx.replica_count() == 2
    ctx.run(4, 0.0, 1e-12)
    for k, (s, f, sd, steps) in enumerate([(sB, fB, 22, 12), (sC, fC, 33, 8)]):
        a = alone(s, f, sd)
        a.run(steps, 0.0, 1e-12)
        pa, ma, ea = a.get_state()
        pb, mb, eb = ctx.replica_state(k)
        assert rel_err(mb, ma) < TOL, k
        assert rel_err(eb, ea) < TOL, k
        assert rel_err(pb, pa) < 1e-14, k
        assert ctx.replica_info(k, clock=True)["steps"] == steps


def test_batch_autotune_keeps_accuracy_model():
    """The autotuner sets every replica's alpha and k_c from its own error
    model for its r_c / L (the chosen candidate scaled by (N_mean / N)^(1/6))."""
    print("batch autotune", flush=True)
    sys_ = [make_system(n, phi, rcf, ov, sd) for n, phi, rcf, ov, _, sd in SYSTEMS]
    ctx = batch(sys_, [0.0] * 3, [1, 2, 3])
    res = ctx.autotune(rc_fracs=(0.25, 0.3), reaches=(0, 2, 3), reps=1)
    n_mean = np.mean([len(s["r"]) for s in sys_])
    for k, s in enumerate(sys_):
        info = ctx.replica_info(k)
        mu = s["cfg"].sat_mag * 4 / 3 * s["cfg"].my_pi * s["r"] ** 3
        # r_c / L of the winning candidate, scaled by the replica's size
        f = min(0.5, max(0.1, res["rc_frac"] * (n_mean / len(s["r"])) ** (1 / 6)))
        alpha, rc, nc = workload.ewald_for(s["cfg"], mu, s["rh"], s["L"], f)
        assert info["alpha"] == pytest.approx(alpha, rel=1e-12)
        assert info["nc"] == nc
        assert info["r_cut"] == pytest.approx(rc, rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_state_bitwise_equal_alone(precision):
    """Positions, moments and easy axes after graph-replayed Philox steps of
    three replicas (two with planted overlaps) are bitwise equal to each
    replica run alone: no sum depends on co-scheduling."""
    print(f"batch bitwise {precision}", flush=True)
    sys_ = [make_system(n, phi, rcf, ov, sd, precision) for n, phi, rcf, ov, _, sd in SYSTEMS]
    fields = [f for *_, f, _ in SYSTEMS]
    seeds = [9 + 77 * k for k in range(3)]
    ctx = batch(sys_, fields, seeds)
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    ctx.run(5, 0.0, 1e-12)
    for k, (s, f, sd) in enumerate(zip(sys_, fields, seeds)):
        a = alone(s, f, sd)
        a.run(5, 0.0, 1e-12)
        for x, y in zip(ctx.replica_state(k), a.get_state()):
            assert np.array_equal(x, y), (k, float(np.max(np.abs(x - y))))


def test_mixed_kspace_path_is_chosen_per_replica():
    """The mixed path puts a replica's S(k) build on the matrix cores when its
    n_c <= 50 and its scatter when n_c <= 32, each replica by its own n_c:
    replicas with n_c = 58 (both on the vector kernels), 39 (matrix-core
    build, vector scatter) and <= 22 (both on the matrix cores) in one
    context give the results each gives alone, bit for bit."""
    print("mixed k-space path per replica", flush=True)
    spec = [(600, 0.02, 0.1, 0, 0.03, 41), (800, 0.05, 0.15, 0, 0.0, 42), (700, 0.03, 0.25, 0, 0.06, 43)]
    sys_ = [make_system(n, phi, rcf, ov, sd, "mixed") for n, phi, rcf, ov, _, sd in spec]
    assert [s["nc"] > 50 for s in sys_] == [True, False, False] and 32 < sys_[1]["nc"] <= 50 and sys_[2]["nc"] <= 32
    fields = [f for *_, f, _ in spec]
    seeds = [5, 6, 7]
    ctx = batch(sys_, fields, seeds)
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    ctx.run(3, 0.0, 1e-12)
    for k, (s, f, sd) in enumerate(zip(sys_, fields, seeds)):
        a = alone(s, f, sd)
        a.run(3, 0.0, 1e-12)
        for x, y in zip(ctx.replica_state(k), a.get_state()):
            assert np.array_equal(x, y), k


def test_replica_stops_at_its_step_count():
    """A replica told to stop at step 3 keeps its 3-step state (and clock and
    Philox counter) while the context runs on: after 7 graph-replayed steps
    it equals itself run alone for 3 steps, bit for bit, and the other
    replica equals itself run alone for 7."""
    print("batch stop_at", flush=True)
    sys_ = [make_system(n, phi, rcf, ov, sd) for n, phi, rcf, ov, _, sd in SYSTEMS[:2]]
    fields = [f for *_, f, _ in SYSTEMS[:2]]
    seeds = [31, 32]
    ctx = batch(sys_, fields, seeds)
    ctx.use_philox(0, 0)
    ctx.replica_stop_at(0, 3)
    ctx.set_graph(True)
    ctx.run(7, 0.0, 1e-12)
    for k, steps in ((0, 3), (1, 7)):
        a = alone(sys_[k], fields[k], seeds[k])
        a.run(steps, 0.0, 1e-12)
        for x, y in zip(ctx.replica_state(k), a.get_state()):
            assert np.array_equal(x, y), k
        info = ctx.replica_info(k, clock=True)
        assert info["steps"] == steps and info["t"] == pytest.approx(steps * 1e-12)
