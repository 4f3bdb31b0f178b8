"""The benchmark configurations that the other GPU tests do not reach at
their own sizes.

* C2: N = 1e4, lognormal cores (median 10 nm radius, sigma_ln 0.2) + 5 nm
  shell, phi = 0.05, 10 mT. The FP64 path against the CPU oracle on every
  particle (1e-10 relative, contact floors); the mixed path within 1e-5 of
  FP64 for field, force and torque, and after 20 graph-replayed Philox steps
  for the orientations.
* C4: 256 replicas in one context, N uniform in [500, 20000], phi uniform in
  [1 %, 10 %], fields on a 0-100 mT grid, own Philox seeds, in FP64 and
  mixed precision. Interaction forces sum to zero in every replica; after graph-replayed steps, eight
  sampled replicas (the smallest, the largest and six others) are bitwise
  equal to themselves run alone.
* C5: a serving mix that holds a request with N = 1e5 beside small and
  medium ones (size-class lanes): each result is bitwise equal to the same
  request solved alone.
"""
import copy
import os

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import serving, workload
from refcfg import rel_err

pytestmark = pytest.mark.gpu


def floors(cfg, r, rh):
    mu0 = 4e-7 * cfg.my_pi
    m = float(cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r.min() ** 3)
    d = 2 * float(rh.min())
    return 3 * mu0 * m * m / (2 * cfg.my_pi * d ** 4), mu0 * m / (2 * cfg.my_pi * d ** 3)


@pytest.fixture(scope="module")
def c2():
    print("C2 system, N = 1e4", flush=True)
    cfg, r, rh, L, pos, mm, ea = workload.system("c2", seed=3)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.3)
    return dict(cfg=cfg, r=r, rh=rh, L=L, pos=pos, mm=mm, ea=ea, mu=mu, alpha=alpha, rc=rc, nc=nc)


def c2_context(s, precision):
    cfg = mds.config_from_string(workload.config_text("c2", seed=3, precision=precision))
    ctx = mds.Context(cfg)
    ctx.set_particles(s["r"], s["rh"], s["L"])
    ctx.set_ewald(s["alpha"], s["rc"], s["nc"])
    ctx.set_state(s["pos"], s["mm"], s["ea"])
    return ctx


def test_c2_fp64_vs_oracle_and_mixed_vs_fp64(c2):
    import oracle
    from refcfg import derived, parse_ini
    ctx = c2_context(c2, "fp64")
    ctx.set_noise(None, None, None)
    F, B, P, X = ctx.comp_interactions()
    ctx.close()
    cfg, r, rh = c2["cfg"], c2["r"], c2["rh"]
    c = derived(parse_ini(workload.config_text("c2", seed=3)))
    props = {"r_mag": r, "r_hydr": rh, "vol_mag": 4 / 3 * cfg.my_pi * r ** 3, "mag_mom": c2["mu"],
             "zeta_rot": 8 * cfg.my_pi * cfg.vis * rh ** 3, "zeta_trans": 6 * cfg.my_pi * cfg.vis * rh}
    s = oracle.System(len(r), c2["L"], c2["alpha"], c2["rc"], 2 * r.max(), c2["nc"], props, c,
                      threads=max(1, min(16, os.cpu_count() or 1)))
    Fo, Bo, Po, Xo = s.comp_interactions(c2["pos"], c2["mm"])
    f_fl, b_fl = floors(cfg, r, rh)
    assert rel_err(X, Xo) < 1e-15
    assert rel_err(F, Fo, f_fl) < 1e-10
    assert rel_err(B, Bo, b_fl) < 1e-10
    assert rel_err(P, Po, f_fl * 2 * float(rh.min())) < 1e-9
    mx = c2_context(c2, "mixed")
    mx.set_noise(None, None, None)
    Fm, Bm, _, _ = mx.comp_interactions()
    mx.close()
    assert rel_err(Fm, F, f_fl) < 1e-5
    assert rel_err(Bm, B, b_fl) < 1e-5
    assert rel_err(np.cross(c2["mm"], Bm), np.cross(c2["mm"], B), b_fl) < 1e-5


def test_c2_mixed_run_tracks_fp64(c2):
    out = {}
    for prec in ("fp64", "mixed"):
        ctx = c2_context(c2, prec)
        ctx.use_philox(77, 0)
        ctx.set_graph(True)
        _, _, mz = ctx.run(20, 0.0, 1e-12, want_mz=True)
        out[prec] = (ctx.get_state(), mz)
        ctx.close()
    (p64, m64, e64), mz64 = out["fp64"]
    (pmx, mmx, emx), mzmx = out["mixed"]
    assert rel_err(mmx, m64) < 1e-5
    assert rel_err(emx, e64) < 1e-5
    assert rel_err(pmx, p64) < 1e-9
    assert np.max(np.abs(mzmx - mz64)) < 1e-5


def c4_systems(cfg, replicas=256, master=20261203):
    """The ensemble: N ~ U[500, 20000], phi ~ U[1 %, 10 %], fields on a
    0-100 mT grid, monodisperse 10 nm cores + 5 nm shell, own seeds."""
    rng = np.random.default_rng(master)
    ns = rng.integers(500, 20001, replicas)
    phis = rng.uniform(0.01, 0.10, replicas)
    fields = np.linspace(0.0, 0.1, replicas)
    out = []
    for k in range(replicas):
        n = int(ns[k])
        r = np.full(n, cfg.r_mag_mean)
        rh = r + cfg.d_shell
        L = (np.sum(4.0 / 3.0 * cfg.my_pi * rh ** 3) / phis[k]) ** (1.0 / 3.0)
        sd = (master * 1000003 + 7919 * k) & 0xFFFFFFFFFFFFFFFF
        pos = mds.init_positions(cfg, rh, L, sd, periodic=True)
        mm, ea = mds.init_orientations(cfg, n, sd ^ 0x5BD1E995)
        mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
        # the autotuner's law: r_c / L scaled by (N_mean / N)^(1/6)
        rcf = min(0.5, max(0.1, 0.3 * (10250.0 / n) ** (1.0 / 6.0)))
        ew = workload.ewald_for(cfg, mu, rh, L, rcf)
        out.append(dict(r_mag=r, r_hydr=rh, l_box=L, field_z=float(fields[k]), seed=sd, pos=pos, mm=mm, ea=ea,
                        ewald=ew))
    return out


@pytest.mark.parametrize("precision,sum_tol", [("fp64", 1e-9), ("mixed", 1e-6)])
def test_c4_ensemble_of_256_replicas(precision, sum_tol):
    """Interaction forces sum to zero per replica: each pair's real-space
    force is added as +F and -F, so the residual is the k-space rounding:
    FP64 sums (1e-9 of the summed magnitudes) or the mixed path's FP32
    matrix-core products (1e-6)."""
    cfg = mds.config_from_string(workload.config_text("c1", seed=1, precision=precision))
    print(f"C4: 256 replicas, {precision}", flush=True)
    systems = c4_systems(cfg)
    ctx = mds.Context(cfg)
    ctx.add_replicas(systems)
    for k, s in enumerate(systems):
        ctx.replica_set_ewald(k, *s["ewald"])
    ctx.set_noise(None, None, None)
    F, _, _, _ = ctx.comp_interactions()
    off = 0
    for k, s in enumerate(systems):
        n = len(s["r_mag"])
        Fr = F[off:off + n]
        assert np.all(np.abs(Fr.sum(axis=0)) <= sum_tol * np.abs(Fr).sum(axis=0)), k
        off += n
    print("C4: graph-replayed steps", flush=True)
    ctx.use_philox(0, 0)
    ctx.set_graph(True)
    steps = 3
    ctx.run(steps, 0.0, 1e-12)
    ns = np.array([len(s["r_mag"]) for s in systems])
    rng = np.random.default_rng(5)
    sample = sorted({int(ns.argmin()), int(ns.argmax())} | set(int(x) for x in rng.choice(256, 6, replace=False)))
    for k in sample:
        s = systems[k]
        c = mds.config_from_string(workload.config_text("c1", seed=1, precision=precision))
        c.mag_flux_dens = s["field_z"]
        a = mds.Context(c)
        a.set_particles(s["r_mag"], s["r_hydr"], s["l_box"])
        a.set_ewald(*s["ewald"])
        a.set_state(s["pos"], s["mm"], s["ea"])
        a.use_philox(s["seed"], 0)
        a.run(steps, 0.0, 1e-12)
        for x, y in zip(ctx.replica_state(k), a.get_state()):
            assert np.array_equal(x, y), (k, len(s["r_mag"]))
        a.close()
    ctx.close()


def test_c5_serving_mix_with_a_large_request():
    print("C5: serving mix with N = 1e5", flush=True)
    # (N, steps, field T, phi, seed, arrival s)
    spec = [(100000, 3, 0.05, 0.10, 501, 0.0), (1200, 30, 0.02, 0.03, 502, 0.0), (3000, 20, 0.08, 0.05, 503, 0.01),
            (20000, 8, 0.0, 0.07, 504, 0.02)]
    reqs = [serving.Request(k, n, st, f, phi, sd, arr) for k, (n, st, f, phi, sd, arr) in enumerate(spec)]
    ls = serving.LaneServer(precision="mixed", chunk_seconds=(0.0, 0.0, 0.0))
    stats = ls.serve(reqs)
    ls.close()
    assert [l["requests"] for l in stats["lanes"]] == [2, 1, 1]
    for q in reqs:
        alone = copy.deepcopy(q)
        alone.state = None
        alone.ewald = {}
        solo = serving.Server(precision="mixed", chunk_seconds=10.0)
        solo.serve([alone])
        solo.close()
        for a, b in zip(q.state, alone.state):
            assert np.array_equal(a, b), (q.rid, q.n, rel_err(a, b))
