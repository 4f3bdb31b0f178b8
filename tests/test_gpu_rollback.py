"""Graph-replayed runs check their flags after every chunk of unchecked
steps; a chunk whose steps raised the overlap-margin flag is rolled back and
redone eagerly with every step checked (the contact-search margin widened
where an overlap move needs it).

Free particles at phi = 0.1 with dt = 1e-7 s move ~2.4 nm per step along a
pair axis relative to each other, so cores that meet overlap by more than
twice the displacement bound (half the initial 2 nm margin) within a few
tens of steps, always at a replayed step (the cores start at least 10 nm
apart). The run must end normally, have redone at least
one chunk, and equal the run whose every step was checked (mds_step)."""
import re

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import workload
from refcfg import rel_err

pytestmark = pytest.mark.gpu


def edit_ini(text, **keys):
    for k, v in keys.items():
        text, n = re.subn(rf"(?m)^{k}\s*=.*$", f"{k} = {v}", text)
        assert n == 1, k
    return text


def make(seed=21, n=2000):
    text = edit_ini(workload.config_text("c1", seed=seed, n=n), satMag="480", anisEn="0", volFrac="0.1",
                    ENABLE_THERM_FIELD="false", ENABLE_THERM_TORQUE="false")
    cfg = mds.config_from_string(text)
    rng = np.random.default_rng(seed)
    r, rh = workload.radii(cfg, n, rng)
    L = workload.box_length(cfg, rh)
    pos = workload.place(rh, L, rng)
    ea = workload.orientations(n, rng, cfg.my_pi)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    ew = workload.ewald_for(cfg, mu, rh, L, 0.3)
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(*ew)
    ctx.set_state(pos, ea, ea)
    ctx.use_philox(9, 0)
    return ctx


def test_flagged_chunk_is_rolled_back_and_redone():
    steps, dt = 60, 1e-7
    print(f"rollback: {steps} graph-replayed steps of {dt} s", flush=True)
    ctx = make()
    ctx.set_graph(True)
    ctx.set_check_interval(20)
    t, _, mz = ctx.run(steps, 0.0, dt, want_mz=True)
    assert ctx.chunks_redone() >= 1
    assert t == pytest.approx(steps * dt)
    ref = make()
    tr, mzr = 0.0, []
    for _ in range(steps):
        tr, _, _, m = ref.step(tr, dt)
        mzr.append(m)
    for a, b in zip(ctx.get_state(), ref.get_state()):
        assert rel_err(a, b) < 1e-9
    assert np.allclose(mz, mzr, rtol=1e-9, atol=1e-12)
    assert ctx.replica_info(0, clock=True)["steps"] == steps
