"""Pair lists (mds_set_pair_list): both evaluations of a step read each
replica's list of the pairs within r_c (1 + skin) of its build positions
instead of searching the cells, and a replica rebuilds its list on the device
once one of its particles has moved half the skin from its build position.

Checks: a run with lists equals the run that searches the cells every
evaluation (skin 0) to rounding (the same pairs, added in another order)
while the lists are rebuilt several times along it; a list region too small
for the list is grown by the checked step that sees the overflow, with the
same results bit for bit; every build is counted; most steps keep every
list and then skip the main grid's sort, and a reservation that moves the
buffers in the middle of a run (the kept lists' streams among them) changes
no bit."""
import re

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import workload
from refcfg import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_small():
    print("C2 settings, N = 3000", flush=True)
    cfg, r, rh, L, pos, mm, ea = workload.system("c2", seed=5, n=3000)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.3)
    return dict(r=r, rh=rh, L=L, pos=pos, mm=mm, ea=ea, alpha=alpha, rc=rc, nc=nc)


def calm(text):
    """Weak moments without anisotropy or thermal torque and field: the
    particles diffuse as in C2, but the moments turn slowly even at the long
    steps used here, so two runs that differ only in rounding stay close."""
    for k, v in dict(satMag="480", anisEn="0", ENABLE_THERM_FIELD="false", ENABLE_THERM_TORQUE="false").items():
        text, n = re.subn(rf"(?m)^{k}\s*=.*$", f"{k} = {v}", text)
        assert n == 1, k
    return text


def make(s, precision, skin, scale=1.0, graph=False, quiet=False):
    text = workload.config_text("c2", seed=5, precision=precision, n=3000)
    cfg = mds.config_from_string(calm(text) if quiet else text)
    ctx = mds.Context(cfg)
    ctx.set_particles(s["r"], s["rh"], s["L"])
    ctx.set_ewald(s["alpha"], s["rc"], s["nc"])
    ctx.set_pair_list(skin, scale)
    ctx.set_state(s["pos"], s["mm"], s["ea"])
    ctx.use_philox(11, 0)
    ctx.set_graph(graph)
    return ctx


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-6)])
def test_lists_equal_the_cell_search_across_rebuilds(c2_small, precision, tol):
    # dt = 3e-10 s: a particle diffuses ~0.1 nm per step against a half skin
    # of 0.005 r_c / 2 ~ 0.6 nm, so the lists are rebuilt every few steps
    steps, dt = 40, 3e-10
    print(f"{precision}: {steps} graph-replayed steps of {dt} s, lists vs search", flush=True)
    a = make(c2_small, precision, 0.005, graph=True, quiet=True)
    b = make(c2_small, precision, 0.0, graph=True, quiet=True)
    _, _, mza = a.run(steps, 0.0, dt, want_mz=True)
    _, _, mzb = b.run(steps, 0.0, dt, want_mz=True)
    st = a.pair_list_stats(0)
    assert 3 <= st["builds"] < steps, st
    assert 0 < st["words"] <= st["capacity"], st
    assert b.pair_list_stats(0)["builds"] == 0
    for x, y in zip(a.get_state(), b.get_state()):
        assert rel_err(x, y) < tol
    assert np.allclose(mza, mzb, rtol=tol, atol=1e-14)
    # one evaluation of the final state: the same pair sums to rounding
    for ctx in (a, b):
        ctx.set_state(*a.get_state())
        ctx.set_noise(None, None, None)
    Fa, Ba, Pa, _ = a.comp_interactions()
    Fb, Bb, Pb, _ = b.comp_interactions()
    sums = 1e-12 if precision == "fp64" else 1e-6
    assert rel_err(Fa, Fb) < sums and rel_err(Ba, Bb) < sums and rel_err(Pa, Pb) < 1e3 * sums
    a.close()
    b.close()


def test_overflowing_region_grows_and_changes_nothing(c2_small):
    print("list region at 5 % of the estimate, eager steps and a graph-replayed run", flush=True)
    ref = make(c2_small, "fp64", 0.005)
    small = make(c2_small, "fp64", 0.005, scale=0.05)
    for k in range(3):
        t0 = k * 1e-12
        ref.step(t0, 1e-12)
        small.step(t0, 1e-12)
    st = small.pair_list_stats(0)
    assert st["words"] <= st["capacity"] and st["capacity"] >= 0.9 * st["words"], st
    for x, y in zip(small.get_state(), ref.get_state()):
        assert np.array_equal(x, y)
    # a fresh context whose first chunk step sees the overflow
    g_ref = make(c2_small, "mixed", 0.005, graph=True)
    g_small = make(c2_small, "mixed", 0.005, scale=0.05, graph=True)
    g_ref.run(12, 0.0, 1e-12)
    g_small.run(12, 0.0, 1e-12)
    for x, y in zip(g_small.get_state(), g_ref.get_state()):
        assert np.array_equal(x, y)
    for c in (ref, small, g_ref, g_small):
        c.close()


def test_kept_lists_survive_a_reservation_mid_run(c2_small):
    """Steps that keep every list reuse the last step's sort, order, work
    items and list streams; mds_reserve of twice the particles between two
    runs moves those buffers, which must re-plan and rebuild the lists from
    their build positions: the same bits as the run without it."""
    print("reserve between two graph-replayed runs", flush=True)
    a = make(c2_small, "mixed", 0.005, graph=True)
    b = make(c2_small, "mixed", 0.005, graph=True)
    ta, dta, _ = a.run(10, 0.0, 1e-12)
    tb, dtb, _ = b.run(10, 0.0, 1e-12)
    b.reserve(2 * len(c2_small["r"]))
    a.run(10, ta, dta)
    b.run(10, tb, dtb)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    assert b.pair_list_stats(0)["builds"] >= a.pair_list_stats(0)["builds"]
    a.close()
    b.close()
