"""Initial suspensions drawn on the device (mds_init_suspension_device):
InitParticleProps' radii and InitCoordsFromSketch's placement
(src/simulation.cpp:118-351) by parallel random sequential addition.

* The C5 maximum, N = 2e5 lognormal cores (sigma_ln 0.25) + 5 nm shells at
  phi = 0.1, periodic: no two hydrodynamic spheres overlap (every pair
  within 2 r_max found by a periodic k-d tree); ln r has the configured
  mean and variance, the easy axes are uniform on the sphere (component
  means, <e_z^2> = 1/3), the positions uniform in the box, each within 5
  standard errors; the box holds the configured volume fraction.
* N = 2e4: the CPU oracle's brute-force pair scan finds no overlap.
* The reference's own rule (not periodic): every particle inside
  [-(L/2 - r_i), L/2 - r_i)^3 and no overlap by plain distance.
* The same seed gives the same suspension; INIT_CONFIG_TRANS 2 the grid.
"""
import re
import time

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import capi, workload

pytestmark = pytest.mark.gpu


def overlaps_periodic(pos, rh, L):
    from scipy.spatial import cKDTree
    x = np.mod(pos + L / 2, L)
    x = np.where(x >= L, 0.0, x)
    tree = cKDTree(x, boxsize=L)
    pairs = tree.query_pairs(2 * rh.max(), output_type="ndarray")
    d = x[pairs[:, 0]] - x[pairs[:, 1]]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(axis=1))
    return int(np.sum(r < rh[pairs[:, 0]] + rh[pairs[:, 1]]))


def test_c5_maximum_suspension_on_device():
    n = 200_000
    cfg = mds.config_from_string(workload.config_text("c3", seed=1, n=n))
    print(f"device init N={n}", flush=True)
    capi.init_suspension_device(cfg, 1000, 1)  # the process's first device work pays the runtime start-up
    t0 = time.perf_counter()
    r, L, pos, mm, ea = capi.init_suspension_device(cfg, n, 12345)
    t_dev = time.perf_counter() - t0
    rh = r + cfg.d_shell
    assert L == pytest.approx(workload.box_length(cfg, rh), rel=1e-12)
    assert overlaps_periodic(pos, rh, L) == 0
    assert np.all(pos >= -L / 2) and np.all(pos < L / 2)
    # radii: ln(r / median) ~ N(0, sigma^2)
    z = np.log(r / cfg.r_mag_median)
    s = cfg.sigma_log
    assert abs(z.mean()) < 5 * s / np.sqrt(n)
    assert abs(z.var() - s * s) < 5 * s * s * np.sqrt(2.0 / n)
    # easy axes uniform on the sphere, moments on them (INIT_CONFIG_ROT 2)
    assert np.allclose(np.linalg.norm(ea, axis=1), 1.0, atol=1e-14)
    assert np.array_equal(mm, ea)
    assert np.all(np.abs(ea.mean(axis=0)) < 5 / np.sqrt(3 * n))
    assert abs(np.mean(ea[:, 2] ** 2) - 1 / 3) < 5 * np.sqrt(4 / 45 / n)
    # positions uniform in the box
    assert np.all(np.abs(pos.mean(axis=0)) < 5 * L / np.sqrt(12 * n))
    assert np.all(np.abs(pos.var(axis=0) - L * L / 12) < 5 * L * L * np.sqrt((1 / 80 - 1 / 144) / n))
    t0 = time.perf_counter()
    capi.init_positions(cfg, rh, L, 12345, periodic=True)
    t_host = time.perf_counter() - t0
    print(f"placement of {n}: device {t_dev:.3f} s (radii, box, positions, orientations), host {t_host:.3f} s",
          flush=True)


def test_oracle_finds_no_overlap():
    import oracle
    from refcfg import derived, parse_ini
    n = 20_000
    text = workload.config_text("c3", seed=2, n=n)
    cfg = mds.config_from_string(text)
    r, L, pos, mm, ea = capi.init_suspension_device(cfg, n, 777)
    rh = r + cfg.d_shell
    c = derived(parse_ini(text))
    # the oracle's overlap scan (CompInteractVecs, src/interaction.cpp:120)
    # counts r < r_i + r_j + 1e-10 m with the radii it is given: hydrodynamic
    # radii less 0.5e-10 m each here, so it counts r < rh_i + rh_j, the
    # placement rule, over every pair within 2 r_max
    rs = rh - 0.5e-10
    props = {"r_mag": rs, "r_hydr": rh, "vol_mag": 4 / 3 * cfg.my_pi * r ** 3, "mag_mom": r,
             "zeta_rot": rh, "zeta_trans": rh}
    s = oracle.System(n, L, 10.0 / L, 2.0 * rh.max() * 1.01, 2 * rh.max(), 2, props, c, threads=16)
    assert s.pairs(pos)["overlaps"] == 0


def test_reference_rule_without_periodic_images():
    n = 2000
    text = re.sub(r"(?m)^volFrac\s*=.*$", "volFrac = 0.05", workload.config_text("c2", seed=3, n=n))
    cfg = mds.config_from_string(text)
    r, L, pos, mm, ea = capi.init_suspension_device(cfg, n, 99, periodic=False)
    rh = r + cfg.d_shell
    assert np.all(np.abs(pos) <= (L / 2 - rh)[:, None])
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(axis=2))
    lim = rh[:, None] + rh[None, :]
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= lim)


def test_seeded_and_grid():
    cfg = mds.config_from_string(workload.config_text("c1", seed=1, n=3000))
    a = capi.init_suspension_device(cfg, 3000, 5)
    b = capi.init_suspension_device(cfg, 3000, 5)
    c = capi.init_suspension_device(cfg, 3000, 6)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[2], c[2])
    g = mds.config_from_string(workload.config_text("c1", seed=1).replace("INIT_CONFIG_TRANS = 1",
                                                                          "INIT_CONFIG_TRANS = 2"))
    r, L, pos, mm, ea = capi.init_suspension_device(g, 1000, 5)
    gap = L / 10
    assert np.allclose(np.sort(np.unique(np.round(pos[:, 0] / gap, 6))), (np.arange(10) + 0.5) - 5)
