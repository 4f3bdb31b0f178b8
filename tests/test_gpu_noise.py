"""The in-kernel thermal noise (Philox4x32-10 + Box-Muller, DefThermFluct's
role, src/simulation.cpp:563-603) on the device.

* At the C3 size (N = 1e5, lognormal radii, three steps, 9e5 draws per
  channel): every draw equals the numpy restatement of the generator
  (oracle/noise.py, itself pinned by the published known-answer vectors)
  times the reference's scale (sigma_T, sigma_B, sigma_F of
  src/simulation.cpp:567, 582, 594); each channel divided by its scale has
  mean 0 and variance 1 within 5 standard errors; channels, neighbouring
  particles and consecutive steps are uncorrelated within 5/sqrt(n).
* Free particles (moments too weak to interact, no field, no anisotropy):
  the mean squared displacement grows as 6 kB T t / zeta_t (the Einstein
  relation the reference's mFiles/eval_diffusion.m:19 fits), within 3 %.
* Non-interacting moments in a field (no anisotropy, Gilbert damping 1):
  the time-averaged <m_z> reaches the Langevin function L(mu B / kB T),
  within 1 %.
"""
import re

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import workload
from oracle import noise

pytestmark = pytest.mark.gpu
SEED = 5
NOISE_SEED = 0x5EED0000C3


def edit_ini(text: str, **keys) -> str:
    for k, v in keys.items():
        text, n = re.subn(rf"(?m)^{k}\s*=.*$", f"{k} = {v}", text)
        assert n == 1, k
    return text


def props(cfg, r, rh):
    """InitParticleProps (src/simulation.cpp:197-202): V_m, zeta_rot, zeta_trans."""
    pi = cfg.my_pi
    return 4.0 / 3.0 * pi * r ** 3, 8.0 * pi * cfg.vis * rh ** 3, 6.0 * pi * cfg.vis * rh


def test_c3_draws_equal_restatement_and_have_the_reference_sigmas():
    print("C3 noise draws, N = 1e5", flush=True)
    cfg, r, rh, L, pos, mm, ea = workload.system("c3", seed=SEED)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.225)
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(alpha, rc, nc)
    ctx.set_state(pos, mm, ea)
    ctx.use_philox(NOISE_SEED, 0)
    dt = 1e-12
    n = len(r)
    vm, zr, zt = props(cfg, r, rh)
    s_t, s_b, s_f = noise.sigmas(cfg.kB, cfg.temp, dt, cfg.mag_damp, cfg.gyro_mr, cfg.sat_mag, zr, vm, zt)
    scaled = []
    t = 0.0
    for step in range(3):
        t, _, du, _ = ctx.step(t, dt)
        assert du == dt
        T, B, F = ctx.get_noise()
        z = noise.normals(NOISE_SEED, np.arange(n), step)
        for got, sig, c0 in ((T, s_t, 0), (B, s_b, 3), (F, s_f, 6)):
            want = z[:, c0:c0 + 3] * sig[:, None]
            assert np.all(np.abs(got - want) <= 1e-13 * sig[:, None] * (1.0 + np.abs(z[:, c0:c0 + 3]))), (step, c0)
        scaled.append(np.concatenate([T / s_t[:, None], B / s_b[:, None], F / s_f[:, None]], axis=1))
    ctx.close()
    zz = np.stack(scaled)  # (3 steps, n, 9 = torque xyz, field xyz, force xyz)
    m = 3 * 3 * n          # draws per channel (steps x axes x particles)
    for ch in range(3):
        v = zz[:, :, 3 * ch:3 * ch + 3].reshape(-1)
        assert abs(v.mean()) < 5 / np.sqrt(m), ch
        assert abs(v.var() - 1.0) < 5 * np.sqrt(2.0 / m), ch
    # correlations: between the 9 components, between neighbours, between steps
    lim = 5 / np.sqrt(3 * n)
    flat = zz.reshape(-1, 9)
    cc = np.corrcoef(flat.T)
    assert np.abs(cc[~np.eye(9, dtype=bool)]).max() < lim
    for k in range(9):
        a = zz[:, :, k]
        assert abs(np.corrcoef(a[:, :-1].reshape(-1), a[:, 1:].reshape(-1))[0, 1]) < 5 / np.sqrt(3 * (n - 1)), k
        assert abs(np.corrcoef(a[:-1].reshape(-1), a[1:].reshape(-1))[0, 1]) < 5 / np.sqrt(2 * n), k


def free_system(n, phi, seed):
    """Monodisperse C1-like particles with moments too weak to interact
    (M_s = 480 A/m: dipole couplings 1e-6 of magnetite's), no anisotropy."""
    text = workload.config_text("c1", seed=seed, n=n)
    text = edit_ini(text, satMag="480", anisEn="0", volFrac=str(phi), ENABLE_THERM_FIELD="false",
                    ENABLE_THERM_TORQUE="false")
    cfg = mds.config_from_string(text)
    rng = np.random.default_rng(seed)
    r, rh = workload.radii(cfg, n, rng)
    L = workload.box_length(cfg, rh)
    pos = workload.place(rh, L, rng)
    ea = workload.orientations(n, rng, cfg.my_pi)
    return cfg, r, rh, L, pos, ea


def test_free_particle_msd_follows_einstein():
    n, steps, dt = 50_000, 200, 1e-10
    print(f"free particles N={n}, {steps} steps of {dt}", flush=True)
    cfg, r, rh, L, pos, ea = free_system(n, 0.01, 11)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.25)
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(alpha, rc, nc)
    ctx.set_state(pos, ea, ea)
    ctx.use_philox(77, 0)
    ctx.observe(True)
    _, _, _, obs = ctx.run_observed(steps, 0.0, dt)
    p_end, _, _ = ctx.get_state()
    ctx.close()
    zeta_t = 6.0 * cfg.my_pi * cfg.vis * rh[0]
    tt = dt * np.arange(1, steps + 1)
    msd = obs[:, 1] + obs[:, 2] + obs[:, 3]
    theory = 6.0 * cfg.kB * cfg.temp * tt / zeta_t
    # the observable against positions unwrapped here (displacements << L)
    d = p_end - pos
    d -= L * np.round(d / L)
    assert msd[-1] == pytest.approx(np.mean(np.sum(d * d, axis=1)), rel=1e-9)
    slope = np.sum(msd * tt) / np.sum(tt * tt)
    assert slope == pytest.approx(6.0 * cfg.kB * cfg.temp / zeta_t, rel=0.03)
    assert msd[-1] == pytest.approx(theory[-1], rel=0.03)
    # each axis alone: 2 kB T t / zeta_t
    for a in range(3):
        assert obs[-1, 1 + a] == pytest.approx(theory[-1] / 3, rel=0.05)


def test_noninteracting_magnetisation_is_langevin():
    n, steps, dt, field = 20_000, 2000, 1e-12, 0.1
    print(f"Langevin magnetisation N={n}, {steps} steps", flush=True)
    text = workload.config_text("c1", seed=3, n=n)
    text = edit_ini(text, satMag="4.8e4", anisEn="0", magDamp="1.0", volFrac="0.005", magFluxDens=str(field))
    cfg = mds.config_from_string(text)
    rng = np.random.default_rng(3)
    r, rh = workload.radii(cfg, n, rng)
    L = workload.box_length(cfg, rh)
    pos = workload.place(rh, L, rng)
    ea = workload.orientations(n, rng, cfg.my_pi)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    alpha, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.25)
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(alpha, rc, nc)
    ctx.set_state(pos, ea, ea)
    ctx.use_philox(2024, 0)
    ctx.set_graph(True)
    _, _, mz = ctx.run(steps, 0.0, dt, want_mz=True)
    ctx.close()
    xi = mu[0] * field / (cfg.kB * cfg.temp)
    langevin = 1.0 / np.tanh(xi) - 1.0 / xi
    got = float(np.mean(mz[steps // 2:]))
    print(f"xi={xi:.3f} L(xi)={langevin:.5f} <m_z>={got:.5f}", flush=True)
    assert got == pytest.approx(langevin, rel=0.01)
