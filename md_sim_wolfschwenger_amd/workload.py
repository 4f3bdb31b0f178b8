"""Synthetic suspensions of the benchmark configurations.

Seeded numpy generators for what InitParticleProps / InitCoordsFromSketch
(src/simulation.cpp:118-351) set up: core radii (equal or lognormal),
hydrodynamic radii (core + shell), uniform random non-overlapping positions
in the periodic box (hydrodynamic radii), easy axes uniform on the sphere
with the reference's sampler (azimuth 2 pi u, elevation 2 asin sqrt(u),
src/simulation.cpp:304-312) and moments along the easy axes
(INIT_CONFIG_ROT = 2). The reference's own EigenRand streams are not
reproduced; runs that must match it take its golden state instead.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi

# Key text of the configurations (sections/keys of configuration.ini plus the
# [mi355x] keys this library adds).
_BASE = """
[initial_configuration]
INIT_CONFIG_ROT = 2
INIT_CONFIG_TRANS = 1
READ_FROM_FILE = false
[size_distribution]
SIZE_DIST = {size_dist}
[mobilization]
ENABLE_MOBILIZATION = true
[thermal_fluctuations]
ENABLE_THERM_FIELD = {noise}
ENABLE_THERM_TORQUE = {noise}
ENABLE_THERM_FORCE = {noise}
seed = {seed}
[interactions]
ENABLE_COATING_POT = false
[adaptive_timestepping_solver]
SOLVER = 0
[particle_properties]
npDim = 10
magDamp = {alpha_g}
temp = 300.0
satMag = 4.8e5
anisEn = 1e4
[assemble_properties]
rMagMean = 1e-8
dShell = 5e-9
volFrac = {phi}
[steric_repulsion]
densSurfMol = 5e17
[vdw_attraction]
hamaker = 33e-21
[electrostat_repulsion]
z_valency = 1
epsilon_r = 78.5
gamma_0 = 0.04
conc = 50
[time_settings]
tMag = 1.0
tRelax = {t_relax}
deltaTinit = 1e-12
[external_magnetic_field_t]
magFluxDens = {field}
[solver_settings]
absTol = 1e-4
deltaTmin = 1e-14
deltaTmax = 1e-9
[long_range_interactions]
errTolEwald = 1e-6
[short_range_interactions]
errTolSR = 0.01
[output_settings]
npCoords = 100
[physical_constants]
my_pi = 3.14159265359
kB = 1.380648e-23
gyroMr = 1.7595e11
avogadro = 6.02214076e23
epsilon_0 = 8.8541e-12
el = 1.6e-19
[faster_calculation]
facEwald = 1e-7
[mi355x]
numbPart = {n}
viscosity = 1e-3
rMagMedian = 1e-8
sigmaLog = {sigma_log}
fixedDt = {fixed}
precision = {precision}
"""

# The benchmark configurations: name -> (N, phi, B (T), Gilbert alpha, sigma_ln, steps)
CONFIGS = {
    "c1": dict(n=1000, phi=0.01, field=0.0, alpha_g=0.1, sigma_log=0.0, steps=10000),
    "c2": dict(n=10000, phi=0.05, field=0.01, alpha_g=0.1, sigma_log=0.2, steps=10000),
    "c3": dict(n=100000, phi=0.10, field=0.05, alpha_g=0.05, sigma_log=0.25, steps=100000),
}


def config_text(name: str, seed: int = 1, noise: bool = True, fixed: bool = True, precision: str = "fp64",
                n: int | None = None) -> str:
    c = dict(CONFIGS[name])
    if n is not None:
        c["n"] = n
    steps = c.pop("steps")
    return _BASE.format(size_dist=1 if c["sigma_log"] > 0 else 0, noise="true" if noise else "false", seed=seed,
                        t_relax=steps * 1e-12, fixed="true" if fixed else "false", precision=precision, **c)


def radii(cfg: capi.Config, n: int, rng: np.random.Generator) -> tuple[np.ndarray, np.ndarray]:
    if cfg.sigma_log > 0:
        r = cfg.r_mag_median * np.exp(cfg.sigma_log * rng.standard_normal(n))
    else:
        r = np.full(n, cfg.r_mag_mean)
    return r, r + cfg.d_shell


def box_length(cfg: capi.Config, r_hydr: np.ndarray) -> float:
    return (np.sum(4.0 / 3.0 * cfg.my_pi * r_hydr ** 3) / cfg.vol_frac) ** (1.0 / 3.0)


def place(r_hydr: np.ndarray, L: float, rng: np.random.Generator) -> np.ndarray:
    """Random sequential addition of spheres of radii r_hydr into the periodic
    box [-L/2, L/2)^3 (no two hydrodynamic spheres overlap, minimum image):
    the library's cell-list placement, seeded from rng."""
    seed = int(rng.integers(0, 2 ** 63))
    cfg = capi.config_from_string("[initial_configuration]\nINIT_CONFIG_TRANS = 1\n")
    return capi.init_positions(cfg, r_hydr, L, seed, periodic=True)


def orientations(n: int, rng: np.random.Generator, my_pi: float) -> np.ndarray:
    az = rng.random(n) * 2 * my_pi
    el = 2.0 * np.arcsin(np.sqrt(rng.random(n)))
    return np.stack([np.sin(el) * np.cos(az), np.sin(el) * np.sin(az), np.cos(el)], axis=1)


def ewald_for(cfg: capi.Config, mag_mom: np.ndarray, r_hydr: np.ndarray, L: float, rc_frac: float):
    """alpha, r_c, n_c for r_c = rc_frac * L by the reference's error model
    (k_c rounded as src/interaction.cpp:304)."""
    import ctypes as C
    a, kc = C.c_double(), C.c_double()
    mu = np.ascontiguousarray(mag_mom)
    rh = np.ascontiguousarray(r_hydr)
    P = C.POINTER(C.c_double)
    st = capi.lib().mds_ewald_params_for_rc(C.byref(cfg), len(mu), mu.ctypes.data_as(P), rh.ctypes.data_as(P), L,
                                           rc_frac * L, C.byref(a), C.byref(kc))
    if st:
        raise capi.MdsError(st, "ewald parameters")
    return a.value, rc_frac * L, int(math.floor(kc.value + 0.5))


def system(name: str, seed: int = 1, noise: bool = True, fixed: bool = True, precision: str = "fp64",
           n: int | None = None, cache_dir: str | None = None):
    """(cfg, r_mag, r_hydr, L, pos, mm, ea) of a benchmark configuration.
    With cache_dir (or $MDS_WORKLOAD_CACHE) the generated arrays are kept in
    an .npz there and reused by later calls with the same name/seed/N."""
    import os
    cfg = capi.config_from_string(config_text(name, seed, noise, fixed, precision, n))
    N = int(cfg.numb_part)
    cache_dir = cache_dir or os.environ.get("MDS_WORKLOAD_CACHE")
    path = os.path.join(cache_dir, f"{name}_{seed}_{N}.npz") if cache_dir else None
    if path and os.path.exists(path):
        z = np.load(path)
        return cfg, z["r"], z["rh"], float(z["L"]), z["pos"], z["ea"].copy(), z["ea"]
    rng = np.random.default_rng(seed)
    r, rh = radii(cfg, N, rng)
    L = box_length(cfg, rh)
    pos = place(rh, L, rng)
    ea = orientations(N, rng, cfg.my_pi)
    if path:
        os.makedirs(cache_dir, exist_ok=True)
        np.savez(path, r=r, rh=rh, L=L, pos=pos, ea=ea)
    return cfg, r, rh, L, pos, ea.copy(), ea
