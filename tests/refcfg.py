"""Test helpers: configuration constants and golden-fixture loading.

The derived constants are computed here independently of the product's C++
parser, following src/configuration.cpp:11-84 of the reference, so a test can
compare the two.
"""
from __future__ import annotations

import configparser
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def parse_ini(text: str) -> dict:
    cp = configparser.ConfigParser(inline_comment_prefixes=(";",), strict=False)
    cp.optionxform = str.lower
    cp.read_string(text)
    out = {}
    for sec in cp.sections():
        for k, v in cp.items(sec):
            out[k] = v
    return out


def _b(v):
    return str(v).strip().lower() in ("true", "yes", "on", "1")


def derived(ini: dict) -> dict:
    """Constants of Configuration::Configuration (src/configuration.cpp:11-84)."""
    g = lambda k: float(ini[k.lower()])
    c = {}
    c["my_pi"] = g("my_pi")
    c["mu0"] = 4.0 * c["my_pi"] * 1e-7
    c["kB"] = g("kB")
    c["gyro_mr"] = g("gyroMr")
    c["temp"] = g("temp")
    c["mag_damp"] = g("magDamp")
    c["sat_mag"] = g("satMag")
    c["anis_en"] = g("anisEn")
    c["fac_ewald"] = g("facEwald")
    c["vel_mm_const"] = c["gyro_mr"] / (1.0 + c["mag_damp"] ** 2)
    c["anis_const"] = 2.0 * c["anis_en"] / c["sat_mag"]
    c["eff_len_surf_mol"] = 0.2 * g("rMagMean")
    c["ster_rep_fac"] = 4 * c["kB"] * c["temp"] * c["my_pi"] * g("densSurfMol") / (2 * c["eff_len_surf_mol"])
    el, z = g("el"), g("z_valency")
    gamma = math.tanh(z * el * g("gamma_0") / (4 * c["kB"] * c["temp"]))
    eps = g("epsilon_0") * g("epsilon_r")
    c["kappa"] = math.sqrt(el * el * g("avogadro") * g("conc") * z * z / (eps * c["kB"] * c["temp"]))
    c["electrostat_rep_fac"] = c["kappa"] * 64 * c["my_pi"] * eps * (c["kB"] * c["temp"] / (z * el)) ** 2 * gamma ** 2
    c["hamaker"] = g("hamaker")
    c["vis"] = 2.414e-5 * 10 ** (247.8 / (c["temp"] - 140))
    c["mag_flux_dens"] = g("magFluxDens")
    c["t_mag"] = g("tMag")
    c["abs_tol"] = g("absTol")
    c["dt_min"] = g("deltaTmin")
    c["dt_max"] = g("deltaTmax")
    c["enable_mobilization"] = int(_b(ini["enable_mobilization"]))
    c["enable_coating_pot"] = int(_b(ini["enable_coating_pot"]))
    c["solver"] = int(ini["solver"])
    c["fixed_dt"] = 0
    return c


def load(name: str) -> dict:
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    d["ini_text"] = bytes(d.pop("ini")).decode()
    d["cfg"] = derived(parse_ini(d["ini_text"]))
    return d


def props(d: dict) -> dict:
    return {"r_mag": d["rMag"], "r_hydr": d["rHydr"], "vol_mag": d["volMag"], "mag_mom": d["magMom"],
            "zeta_rot": d["zetaRot"], "zeta_trans": d["zetaTrans"]}


def oracle_system(d: dict, threads: int = 1):
    import oracle
    return oracle.System(int(d["N"][0]), float(d["lBox"][0]), float(d["alpha"][0]), float(d["rCutLR"][0]),
                         float(d["rCutSR"][0]), int(d["nc"][0]), props(d), d["cfg"], threads=threads)


def rel_err(a, b, floor: float = 0.0) -> float:
    """max |a-b| / max(max |b|, floor) over the whole array (norm-wise
    relative error; `floor` guards sums that cancel to ~0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.abs(b).max()), floor)
    if den == 0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / den)


def contact_scales(d: dict) -> tuple:
    """Force, field and torque at contact of two of the smallest particles, the
    normalisers of the reference's CheckEwaldSum (src/interaction.cpp:743-753)."""
    c = d["cfg"]
    m = float(d["magMom"].min())
    rh = float(d["rHydr"].min())
    f_al = 3 * c["mu0"] * m * m / (2 * c["my_pi"] * (2 * rh) ** 4)
    b_al = c["mu0"] * m / (2 * c["my_pi"] * (2 * rh) ** 3)
    t_al = m * m * c["mu0"] / (16.0 * c["my_pi"] * rh ** 3)
    return f_al, b_al, t_al


CASES = ["heun_n125", "heun_n125_nonoise", "overlap_n64", "bs_lognormal_n125", "dp_immobile_n64",
         "heun_reject_n64", "coating_n27", "grid_sc_n216", "heun_n1000", "single_n1"]
