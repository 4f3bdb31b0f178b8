"""Write the golden vectors for the Brownian/LLG step from the reference itself.

Runs in the build container only (it needs the reference checkout): builds
``oracle/_ref/ref_harness`` (the reference's unmodified sources linked with the
project's own driver, ``oracle/ref_harness.cpp``), runs it once per case in a
scratch directory holding that case's ``configuration.ini``, and packs what it
dumps into ``tests/golden/<case>.npz`` (plain arrays, no pickles).

Every array is the reference's own output: particle properties and initial
state from ``InitParticleProps``/``InitCoordsFromSketch``
(src/simulation.cpp:118-351), thermal draws from ``DefThermFluct`` (563-603),
forces/fields/pressure from ``CompInteractions`` (628-680), rotation from
``IntegrationRot`` (707-868), positions from ``IntegrationTrans`` +
``ApplyBoundaryCondArr`` (959-962, 610-620), the transport observables of
``EvalTransCoeffs`` (src/postprocessing.cpp:167-187) and the text files the
reference writes (``saveData``, src/postprocessing.cpp:15-32), kept as bytes.

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

# Keys and sections of the reference's configuration.ini, with the values the
# cases start from (src/configuration.ini:1-79 defines the keys).
BASE = {
    "initial_configuration": {"INIT_CONFIG_ROT": 2, "INIT_CONFIG_TRANS": 1, "READ_FROM_FILE": "false"},
    "size_distribution": {"SIZE_DIST": 0},
    "mobilization": {"ENABLE_MOBILIZATION": "true"},
    "thermal_fluctuations": {"ENABLE_THERM_FIELD": "true", "ENABLE_THERM_TORQUE": "true",
                             "ENABLE_THERM_FORCE": "true", "seed": 1},
    "interactions": {"ENABLE_COATING_POT": "false"},
    "adaptive_timestepping_solver": {"SOLVER": 0},
    "particle_properties": {"npDim": 5, "magDamp": 0.1, "temp": 300.0, "satMag": 4.8e5, "anisEn": 1e4},
    "assemble_properties": {"rMagMean": 1e-8, "dShell": 5e-9, "volFrac": 0.01},
    "steric_repulsion": {"densSurfMol": 5e17},
    "vdw_attraction": {"hamaker": 33e-21},
    "electrostat_repulsion": {"z_valency": 1, "epsilon_r": 78.5, "gamma_0": 0.04, "conc": 50},
    "time_settings": {"tMag": 0, "tRelax": 1e-8, "deltaTinit": 1e-12},
    "external_magnetic_field_t": {"magFluxDens": 0.0},
    "solver_settings": {"absTol": 1e-4, "deltaTmin": 1e-14, "deltaTmax": 1e-9},
    "long_range_interactions": {"errTolEwald": 1e-6},
    "short_range_interactions": {"errTolSR": 0.01},
    "output_settings": {"npCoords": 100},
    "physical_constants": {"my_pi": "3.14159265359", "kB": "1.380648e-23", "gyroMr": "1.7595e11",
                           "avogadro": "6.02214076e23", "epsilon_0": "8.8541e-12", "el": "1.6e-19"},
    "faster_calculation": {"facEwald": "1e-7"},
}


def ini_text(over: dict) -> str:
    out = []
    for sec, kv in BASE.items():
        out.append(f"[{sec}]")
        merged = dict(kv)
        merged.update(over.get(sec, {}))
        for k, v in merged.items():
            out.append(f"{k} = {v}\t\t; case value")
        out.append("")
    return "\n".join(out)


def overlap_state(seed: int, n_dim: int, lbox: float, r_mag: float) -> np.ndarray:
    """Grid positions with a few pairs and one three-particle chain pushed into
    core overlap, one pair overlapping across the periodic boundary."""
    rng = np.random.default_rng(seed)
    gap = lbox / n_dim
    g = (np.arange(n_dim) + 0.5) * gap - lbox / 2
    pos = np.array([[x, y, z] for x in g for y in g for z in g], dtype=np.float64)
    pos += rng.uniform(-0.05, 0.05, pos.shape) * gap
    n = pos.shape[0]
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    # pairs (0,1), (5,9): overlap by 5-10 % of the contact distance
    for a, b, f in [(0, 1, 0.93), (5, 9, 0.9)]:
        pos[b] = pos[a] + u[a] * 2 * r_mag * f
    # chain 20 - 21 - 22 along a line, both contacts overlapping
    pos[21] = pos[20] + u[20] * 2 * r_mag * 0.95
    pos[22] = pos[21] + u[20] * 2 * r_mag * 0.92
    # pair across the +x face of the box (cases with enough particles)
    if n > 31:
        pos[30] = np.array([lbox / 2 - 0.3 * r_mag, 0.1 * lbox, -0.2 * lbox])
        pos[31] = np.array([-lbox / 2 + 0.5 * r_mag, 0.1 * lbox, -0.2 * lbox])
    # keep every position inside [-L/2, L/2] as the reference's wrap does
    pos = np.where(pos > lbox / 2, pos - lbox, pos)
    pos = np.where(pos < -lbox / 2, pos + lbox, pos)
    mm = rng.normal(size=(n, 3))
    mm /= np.linalg.norm(mm, axis=1, keepdims=True)
    ea = rng.normal(size=(n, 3))
    ea /= np.linalg.norm(ea, axis=1, keepdims=True)
    return np.stack([pos, mm, ea])


# name -> (ini overrides, alphaL, rcFrac, nc, steps, state builder or None)
CASES = {
    # C1-like: monodisperse magnetite, phi=1 %, thermal noise on, Heun-Euler
    "heun_n125": ({"particle_properties": {"npDim": 5}}, 5.85, 0.5, 5, 6, None),
    # same, no thermal noise (deterministic path)
    "heun_n125_nonoise": ({"particle_properties": {"npDim": 5},
                           "thermal_fluctuations": {"ENABLE_THERM_FIELD": "false", "ENABLE_THERM_TORQUE": "false",
                                                    "ENABLE_THERM_FORCE": "false"}}, 7.0, 0.4, 7, 4, None),
    # forced core overlaps (sequential overlap correction), field on
    "overlap_n64": ({"particle_properties": {"npDim": 4}, "assemble_properties": {"volFrac": 0.05},
                     "external_magnetic_field_t": {"magFluxDens": 0.02}, "time_settings": {"tMag": 1e-6}},
                    8.0, 0.4, 7, 3, "overlap"),
    # lognormal sizes, Bogacki-Shampine, tight tolerance so steps are rejected
    "bs_lognormal_n125": ({"size_distribution": {"SIZE_DIST": 1}, "adaptive_timestepping_solver": {"SOLVER": 1},
                           "assemble_properties": {"volFrac": 0.05}, "particle_properties": {"magDamp": 0.05},
                           "external_magnetic_field_t": {"magFluxDens": 0.05}, "time_settings": {"tMag": 1e-6},
                           "solver_settings": {"absTol": 2e-6}}, 10.0, 0.3, 9, 4, None),
    # Dormand-Prince, immobilised easy axes, adaptive step growth
    "dp_immobile_n64": ({"particle_properties": {"npDim": 4, "anisEn": 2e4}, "adaptive_timestepping_solver": {"SOLVER": 2},
                         "mobilization": {"ENABLE_MOBILIZATION": "false"},
                         "external_magnetic_field_t": {"magFluxDens": 0.01}, "time_settings": {"tMag": 1e-6}},
                        6.0, 0.45, 6, 4, None),
    # Heun-Euler with a tolerance that forces rejections (noise rescaling quirk)
    "heun_reject_n64": ({"particle_properties": {"npDim": 4}, "solver_settings": {"absTol": 2e-5},
                         "time_settings": {"deltaTinit": 5e-11}}, 6.0, 0.45, 6, 4, None),
    # coating potentials (steric + electrostatic + vdW) with close pairs
    "coating_n27": ({"particle_properties": {"npDim": 3}, "interactions": {"ENABLE_COATING_POT": "true"},
                     "assemble_properties": {"rMagMean": 8e-9, "dShell": 0, "volFrac": 0.05}},
                    6.0, 0.45, 5, 1, "overlap_coating"),
    # simple-cubic grid, all moments along z: Ewald total field and force vanish
    "grid_sc_n216": ({"particle_properties": {"npDim": 6}, "initial_configuration": {"INIT_CONFIG_TRANS": 2,
                                                                                  "INIT_CONFIG_ROT": 3},
                      "thermal_fluctuations": {"ENABLE_THERM_FIELD": "false", "ENABLE_THERM_TORQUE": "false",
                                               "ENABLE_THERM_FORCE": "false"}}, 9.0, 0.4, 10, 1, None),
    # C1 configuration at its full size: N=1000, phi=1 %
    "heun_n1000": ({"particle_properties": {"npDim": 10}}, 5.85, 0.5, 5, 2, None),
    # a single particle with its periodic images, field on, mobile easy axis
    "single_n1": ({"particle_properties": {"npDim": 1}, "external_magnetic_field_t": {"magFluxDens": 0.02},
                   "time_settings": {"tMag": 1e-6}}, 6.0, 0.45, 6, 4, None),
}


# files the reference writes through saveData (src/postprocessing.cpp:15-32)
TEXT_OUTPUTS = ("data.txt", "coords.txt", "mz.txt", "t.txt", "dtMean.txt", "Pxy.txt", "Pxz.txt", "Pyz.txt",
                "MSDx.txt", "MSDy.txt", "MSDz.txt")


def read_dump(prefix: str) -> dict:
    out = {}
    with open(prefix + ".idx") as f, open(prefix + ".bin", "rb") as b:
        for line in f:
            name, r, c = line.split()
            r, c = int(r), int(c)
            a = np.frombuffer(b.read(8 * r * c), dtype="<f8").reshape((r, c), order="F")
            out[name] = a[:, 0].copy() if c == 1 else a.copy()
    return out


def run_case(name: str) -> str:
    over, alpha_l, rc, nc, steps, state = CASES[name]
    with tempfile.TemporaryDirectory() as tmp:
        run = os.path.join(tmp, "run")
        os.makedirs(run)
        with open(os.path.join(tmp, "configuration.ini"), "w") as f:
            f.write(ini_text(over))
        args = [HARNESS, os.path.join(run, "out"), repr(alpha_l), repr(rc), str(nc), str(steps)]
        if state is not None:
            # first pass with no steps to learn the box and radii
            pdir = os.path.join(tmp, "probe")
            os.makedirs(pdir)
            subprocess.run([HARNESS, os.path.join(pdir, "probe"), repr(alpha_l), repr(rc), str(nc), "0"],
                           cwd=pdir, check=True, stdout=subprocess.DEVNULL, timeout=120)
            probe = read_dump(os.path.join(pdir, "probe"))
            n_dim = round(probe["N"][0] ** (1 / 3))
            r_mag = float(probe["rMag"].min())
            st = overlap_state(7 if state == "overlap" else 11, n_dim, float(probe["lBox"][0]), r_mag)
            if state == "overlap_coating":
                # coating case: pull pairs to just outside core contact instead
                st[0, 1] = st[0, 0] + (st[0, 1] - st[0, 0]) * 1.12
            sp = os.path.join(tmp, "state.bin")
            # column-major N x 3 arrays, as Eigen stores them
            with open(sp, "wb") as f:
                for a in st:
                    f.write(np.asfortranarray(a).tobytes(order="F"))
            args.append(sp)
        # the reference loops forever when a step is rejected at deltaTmin
        # (src/simulation.cpp:917-927, 725); a case that does is a bad case
        subprocess.run(args, cwd=run, check=True, stdout=subprocess.DEVNULL, timeout=300)
        data = read_dump(os.path.join(run, "out"))
        # the reference's text outputs, byte for byte
        for fn in TEXT_OUTPUTS:
            fp = os.path.join(run, fn)
            if os.path.exists(fp):
                with open(fp, "rb") as f:
                    data["txt_" + fn[:-4]] = np.frombuffer(f.read(), dtype=np.uint8)
    data["ini"] = np.frombuffer(ini_text(over).encode(), dtype=np.uint8)
    data["steps"] = np.array([steps], dtype=np.float64)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **data)
    return path


def main(argv):
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True, stdout=subprocess.DEVNULL)
    names = argv or list(CASES)
    meta = {"compiler": subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
            "flags": "-std=c++14 -O2 (x86-64 default ISA: SSE2, so Vmt19937_64 = Pmt19937_64<Packet4i>)",
            "cases": {}}
    for n in names:
        p = run_case(n)
        meta["cases"][n] = {"alphaL": CASES[n][1], "rcFrac": CASES[n][2], "nc": CASES[n][3],
                            "steps": CASES[n][4], "bytes": os.path.getsize(p)}
        print(n, os.path.getsize(p))
    mp = os.path.join(HERE, "manifest.json")
    old = json.load(open(mp)) if os.path.exists(mp) else {"cases": {}}
    old["cases"].update(meta["cases"])
    meta["cases"] = old["cases"]
    with open(mp, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
