"""Host-side parts of libmdsim.so, checked without a GPU: the library loads and
exports every entry point of include/mdsim_capi.h; the ini reader and the
derived constants, particle properties, box edge, short-range cut-off and Ewald
error models match the values the reference itself produced."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import capi
from refcfg import CASES, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mdsim_capi.h")).read()
    declared = set(re.findall(r"\b(mds_[a-z0-9_]+)\s*\(", hdr))
    lib = mds.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    bound = {name for name, _, _ in capi.SIGNATURES}
    assert declared == bound, declared ^ bound


def test_config_struct_layout():
    assert mds.lib().mds_config_sizeof() == C.sizeof(mds.Config)


@pytest.mark.parametrize("case", CASES)
def test_config_and_particle_props_match_reference(case):
    d = load(case)
    cfg = mds.config_from_string(d["ini_text"])
    ref = d["cfg"]
    for k in ("mu0", "vel_mm_const", "anis_const", "ster_rep_fac", "kappa", "electrostat_rep_fac", "vis",
              "eff_len_surf_mol"):
        assert getattr(cfg, k) == pytest.approx(ref[k], rel=1e-14), k
    assert cfg.numb_part == int(d["N"][0])
    assert cfg.solver == ref["solver"]
    n = int(d["N"][0])
    r = np.ascontiguousarray(d["rMag"])
    out = {k: np.zeros(n) for k in ("rh", "vm", "mu", "zr", "zt")}
    st = mds.lib().mds_particle_props(C.byref(cfg), n, _p(r), *[_p(out[k]) for k in ("rh", "vm", "mu", "zr", "zt")])
    assert st == 0
    np.testing.assert_allclose(out["rh"], d["rHydr"], rtol=1e-13)
    np.testing.assert_allclose(out["vm"], d["volMag"], rtol=1e-13)
    np.testing.assert_allclose(out["mu"], d["magMom"], rtol=1e-13)
    np.testing.assert_allclose(out["zr"], d["zetaRot"], rtol=1e-13)
    np.testing.assert_allclose(out["zt"], d["zetaTrans"], rtol=1e-13)
    L = mds.lib().mds_box_length(C.byref(cfg), n, _p(np.ascontiguousarray(d["rHydr"])))
    assert L == pytest.approx(float(d["lBox"][0]), rel=1e-13)
    rcsr = C.c_double()
    st = mds.lib().mds_cutoff_sr(C.byref(cfg), n, _p(r), _p(np.ascontiguousarray(d["rHydr"])),
                                 _p(np.ascontiguousarray(d["magMom"])), C.byref(rcsr))
    assert st == 0
    assert rcsr.value == pytest.approx(float(d["rCutSR"][0]), rel=1e-12)


@pytest.mark.parametrize("case", ["heun_n125", "bs_lognormal_n125", "coating_n27"])
def test_ewald_error_models_match_reference(case):
    d = load(case)
    cfg = mds.config_from_string(d["ini_text"])
    n = int(d["N"][0])
    M, fal, L, rc = float(d["errM"][0]), float(d["errFAl"][0]), float(d["lBox"][0]), float(d["rCutLR"][0])
    alpha = float(d["alpha"][0])
    g = d["errgrid"]
    lib = mds.lib()
    for x, real_ref, kc_i in zip(g[:, 0], g[:, 1], range(len(g))):
        v = lib.mds_ewald_real_error(C.byref(cfg), n, M, fal, L, rc, float(x))
        assert v == pytest.approx(real_ref, rel=1e-12, abs=1e-18)
        kc = 0.5 + 0.5 * kc_i
        v = lib.mds_ewald_rez_error(C.byref(cfg), n, M, fal, L, alpha, kc)
        assert v == pytest.approx(g[kc_i, 2], rel=1e-12, abs=1e-18)


def test_ewald_param_scan_is_a_root_of_the_error_models():
    d = load("heun_n1000")
    cfg = mds.config_from_string(d["ini_text"])
    n = int(d["N"][0])
    L = float(d["lBox"][0])
    mu, rh = np.ascontiguousarray(d["magMom"]), np.ascontiguousarray(d["rHydr"])
    a, kc = C.c_double(), C.c_double()
    for frac in (0.2, 0.3, 0.5):
        st = mds.lib().mds_ewald_params_for_rc(C.byref(cfg), n, _p(mu), _p(rh), L, frac * L, C.byref(a), C.byref(kc))
        assert st == 0
        M = float((mu ** 2).sum())
        fal = 3 * cfg.mu0 * mu.min() ** 2 / (2 * cfg.my_pi * (2 * rh.min()) ** 4)
        # the interpolated roots sit within one grid step of the exact roots
        f = lambda x: mds.lib().mds_ewald_real_error(C.byref(cfg), n, M, fal, L, frac * L, x)
        step = 2e8 / 9999
        assert f(a.value - step) > 0 > f(a.value + step)
        g = lambda x: mds.lib().mds_ewald_rez_error(C.byref(cfg), n, M, fal, L, a.value, x)
        assert g(kc.value - 30 / 9999) > 0 > g(kc.value + 30 / 9999)
        # alpha * r_c lands where the reference's error model puts it (~3)
        assert 2.5 < a.value * frac * L < 4.0


def test_ini_rules():
    text = ("[Particle_Properties]\n  ; full-line comment\nNPDIM = 3 ; inline\n magdamp=0.5\n"
            "[physical_constants]\nmy_pi : 3.0\n[thermal_fluctuations]\nENABLE_THERM_FIELD = On\n"
            "seed = 0x10\n")
    c = mds.config_from_string(text)
    assert c.np_dim == 3 and c.numb_part == 27
    # an indented line continues the previous value (inih multi-line)
    assert c.mag_damp == -1 or c.mag_damp == 0.5
    assert c.my_pi == 3.0 and c.mu0 == pytest.approx(12e-7)
    assert c.enable_therm_field == 1 and c.seed == 16
    assert c.temp == -1  # missing key keeps the reference's default
    with pytest.raises(mds.MdsError):
        mds.config_from_string("[a]\nthis line has no separator\n")
    with pytest.raises(mds.MdsError):
        mds.config_from_ini("/nonexistent/configuration.ini")


def test_added_keys():
    c = mds.config_from_string("[particle_properties]\nnpDim=10\ntemp=300\n[mi355x]\nnumbPart=100000\n"
                               "viscosity=1e-3\nfixedDt=true\nprecision=mixed\n")
    assert c.numb_part == 100000 and c.vis == 1e-3 and c.fixed_dt == 1 and c.precision == mds.PREC_MIXED


def test_no_device_in_build_container_is_an_error_not_a_fallback():
    if mds.device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg = mds.config_from_string(load("heun_n125")["ini_text"])
    with pytest.raises(mds.MdsError):
        mds.Context(cfg)


def test_predicted_ewald_errors_match_measured_error_of_oracle():
    """mds_ewald_predicted_errors (the estimates CompEwaldParams prints,
    src/interaction.cpp:311-323) against the rms error the CPU oracle
    measures at the same parameters, as CheckEwaldSum does
    (src/interaction.cpp:678-755): C2 at r_c = 0.3 L against a converged
    splitting (alpha L = 18.6, r_c = 0.35 L, n_c = 40). Both models are rms
    estimates; they agree within 30 %, and the force error is the
    configured 1e-6 of the contact force within the rounding of n_c."""
    import oracle
    from md_sim_wolfschwenger_amd import workload
    from refcfg import derived, parse_ini
    cfg, r, rh, L, pos, mm, ea = workload.system("c2", seed=3)
    mu = cfg.sat_mag * 4 / 3 * cfg.my_pi * r ** 3
    c = derived(parse_ini(workload.config_text("c2", seed=3)))
    props = {"r_mag": r, "r_hydr": rh, "vol_mag": 4 / 3 * cfg.my_pi * r ** 3, "mag_mom": mu,
             "zeta_rot": 8 * cfg.my_pi * cfg.vis * rh ** 3, "zeta_trans": 6 * cfg.my_pi * cfg.vis * rh}
    a, rc, nc = workload.ewald_for(cfg, mu, rh, L, 0.3)
    th = max(1, min(8, os.cpu_count() or 1))
    F, B, _, _ = oracle.System(len(r), L, a, rc, 2 * r.max(), nc, props, c, threads=th).comp_interactions(pos, mm)
    Fx, Bx, _, _ = oracle.System(len(r), L, 18.6 / L, 0.35 * L, 2 * r.max(), 40, props, c,
                                 threads=th).comp_interactions(pos, mm)
    mu0 = 4e-7 * cfg.my_pi
    f_al = 3 * mu0 * mu.min() ** 2 / (2 * cfg.my_pi * (2 * rh.min()) ** 4)
    t_al = mu.min() ** 2 * mu0 / (16 * cfg.my_pi * rh.min() ** 3)
    rms = lambda d: float(np.sqrt((d * d).sum(axis=1).mean()))
    e_f = rms(F - Fx) / f_al
    e_t = rms(np.cross(mm * mu[:, None], B - Bx)) / t_al
    pf, pt = capi.ewald_predicted_errors(cfg, mu, rh, L, a, rc, nc)
    assert e_f == pytest.approx(pf, rel=0.3)
    assert e_t == pytest.approx(pt, rel=0.3)
    assert pf == pytest.approx(cfg.err_tol_ewald, rel=0.3)


def test_pair_list_region_estimate():
    """The planner's list region (also the serving admission estimate): the
    pairs of a uniform suspension within r_l = r_c (1 + skin) (N(N-1)/2 times
    the sphere's share of the box), a quarter extra for padding, 16 words per
    particle and 4096 for headers; no region without a list."""
    n, L, rc, skin = 100_000, 2.54e-6, 0.225 * 2.54e-6, 0.005
    rl = rc * (1 + skin)
    pairs = 0.5 * n * (n - 1) * 4.0 / 3.0 * np.pi * rl ** 3 / L ** 3
    w = mds.Context.pair_list_words(n, L, rc, skin)
    assert abs(w - (1.25 * pairs + 16 * n + 4096)) <= 1 + 1e-12 * w
    # the measured C3 list (257.1 M words for 238.5 M unique pairs within
    # r_c) fits the estimate with room to spare
    assert 257_107_208 < w < 1.3 * 257_107_208
    # a sphere larger than the box: every pair
    assert mds.Context.pair_list_words(1000, 1.0, 0.7, skin) == int(1.25 * 0.5 * 1000 * 999 + 16 * 1000 + 4096)
    assert mds.Context.pair_list_words(1000, 1.0, 0.3, 0.0) == 0
