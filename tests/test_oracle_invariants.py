"""Properties the mathematics fixes for the CPU oracle, beyond the golden
vectors: the Ewald sum does not depend on its splitting parameter alpha
(a wrong sign, weight or dropped term in the real-space, k-space, self or
surface part breaks this), the interaction forces sum to zero (Newton's
third law, real and k space), the field of one dipole in a large box tends
to the point-dipole field, and RKCoeffs of one free moment is the LLG
right-hand side in closed form. No GPU."""
import math

import numpy as np

from refcfg import load, oracle_system, props, rel_err


def _system(d, alpha_l, nc, rc_frac):
    import oracle
    L = float(d["lBox"][0])
    return oracle.System(int(d["N"][0]), L, alpha_l / L, rc_frac * L, float(d["rCutSR"][0]), nc, props(d), d["cfg"],
                         threads=4)


def test_ewald_sum_independent_of_alpha():
    d = load("heun_n125_nonoise")
    pos, mm = d["s0_pos_in"], d["s0_mm_in"]
    # two splittings, each converged far below the tolerance checked
    F1, B1, _, _ = _system(d, 8.0, 20, 0.5).comp_interactions(pos.copy(), mm)
    F2, B2, _, _ = _system(d, 10.0, 24, 0.5).comp_interactions(pos.copy(), mm)
    assert rel_err(B1, B2) < 1e-7
    assert rel_err(F1, F2) < 1e-7
    # a splitting that is not converged in k space differs visibly: the
    # check above is able to fail
    F3, B3, _, _ = _system(d, 9.0, 6, 0.5).comp_interactions(pos.copy(), mm)
    assert rel_err(B3, B2) > 1e-5


def test_forces_sum_to_zero():
    d = load("heun_n125_nonoise")
    s = oracle_system(d)
    F, _, _, _ = s.comp_interactions(d["s0_pos_in"].copy(), d["s0_mm_in"], None)
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * np.abs(F).max() * math.sqrt(len(F))


def test_single_dipole_field_tends_to_point_dipole():
    """Two particles at distance r in a box of L >> r: the field at one is the
    point-dipole field mu0/(4 pi) (3 (m.u) u - m) / r^3 of the other plus the
    images and self terms, which vanish as (r/L)^3."""
    import oracle
    d = load("heun_n125_nonoise")
    c = dict(d["cfg"])
    r = 40e-9
    out = []
    for L in (2e-6, 4e-6):
        pos = np.array([[0.0, 0.0, 0.0], [r * 0.6, r * 0.0, r * 0.8]])
        mm = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
        pr = {k: v[:2].copy() for k, v in props(d).items()}
        s = oracle.System(2, L, 7.0 / L, 0.5 * L, 1e-9, 14, pr, c)
        _, B, _, _ = s.comp_interactions(pos.copy(), mm, None)
        u = (pos[1] - pos[0]) / r
        m1 = mm[1] * pr["mag_mom"][1]
        mu0 = 4 * c["my_pi"] * 1e-7
        Bd = mu0 / (4 * c["my_pi"]) * (3 * (m1 @ u) * u - m1) / r ** 3
        out.append(np.abs(B[0] - Bd).max() / np.abs(Bd).max())
    # images, self and surface terms fall with the box volume
    assert out[0] < 5e-3 and out[1] < out[0] / 4


def test_rk_coeffs_closed_form():
    d = load("heun_n125_nonoise")
    s = oracle_system(d)
    n = int(d["N"][0])
    c = d["cfg"]
    rng = np.random.default_rng(1)
    mm = rng.standard_normal((n, 3))
    mm /= np.linalg.norm(mm, axis=1)[:, None]
    ea = rng.standard_normal((n, 3))
    ea /= np.linalg.norm(ea, axis=1)[:, None]
    B = 1e-3 * rng.standard_normal((n, 3))
    z = np.zeros((n, 3))
    bz = 0.02
    km, ke = s.rk_coeffs(bz, mm, ea, B, z, z)
    dot = (mm * ea).sum(axis=1)[:, None]
    Be = B + np.array([0, 0, bz]) + c["anis_const"] * dot * ea
    g = c["vel_mm_const"]
    a = c["mag_damp"]
    km_ref = -g * (np.cross(mm, Be) + a * np.cross(mm, np.cross(mm, Be)))
    assert rel_err(km, km_ref) < 1e-13
    # easy axis relaxes toward the moment: (m - (m.e) e) c1 (m.e)
    c1 = 2 * c["anis_en"] * d["volMag"] / d["zetaRot"]
    ke_ref = (mm - dot * ea) * (c1[:, None] * dot)
    if c["enable_mobilization"]:
        assert rel_err(ke, ke_ref) < 1e-13
    else:
        assert np.abs(ke).max() == 0
    # |m| is conserved by the LLG right-hand side
    assert np.abs((km * mm).sum(axis=1)).max() < 1e-12 * np.abs(km).max()
