"""The CPU oracle against the golden vectors the reference itself wrote.

Each step is started from the reference's own state and thermal draws for that
step (tests/golden/make_golden.py), so every comparison is per evaluation /
per step. Tolerance: 1e-10 relative (max-norm over the array).

Every case runs with threads = 1 (the stored pair list, as the reference) and
threads = 4: without an overlap move and without the coating potential, the
oracle then streams the pairs over OpenMP threads instead of storing them
(the form that checks the device at N = 1e4 .. 1e5), so that form is held to
the same vectors.
"""
import numpy as np
import pytest

from refcfg import CASES, contact_scales, load, oracle_system, rel_err

TOL = 1e-10


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("case", CASES)
def test_comp_interactions_matches_reference(case, threads):
    d = load(case)
    s = oracle_system(d, threads=threads)
    f_al, b_al, _ = contact_scales(d)
    F, B, P, pos = s.comp_interactions(d["s0_pos_in"], d["s0_mm_in"], d["s0_nF"])
    assert rel_err(F, d["s0_force"], f_al) < TOL
    assert rel_err(B, d["s0_field"], b_al) < TOL
    assert rel_err(P, d["s0_press"], f_al * 2 * float(d["rHydr"].min())) < TOL
    assert rel_err(pos, d["s0_pos_mid"]) < 1e-14
    pl = s.pairs(d["s0_pos_in"])
    assert len(pl["i"]) == int(d["s0_pairs"][0])


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("case", CASES)
def test_steps_match_reference(case, threads):
    d = load(case)
    s = oracle_system(d, threads=threads)
    _, b_al, _ = contact_scales(d)
    for k in range(int(d["steps"][0])):
        p = f"s{k}_"
        out = s.step(d[p + "pos_in"], d[p + "mm_in"], d[p + "ea_in"], float(d[p + "t_in"][0]),
                     float(d[p + "dt_in"][0]), d[p + "nT"], d[p + "nB"], d[p + "nF"])
        assert out["status"] == 0
        assert out["dt_used"] == pytest.approx(float(d[p + "dt_used"][0]), rel=1e-12)
        assert out["dt"] == pytest.approx(float(d[p + "dt_out"][0]), rel=1e-10)
        assert rel_err(out["mm"], d[p + "mm_out"]) < TOL, k
        assert rel_err(out["ea"], d[p + "ea_out"]) < TOL, k
        assert rel_err(out["pos"], d[p + "pos_out"]) < 1e-13, k
        assert rel_err(out["field_out"], d[p + "field_out"], b_al) < TOL, k
        assert rel_err(out["n_t"], d[p + "nT_out"]) < 1e-13, k
        assert rel_err(out["n_b"], d[p + "nB_out"]) < 1e-13, k
        assert out["mz"] == pytest.approx(float(d[p + "mz"][0]), rel=1e-12, abs=1e-14)


def test_fixture_exercises_quirks():
    """The golden set covers overlap moves, step rejections and every solver."""
    d = load("overlap_n64")
    assert np.abs(d["s0_pos_mid"] - d["s0_pos_in"]).max() > 0
    d = load("heun_reject_n64")
    assert d["s0_dt_used"][0] < d["s0_dt_in"][0]
    solvers = {load(c)["cfg"]["solver"] for c in CASES}
    assert solvers == {0, 1, 2}
