"""Output files and transport observables against what the reference itself
wrote (tests/golden/make_golden.py runs the reference's RunSimulation
post-processing and keeps its text files as bytes).

* The text writers reproduce the reference's files byte for byte from the
  reference's own values: data.txt from the parameters, coords.txt (a block
  after step 0, the progress point at t = 0, and one appended at the end),
  t.txt, mz.txt, MSD*.txt, P*.txt and dtMean.txt.
* The oracle's restatement of EvalTransCoeffs (unwrap, per-axis MSD, P / V)
  reproduces the reference's unwrapped positions bit for bit and its MSD and
  pressure series, replayed over the reference's own trajectory.
"""
import os

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import output
from refcfg import CASES, load, rel_err


def _steps(d):
    return int(d["steps"][0])


def _txt(d, name):
    return bytes(d["txt_" + name]).decode()


def _read(path):
    with open(path, newline="") as f:
        return f.read()


@pytest.mark.parametrize("case", CASES)
def test_text_files_match_reference_bytes(case, tmp_path):
    d = load(case)
    n, k = int(d["N"][0]), _steps(d)
    cfg = mds.config_from_string(d["ini_text"])
    L = float(d["lBox"][0])
    p = str(tmp_path)

    if n >= 22:
        output.write_data(os.path.join(p, "data.txt"), d["rHydr"], d["rMag"], L, cfg)
        assert _read(os.path.join(p, "data.txt")) == _txt(d, "data"), "data.txt"
    else:  # the reference indexes past the matrix; the harness does not call it
        assert "txt_data" not in d
        with pytest.raises(ValueError):
            output.write_data(os.path.join(p, "data.txt"), d["rHydr"], d["rMag"], L, cfg)

    # progress point of step 0 (t = 0 replaces the file), final block appended
    coords = os.path.join(p, "coords.txt")
    with open(coords, "w") as f:
        f.write("stale content\n")
    output.write_coords(coords, d["s0_pos_out"], d["s0_mm_out"], d["s0_ea_out"], float(d["s0_t_in"][0]))
    last = f"s{k - 1}_"
    output.write_coords(coords, d[last + "pos_out"], d[last + "mm_out"], d[last + "ea_out"], float(d["t_end"][0]))
    assert _read(coords) == _txt(d, "coords"), "coords.txt"

    t = np.array([d[f"s{i}_t_in"][0] for i in range(k)])
    mz = np.array([d[f"s{i}_mz"][0] for i in range(k)])
    obs = np.array([[d[f"s{i}_t_in"][0], *d[f"s{i}_msd"], *d[f"s{i}_press_vol"]] for i in range(k)])
    output.export_series(p, t, mz, obs, float(d["t_end"][0]), k)
    for name in ("mz", "t", "dtMean", "Pxy", "Pxz", "Pyz", "MSDx", "MSDy", "MSDz"):
        assert _read(os.path.join(p, name + ".txt")) == _txt(d, name), name


def test_format_matrix_alignment_and_special_values():
    # widths follow the widest entry; %g switches to exponent form at 1e-5 / 1e6
    assert output.format_matrix([[1.0, -2.5], [1e-5, 123456789.0]]) == ("          1        -2.5\n"
                                                                                  "      1e-05 1.23457e+08")
    assert output.format_matrix(np.zeros(0)) == ""
    assert output.format_matrix(np.array([[np.inf, -np.inf]])) == " inf -inf"
    assert output.format_matrix([0.0, -0.0]) == " 0\n-0"


def test_save_data_replaces_at_t0_and_appends_otherwise(tmp_path):
    f = str(tmp_path / "x.txt")
    output.save_data(f, [1.0], 0)
    output.save_data(f, [2.0], 3e-9)
    assert _read(f) == "1\n\n2\n\n"
    output.save_data(f, [3.0], 0)
    assert _read(f) == "3\n\n"
    assert not output.save_data(str(tmp_path / "missing" / "y.txt"), [1.0], 0)


def test_write_data_needs_22_rows(tmp_path):
    cfg = mds.config_from_string(load("coating_n27")["ini_text"])
    with pytest.raises(ValueError):
        output.write_data(str(tmp_path / "data.txt"), np.ones(21), np.ones(21), 1.0, cfg)


@pytest.mark.parametrize("case", CASES)
def test_oracle_trans_coeffs_match_reference(case):
    import oracle
    d = load(case)
    L = float(d["lBox"][0])
    ref = d["s0_pos_in"]
    true_pos = ref.copy()
    for i in range(_steps(d)):
        p = f"s{i}_"
        true_pos, msd, pv = oracle.trans_coeffs(d[p + "pos_out"], true_pos, ref, L, d[p + "press"])
        assert np.array_equal(true_pos, d[p + "true_pos"]), i
        assert rel_err(msd, d[p + "msd"]) < 1e-13, i
        assert np.array_equal(pv, d[p + "press_vol"]), i


def test_oracle_trans_coeffs_unwraps_boundary_crossings():
    """A particle leaving through a face re-enters on the other side; the
    unwrapped track continues and the MSD sees the true displacement."""
    import oracle
    L = 1.0
    ref = np.array([[0.49, 0.0, -0.49], [0.0, 0.2, 0.0]])
    true_pos = ref.copy()
    # step 1: +0.02 in x through +L/2 and -0.02 in z through -L/2
    pos = np.array([[0.51 - L, 0.0, -0.51 + L], [0.0, 0.2, 0.0]])
    true_pos, msd, _ = oracle.trans_coeffs(pos, true_pos, ref, L, np.zeros(3))
    assert np.allclose(true_pos, [[0.51, 0.0, -0.51], [0.0, 0.2, 0.0]], atol=1e-15)
    assert np.allclose(msd, [0.0002, 0.0, 0.0002], rtol=1e-9)
    # step 2: back across the face
    true_pos, msd, _ = oracle.trans_coeffs(ref.copy(), true_pos, ref, L, np.zeros(3))
    assert np.allclose(true_pos, ref, atol=1e-15)
    assert np.allclose(msd, 0.0, atol=1e-30)


def _as_printed(a):
    """What strtod returns for each value as saveData prints it."""
    return np.vectorize(lambda v: float("%.6g" % v))(a)


def test_read_from_file_takes_last_block_of_reference_coords(tmp_path):
    """mds_init_coords_from_sim on the coords.txt the reference wrote (two
    blocks: step 0 and the end): the last block, posMm from columns 3-5 and
    posEa from 6-8, every value as the text holds it."""
    d = load("heun_n125")
    n, k = int(d["N"][0]), _steps(d)
    f = tmp_path / "coords.txt"
    f.write_bytes(bytes(d["txt_coords"]))
    pos, mm, ea, t = mds.init_coords_from_sim(str(f), n)
    last = f"s{k - 1}_"
    assert np.array_equal(pos, _as_printed(d[last + "pos_out"]))
    assert np.array_equal(mm, _as_printed(d[last + "mm_out"]))
    assert np.array_equal(ea, _as_printed(d[last + "ea_out"]))
    assert t == float("%.6g" % d["t_end"][0])
    assert rel_err(mm, d[last + "mm_out"]) < 5e-6


def test_read_from_file_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(4)
    n = 30
    pos, mm, ea = rng.standard_normal((3, n, 3)) * np.array([1e-7, 1.0, 1.0])[:, None, None]
    f = str(tmp_path / "coords_final.txt")
    output.write_coords(f, pos, mm, ea, 0)          # the reference's coords_final.txt: one block
    p2, m2, e2, t = mds.init_coords_from_sim(f, n)
    assert np.array_equal(p2, _as_printed(pos)) and np.array_equal(m2, _as_printed(mm))
    assert np.array_equal(e2, _as_printed(ea)) and t == 0.0
    with pytest.raises(mds.MdsError):
        mds.init_coords_from_sim(f, n + 1)          # block of another size
    with pytest.raises(mds.MdsError):
        mds.init_coords_from_sim(str(tmp_path / "missing.txt"), n)
    g = tmp_path / "nine.txt"
    g.write_text(output.format_matrix(np.zeros((n, 9))) + "\n\n")
    with pytest.raises(mds.MdsError):
        mds.init_coords_from_sim(str(g), n)         # not WriteCoords2TXT's 10 columns
    h = tmp_path / "junk.txt"
    h.write_text(output.format_matrix(np.zeros((n, 10))).replace("0", "x", 1) + "\n")
    with pytest.raises(mds.MdsError):
        mds.init_coords_from_sim(str(h), n)


class _FakeContext:
    """Stands in for a device context in the driver's control logic: steps
    advance the clock by dt and (adaptive) change dt by a seeded factor in
    [0.2, 5], the controller's bounds (src/simulation.cpp:935-950); the state
    records the number of steps taken."""

    def __init__(self, n, adaptive, seed):
        self.n, self.adaptive = n, adaptive
        self.seed = seed
        self.steps = 0

    def _factor(self):
        # a fixed pseudo-random factor in [0.2, 5] per step index
        return 0.2 + 4.8 * ((self.steps * 0.6180339887498949 + self.seed * 0.1) % 1.0)

    def save_checkpoint(self, path, t, dt):
        np.savez(path, steps=self.steps, t=t, dt=dt)

    def load_checkpoint(self, path):
        z = np.load(path if path.endswith(".npz") else path + ".npz")
        self.steps = int(z["steps"])
        return float(z["t"]), float(z["dt"])

    def replica_count(self):
        return 1

    def observe(self, on=True):
        pass

    def run_observed(self, k, t, dt):
        obs, mz = np.zeros((k, 7)), np.zeros(k)
        for i in range(k):
            obs[i, 0] = t
            obs[i, 1:] = self.steps
            mz[i] = self.steps
            t += dt
            if self.adaptive:
                dt *= self._factor()
            self.steps += 1
        return t, dt, mz, obs

    def get_state(self):
        a = np.full((self.n, 3), float(self.steps))
        return a, a, a


@pytest.mark.parametrize("adaptive", [False, True])
def test_run_driver_writes_coords_where_the_reference_loop_does(adaptive, tmp_path):
    """output.Run over a stand-in context: coords.txt blocks, their time
    labels and the step after which each is taken match RunSimulation's loop
    (src/simulation.cpp:464-519) replayed over the same sequence of dt."""
    d = load("heun_n125")
    cfg = mds.config_from_string(d["ini_text"])
    cfg.fixed_dt = 0 if adaptive else 1
    cfg.np_coords = 13.0
    cfg.delta_t_init = 1e-12
    t_end = 400e-12 if not adaptive else 3e-9
    n = 30
    run = output.Run(_FakeContext(n, adaptive, 5), cfg, str(tmp_path), np.ones(n), np.ones(n), 1.0, t_end=t_end,
                     chunk=64)
    run.run()
    # the reference loop over the same dt sequence
    ref = _FakeContext(n, adaptive, 5)
    t, dt, count, expect = 0.0, 1e-12, 0, []
    while t <= t_end:
        due = np.floor(t / t_end * 13.0) == count
        t0 = t
        t, dt, _, _ = ref.run_observed(1, t, dt)
        if due:
            count += 1
            expect.append((t0, ref.steps))
    expect.append((t, ref.steps))
    assert run.steps == ref.steps
    blocks = (tmp_path / "coords.txt").read_text().split("\n\n")[:-1]
    assert len(blocks) == len(expect) and len(expect) > 3
    for b, (lab, steps) in zip(blocks, expect):
        rows = [list(map(float, ln.split())) for ln in b.strip("\n").split("\n")]
        assert rows[0][9] == float("%.6g" % lab)
        assert rows[0][0] == float("%.6g" % steps)
    assert len((tmp_path / "t.txt").read_text().split()) == ref.steps


@pytest.mark.parametrize("adaptive", [False, True])
def test_run_driver_resumed_writes_the_same_files(adaptive, tmp_path):
    """A run stopped after some steps, checkpointed and resumed on a fresh
    context writes the same bytes as the run that never stopped."""
    d = load("heun_n125")
    cfg = mds.config_from_string(d["ini_text"])
    cfg.fixed_dt = 0 if adaptive else 1
    cfg.np_coords = 11.0
    cfg.delta_t_init = 1e-12
    t_end = 300e-12 if not adaptive else 3e-9
    n = 25
    a, b = tmp_path / "a", tmp_path / "b"
    whole = output.Run(_FakeContext(n, adaptive, 3), cfg, str(a), np.ones(n), np.ones(n), 1.0, t_end=t_end,
                       chunk=16)
    whole.run()
    half = whole.steps // 2
    assert half > 5
    first = output.Run(_FakeContext(n, adaptive, 3), cfg, str(b), np.ones(n), np.ones(n), 1.0, t_end=t_end,
                       chunk=16)
    assert not first.run(max_steps=half)
    ck = str(tmp_path / "run.ckpt")
    first.checkpoint(ck)
    second = output.Run.resume(_FakeContext(n, adaptive, 3), cfg, str(b), ck)
    assert second.steps == half
    assert second.run()
    assert second.steps == whole.steps
    for name in output.TEXT_OUTPUTS:
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
