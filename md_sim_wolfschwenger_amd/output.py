"""Output files of a run in the reference's text formats, and a run driver
that writes them at the points the reference does.

Every file the reference writes goes through saveData
(src/postprocessing.cpp:15-32): the matrix in Eigen's default stream format
(IOFormat(): six significant digits as ``%g`` prints them, every coefficient
right-aligned to the widest one's width, columns separated by one space,
rows by a newline, src/lib/Eigen/src/Core/IO.h:131-227), followed by two
newlines. A call with ``tInt == 0`` replaces the file, any other appends to
it. Vectors are written as one value per line.

Files (names as the reference writes them, in the run directory):
  data.txt    N x 3: r_hydr, r_mag, run parameters in column 2 (rows 0-21),
              written once before the loop (WriteData2TXT, 105-137)
  coords.txt  N x 10: pos, posMm, posEa, tInt in (0, 9); one block at every
              progress point and one at the end (WriteCoords2TXT, 80-100;
              src/simulation.cpp:492-505, 529)
  mz.txt, t.txt         <m_z> after each step and the step's start time
                        (src/simulation.cpp:514-515, 536-537)
  Pxy.txt, Pxz.txt, Pyz.txt, MSDx.txt, MSDy.txt, MSDz.txt, dtMean.txt
              EvalTransCoeffs' series and tEnd / steps (ExportDiffVis, 219-232)
"""
from __future__ import annotations

import math
import os

import numpy as np

TEXT_OUTPUTS = ("data.txt", "coords.txt", "mz.txt", "t.txt", "dtMean.txt", "Pxy.txt", "Pxz.txt", "Pyz.txt",
                "MSDx.txt", "MSDy.txt", "MSDz.txt")


def _g(v: float) -> str:
    # what an ostream with default flags and precision 6 prints for a double
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    if math.isinf(v):
        return "-inf" if v < 0 else "inf"
    return "%.6g" % v


def format_matrix(a) -> str:
    """A matrix (or a vector, as one column) in Eigen's default format."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.size == 0:
        return ""
    cells = [[_g(float(v)) for v in row] for row in a]
    width = max(len(s) for row in cells for s in row)
    return "\n".join(" ".join(s.rjust(width) for s in row) for row in cells)


def save_data(filename: str, matrix, t_int: float) -> bool:
    """saveData (src/postprocessing.cpp:15-32): write the matrix and two
    newlines, replacing the file when t_int == 0, appending otherwise.
    Returns False, writing nothing, where the reference's stream would not
    open."""
    try:
        with open(filename, "w" if t_int == 0 else "a", newline="") as f:
            f.write(format_matrix(matrix) + "\n\n")
    except OSError:
        return False
    return True


def write_coords(filename: str, pos, mm, ea, t_int: float) -> bool:
    """WriteCoords2TXT (src/postprocessing.cpp:80-100)."""
    pos, mm, ea = (np.asarray(a, dtype=np.float64) for a in (pos, mm, ea))
    res = np.zeros((pos.shape[0], 10))
    res[:, 0:3], res[:, 3:6], res[:, 6:9] = pos, mm, ea
    res[0, 9] = t_int
    return save_data(filename, res, t_int)


def run_parameters(cfg, n: int, l_box: float) -> list:
    """Column 2 of data.txt, rows 0-21 (src/postprocessing.cpp:114-135)."""
    return [l_box, n, cfg.mag_flux_dens, cfg.sat_mag, cfg.temp, cfg.t_mag, cfg.t_relax, cfg.t_end, cfg.vol_frac,
            cfg.init_config_rot, cfg.init_config_trans, cfg.seed, cfg.anis_en, cfg.dens_surf_mol,
            cfg.eff_len_surf_mol, cfg.hamaker, cfg.z_valency, cfg.epsilon_r, cfg.gamma_0, cfg.abs_tol,
            cfg.err_tol_ewald, cfg.err_tol_sr]


def write_data(filename: str, r_hydr, r_mag, l_box: float, cfg) -> bool:
    """WriteData2TXT (src/postprocessing.cpp:105-137). The parameter column
    needs N >= 22 rows; the reference indexes past a shorter matrix, so a
    smaller N is an error here."""
    r_hydr = np.asarray(r_hydr, dtype=np.float64)
    n = r_hydr.shape[0]
    if n < 22:
        raise ValueError("data.txt holds 22 run parameters in its third column: N >= 22 needed")
    res = np.zeros((n, 3))
    res[:, 0], res[:, 1] = r_hydr, r_mag
    res[:22, 2] = [float(v) for v in run_parameters(cfg, n, l_box)]
    return save_data(filename, res, 0)


def export_series(directory: str, t, mz, obs, t_int: float, steps: int) -> None:
    """ExportDiffVis (src/postprocessing.cpp:219-232) and the <m_z> and time
    series (src/simulation.cpp:536-537). ``obs`` holds one row per step as
    returned by Context.run_observed: (t, MSDx, MSDy, MSDz, Pxy, Pxz, Pyz)."""
    obs = np.asarray(obs, dtype=np.float64).reshape(-1, 7)
    j = lambda name: os.path.join(directory, name)  # noqa: E731
    with np.errstate(divide="ignore", invalid="ignore"):
        dt_mean = np.float64(t_int) / np.float64(steps)
    save_data(j("dtMean.txt"), np.array([[dt_mean]]), 0)
    for k, name in ((4, "Pxy.txt"), (5, "Pxz.txt"), (6, "Pyz.txt"), (1, "MSDx.txt"), (2, "MSDy.txt"),
                    (3, "MSDz.txt")):
        save_data(j(name), obs[:, k], 0)
    save_data(j("mz.txt"), np.asarray(mz, dtype=np.float64), 0)
    save_data(j("t.txt"), np.asarray(t, dtype=np.float64), 0)


class Run:
    """RunSimulation's loop (src/simulation.cpp:441-538) over a Context:
    steps until the start time passes t_end, the observables of every step
    kept, coords.txt written after the step whose start time reaches the
    next progress point (floor(t / tEnd * npCoords) == count), and every file
    above written in ``directory``.

    Steps between progress points run as one Context.run_observed call (a
    replayed graph for fixed dt); a call never crosses a progress point or
    t_end: at fixed dt its length follows from dt; with the adaptive
    controller, whose step may grow five-fold per step and is not held below
    deltaTmax (src/simulation.cpp:935-950), from that growth bound. One
    replica per context.

    A run can stop (``run(max_steps=...)``), be saved (``checkpoint``: the
    context's binary checkpoint plus the loop's own state, ``<path>.run.npz``)
    and continue in another process (``Run.resume``); the files it then
    writes are those of the run that never stopped."""

    def __init__(self, ctx, cfg, directory: str, r_mag, r_hydr, l_box: float, t_end: float | None = None,
                 chunk: int = 1000, _resume: bool = False):
        self.ctx, self.cfg, self.dir = ctx, cfg, directory
        self.t_end = float(cfg.t_end if t_end is None else t_end)
        if not self.t_end > 0:
            raise ValueError("t_end must be positive")
        self.chunk = int(chunk)
        self.t, self.dt = 0.0, float(cfg.delta_t_init)
        self.steps, self.count = 0, 0
        self._t, self._mz, self._obs = [], [], []
        os.makedirs(directory, exist_ok=True)
        if ctx.replica_count() > 1:
            raise ValueError("one replica per context")
        if not _resume:
            write_data(os.path.join(directory, "data.txt"), r_hydr, r_mag, l_box, cfg)
            ctx.observe(True)

    def checkpoint(self, path: str) -> None:
        """The context's checkpoint at `path` and the loop's state beside it."""
        self.ctx.save_checkpoint(path, self.t, self.dt)
        np.savez(path + ".run.npz", t=self.t, dt=self.dt, steps=self.steps, count=self.count, t_end=self.t_end,
                 chunk=self.chunk, series_t=self.series("t"), series_mz=self.series("mz"),
                 series_obs=self.series("obs"))

    @classmethod
    def resume(cls, ctx, cfg, directory: str, path: str) -> "Run":
        """Continue a checkpointed run in `directory` (whose data.txt and
        coords.txt the first part wrote) on a context created from `cfg`."""
        z = np.load(path + ".run.npz")
        ctx.load_checkpoint(path)
        run = cls(ctx, cfg, directory, None, None, 0.0, t_end=float(z["t_end"]), chunk=int(z["chunk"]),
                  _resume=True)
        run.t, run.dt = float(z["t"]), float(z["dt"])
        run.steps, run.count = int(z["steps"]), int(z["count"])
        run._t, run._mz, run._obs = [z["series_t"]], [z["series_mz"]], [z["series_obs"].reshape(-1, 7)]
        return run

    def _progress_due(self) -> bool:
        return math.floor(self.t / self.t_end * self.cfg.np_coords) == self.count

    def _free_steps(self) -> int:
        """Steps that surely start before the next progress point and at or
        before t_end (one step of margin against rounding)."""
        limit = self.t_end
        level = math.floor(self.t / self.t_end * self.cfg.np_coords)
        if level == self.count:
            return 0
        if level < self.count:
            limit = min(limit, self.count * self.t_end / self.cfg.np_coords)
        gap = limit - self.t
        if gap <= 0 or self.dt <= 0:
            return 0
        if self.cfg.fixed_dt:
            k = int(gap / self.dt) - 1
        else:
            # k steps advance at most dt (1 + 5 + ... + 5^(k-1)) = dt (5^k - 1) / 4
            k = int(math.log(4.0 * gap / self.dt + 1.0) / math.log(5.0)) - 1
        return max(0, min(self.chunk, k))

    def _advance(self, k: int) -> None:
        t, dt, mz, obs = self.ctx.run_observed(k, self.t, self.dt)
        self._t.append(obs[:, 0].copy())
        self._mz.append(np.asarray(mz).reshape(-1))
        self._obs.append(obs)
        self.t, self.dt = t, dt
        self.steps += k

    def run(self, max_steps: int | None = None) -> bool:
        """Step to the end and write the end-of-run files; with max_steps,
        stop once that many steps are done in total (no end files). Returns
        whether the run is finished."""
        coords = os.path.join(self.dir, "coords.txt")
        while self.t <= self.t_end:
            if max_steps is not None and self.steps >= max_steps:
                return False
            k = self._free_steps()
            if max_steps is not None:
                k = min(k, max_steps - self.steps)
            if k > 0:
                self._advance(k)
                continue
            t0, due = self.t, self._progress_due()
            self._advance(1)
            if due:
                self.count += 1
                pos, mm, ea = self.ctx.get_state()
                write_coords(coords, pos, mm, ea, t0)
        pos, mm, ea = self.ctx.get_state()
        write_coords(coords, pos, mm, ea, self.t)
        export_series(self.dir, self.series("t"), self.series("mz"), self.series("obs"), self.t, self.steps)
        return True

    def series(self, name: str) -> np.ndarray:
        parts = {"t": self._t, "mz": self._mz, "obs": self._obs}[name]
        if not parts:
            return np.zeros((0, 7)) if name == "obs" else np.zeros(0)
        return np.concatenate(parts)
