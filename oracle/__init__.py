"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's CPU-baseline leg, as the checker of the HIP path; never by the
product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("l_box", C.c_double),
        ("alpha", C.c_double), ("r_cut_lr", C.c_double), ("r_cut_sr", C.c_double), ("nc", C.c_int32),
        ("r_mag", _dp), ("r_hydr", _dp), ("vol_mag", _dp), ("mag_mom", _dp), ("zeta_rot", _dp),
        ("zeta_trans", _dp),
        ("my_pi", C.c_double), ("fac_ewald", C.c_double), ("kB", C.c_double), ("temp", C.c_double),
        ("gyro_mr", C.c_double), ("mag_damp", C.c_double), ("sat_mag", C.c_double), ("anis_en", C.c_double),
        ("anis_const", C.c_double), ("vel_mm_const", C.c_double), ("mag_flux_dens", C.c_double),
        ("t_mag", C.c_double), ("abs_tol", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("eff_len_surf_mol", C.c_double), ("ster_rep_fac", C.c_double), ("electrostat_rep_fac", C.c_double),
        ("kappa", C.c_double), ("hamaker", C.c_double),
        ("enable_mobilization", C.c_int32), ("enable_coating_pot", C.c_int32), ("solver", C.c_int32),
        ("fixed_dt", C.c_int32), ("threads", C.c_int32),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_pairs_build.restype = vp
        L.orc_pairs_build.argtypes = [C.POINTER(OrcParams), _dp]
        L.orc_pairs_free.argtypes = [vp]
        L.orc_pairs_count.restype = C.c_int64
        L.orc_pairs_count.argtypes = [vp]
        L.orc_pairs_overlaps.restype = C.c_int64
        L.orc_pairs_overlaps.argtypes = [vp]
        L.orc_pairs_get.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _dp, _dp, C.POINTER(C.c_int32)]
        L.orc_comp_interactions.argtypes = [C.POINTER(OrcParams), _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(vp)]
        L.orc_comp_demag_field.argtypes = [C.POINTER(OrcParams), vp, _dp, _dp, _dp]
        L.orc_kspace.argtypes = [C.POINTER(OrcParams), _dp, _dp, _dp, _dp, _dp]
        L.orc_rk_coeffs.argtypes = [C.POINTER(OrcParams), C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_step.argtypes = [C.POINTER(OrcParams), _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               _dp, _dp, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def colmajor(a: np.ndarray) -> np.ndarray:
    """(N,3) array -> a new contiguous column-major buffer as a flat (3N,)
    array. Always a copy: the C side writes into the state buffers, and for
    N = 1 the transpose of a row-major array is already contiguous, so a view
    would hand the caller's own array (and, for the same array passed twice,
    one buffer for two arguments) to the C side."""
    return np.array(np.asarray(a, dtype=np.float64).T, order="C", copy=True).reshape(-1)


def rowmajor(flat: np.ndarray, n: int) -> np.ndarray:
    return flat.reshape(3, n).T.copy()


class System:
    """Particle set + constants for the oracle. ``props`` holds per-particle
    arrays r_mag, r_hydr, vol_mag, mag_mom, zeta_rot, zeta_trans; ``cfg`` is a
    dict of the configuration constants (names as OrcParams fields)."""

    def __init__(self, n, l_box, alpha, r_cut_lr, r_cut_sr, nc, props, cfg, threads=1):
        self.n = int(n)
        self._arrays = {k: np.ascontiguousarray(props[k], dtype=np.float64) for k in
                        ("r_mag", "r_hydr", "vol_mag", "mag_mom", "zeta_rot", "zeta_trans")}
        p = OrcParams()
        p.n, p.l_box, p.alpha, p.r_cut_lr, p.r_cut_sr, p.nc = self.n, l_box, alpha, r_cut_lr, r_cut_sr, int(nc)
        for k, a in self._arrays.items():
            setattr(p, k, _p(a))
        for name, _ in OrcParams._fields_[12:-2]:
            setattr(p, name, cfg[name])
        p.threads = int(threads)
        self.p = p

    def pairs(self, pos):
        """CompInteractVecs on (N,3) positions -> dict of the pair list and moved positions."""
        L = lib()
        x = colmajor(pos)
        h = L.orc_pairs_build(C.byref(self.p), _p(x))
        P = L.orc_pairs_count(h)
        i = np.zeros(P, np.int64)
        j = np.zeros(P, np.int64)
        r = np.zeros(P)
        u = np.zeros(3 * P)
        sr = np.zeros(P, np.int32)
        L.orc_pairs_get(h, i.ctypes.data_as(C.POINTER(C.c_int64)), j.ctypes.data_as(C.POINTER(C.c_int64)), _p(r),
                        _p(u), sr.ctypes.data_as(C.POINTER(C.c_int32)))
        ov = L.orc_pairs_overlaps(h)
        L.orc_pairs_free(h)
        return {"i": i, "j": j, "r": r, "u": u.reshape(3, P).T.copy(), "sr": sr, "overlaps": ov,
                "pos": rowmajor(x, self.n)}

    def comp_interactions(self, pos, mm, therm_force=None):
        L = lib()
        n = self.n
        x = colmajor(pos)
        m = colmajor(mm)
        tf = None if therm_force is None else colmajor(therm_force)
        F = np.zeros(3 * n)
        B = np.zeros(3 * n)
        P = np.zeros(3)
        L.orc_comp_interactions(C.byref(self.p), _p(x), _p(m), _p(tf), _p(F), _p(B), _p(P), None)
        return rowmajor(F, n), rowmajor(B, n), P, rowmajor(x, n)

    def demag_field(self, pos, mm_stage):
        """CompDemagField for stage moments with the pair list of `pos`."""
        L = lib()
        n = self.n
        x = colmajor(pos)
        xk = x.copy()
        h = L.orc_pairs_build(C.byref(self.p), _p(x))
        m = colmajor(mm_stage)
        B = np.zeros(3 * n)
        L.orc_comp_demag_field(C.byref(self.p), h, _p(xk), _p(m), _p(B))
        L.orc_pairs_free(h)
        return rowmajor(B, n)

    def kspace(self, pos, mom):
        L = lib()
        n = self.n
        x = colmajor(pos)
        m = colmajor(mom)
        ms = np.asarray(mom, dtype=np.float64).sum(axis=0)
        F = np.zeros(3 * n)
        B = np.zeros(3 * n)
        L.orc_kspace(C.byref(self.p), _p(x), _p(m), _p(np.ascontiguousarray(ms)), _p(F), _p(B))
        return rowmajor(F, n), rowmajor(B, n)

    def rk_coeffs(self, bext_z, mm, ea, field, therm_torque, therm_field):
        L = lib()
        n = self.n
        args = [colmajor(a) for a in (mm, ea, field, therm_torque, therm_field)]
        kmm = np.zeros(3 * n)
        kea = np.zeros(3 * n)
        L.orc_rk_coeffs(C.byref(self.p), float(bext_z), *[_p(a) for a in args], _p(kmm), _p(kea))
        return rowmajor(kmm, n), rowmajor(kea, n)

    def step(self, pos, mm, ea, t, dt, n_t, n_b, n_f, max_tries=200):
        """One RunSimulation iteration; returns a dict (state, t, dt, dt_used,
        mz, force, field, press, field_out, rescaled noise, status)."""
        L = lib()
        n = self.n
        x, m, e = colmajor(pos), colmajor(mm), colmajor(ea)
        nt, nb, nf = colmajor(n_t), colmajor(n_b), colmajor(n_f)
        tt = np.array([t], dtype=np.float64)
        dd = np.array([dt], dtype=np.float64)
        du = np.zeros(1)
        mz = np.zeros(1)
        F, B, Bo = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        P = np.zeros(3)
        st = L.orc_step(C.byref(self.p), _p(x), _p(m), _p(e), _p(tt), _p(dd), _p(du), _p(mz), _p(nt), _p(nb),
                        _p(nf), _p(F), _p(B), _p(P), _p(Bo), int(max_tries))
        return {"pos": rowmajor(x, n), "mm": rowmajor(m, n), "ea": rowmajor(e, n), "t": tt[0], "dt": dd[0],
                "dt_used": du[0], "mz": mz[0], "force": rowmajor(F, n), "field": rowmajor(B, n), "press": P,
                "field_out": rowmajor(Bo, n), "n_t": rowmajor(nt, n), "n_b": rowmajor(nb, n), "status": st}


def trans_coeffs(pos, true_pos, pos_ref, l_box, press):
    """EvalTransCoeffs (src/postprocessing.cpp:167-187): unwrap the periodic
    positions against the previous unwrapped ones (Rapaport's minimum-image
    continuation, line 171; Eigen's ``round`` rounds halves away from zero),
    the per-axis mean square displacement from the reference positions
    (173-176) and the step's pressure-tensor sums divided by the box volume
    (179-181). Returns (true_pos, msd[3], p[3])."""
    q = (np.asarray(true_pos) - pos) / l_box
    true_pos = pos + np.where(q >= 0, np.floor(q + 0.5), np.ceil(q - 0.5)) * l_box
    d = true_pos - pos_ref
    return true_pos, (d * d).mean(axis=0), np.asarray(press, dtype=np.float64) / l_box ** 3.0
