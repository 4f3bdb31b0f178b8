"""ctypes binding of libmdsim.so (include/mdsim_capi.h).

The product is the C ABI; this module is the thin Python mirror of it that
the tests and bench.py drive. Arrays cross the boundary as (N, 3) float64
numpy arrays here and as column-major N x 3 buffers in C, the layout of the
reference's Eigen arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdsim.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)

STATUS = {0: "MDS_OK", -1: "MDS_EINVAL", -2: "MDS_ENOMEM", -3: "MDS_EDEVICE", -4: "MDS_ECONFIG",
          -5: "MDS_ENOCONV", -6: "MDS_EOVERLAP", -7: "MDS_ENONFINITE", -8: "MDS_ECAPACITY"}
HEUN_EULER, BOGACKI_SHAMPINE, DORMAND_PRINCE = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1


class MdsError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    """Mirror of mds_config (field order and types as in the header)."""
    _fields_ = [
        ("init_config_rot", C.c_int32), ("init_config_trans", C.c_int32), ("read_from_file", C.c_int32),
        ("size_dist", C.c_int32), ("enable_mobilization", C.c_int32),
        ("enable_therm_field", C.c_int32), ("enable_therm_torque", C.c_int32), ("enable_therm_force", C.c_int32),
        ("seed", C.c_int64), ("enable_coating_pot", C.c_int32), ("solver", C.c_int32), ("np_dim", C.c_int32),
        ("mag_damp", C.c_double), ("temp", C.c_double), ("sat_mag", C.c_double), ("anis_en", C.c_double),
        ("r_mag_mean", C.c_double), ("d_shell", C.c_double), ("vol_frac", C.c_double),
        ("dens_surf_mol", C.c_double), ("hamaker", C.c_double), ("z_valency", C.c_double),
        ("epsilon_r", C.c_double), ("gamma_0", C.c_double), ("conc", C.c_double),
        ("t_mag", C.c_double), ("t_relax", C.c_double), ("delta_t_init", C.c_double),
        ("mag_flux_dens", C.c_double),
        ("abs_tol", C.c_double), ("delta_t_min", C.c_double), ("delta_t_max", C.c_double),
        ("err_tol_ewald", C.c_double), ("err_tol_sr", C.c_double), ("np_coords", C.c_double),
        ("my_pi", C.c_double), ("kB", C.c_double), ("gyro_mr", C.c_double), ("avogadro", C.c_double),
        ("epsilon_0", C.c_double), ("el", C.c_double), ("fac_ewald", C.c_double),
        ("numb_part", C.c_int64),
        ("mu0", C.c_double), ("eff_len_surf_mol", C.c_double), ("vel_mm_const", C.c_double),
        ("anis_const", C.c_double), ("ster_rep_fac", C.c_double), ("gamma_el", C.c_double),
        ("epsilon", C.c_double), ("kappa", C.c_double), ("electrostat_rep_fac", C.c_double),
        ("ir_pi", C.c_double), ("vis", C.c_double), ("sigma", C.c_double), ("t_end", C.c_double),
        ("numb_part_override", C.c_int64), ("viscosity", C.c_double), ("r_mag_median", C.c_double),
        ("sigma_log", C.c_double), ("fixed_dt", C.c_int32), ("precision", C.c_int32),
        ("ewald_alpha_l", C.c_double), ("ewald_rc_frac", C.c_double), ("ewald_nc", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported entry point of the header
SIGNATURES = [
    ("mds_config_from_ini", C.c_int, [C.c_char_p, C.POINTER(Config)]),
    ("mds_config_from_string", C.c_int, [C.c_char_p, C.POINTER(Config)]),
    ("mds_config_derive", None, [C.POINTER(Config)]),
    ("mds_config_sizeof", C.c_int64, []),
    ("mds_particle_props", C.c_int, [C.POINTER(Config), C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("mds_box_length", C.c_double, [C.POINTER(Config), C.c_int64, _dp]),
    ("mds_cutoff_sr", C.c_int, [C.POINTER(Config), C.c_int64, _dp, _dp, _dp, _dp]),
    ("mds_ewald_real_error", C.c_double, [C.POINTER(Config), C.c_int64, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double]),
    ("mds_ewald_rez_error", C.c_double, [C.POINTER(Config), C.c_int64, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double]),
    ("mds_ewald_params_for_rc", C.c_int, [C.POINTER(Config), C.c_int64, _dp, _dp, C.c_double, C.c_double, _dp,
                                          _dp]),
    ("mds_ewald_predicted_errors", C.c_int, [C.POINTER(Config), C.c_int64, _dp, _dp, C.c_double, C.c_double,
                                             C.c_double, C.c_int32, _dp, _dp]),
    ("mds_init_radii", C.c_int, [C.POINTER(Config), C.c_int64, C.c_uint64, _dp]),
    ("mds_init_positions", C.c_int, [C.POINTER(Config), C.c_int64, _dp, C.c_double, C.c_uint64, C.c_int32, _dp]),
    ("mds_init_orientations", C.c_int, [C.POINTER(Config), C.c_int64, C.c_uint64, _dp, _dp]),
    ("mds_init_coords_from_sim", C.c_int, [C.c_char_p, C.c_int64, _dp, _dp, _dp, _dp]),
    ("mds_init_suspension_device", C.c_int, [C.POINTER(Config), C.c_int, C.c_int64, C.c_uint64, C.c_int32, _dp, _dp,
                                             _dp, _dp, _dp]),
    ("mds_create", C.c_int, [C.POINTER(Config), C.c_int, C.POINTER(C.c_void_p)]),
    ("mds_destroy", None, [C.c_void_p]),
    ("mds_last_error", C.c_char_p, [C.c_void_p]),
    ("mds_device_count", C.c_int, []),
    ("mds_set_particles", C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, C.c_double]),
    ("mds_set_ewald", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32]),
    ("mds_set_tiling", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_get_tiling", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("mds_set_pair_list", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("mds_set_kspace_waves", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_pair_list_words", C.c_int64, [C.c_int64, C.c_double, C.c_double, C.c_double]),
    ("mds_pair_list_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]),
    ("mds_autotune", C.c_int, [C.c_void_p, _dp, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_int32, _dp,
                               C.POINTER(C.c_int32), _dp]),
    ("mds_get_params", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32)]),
    ("mds_set_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("mds_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("mds_set_noise", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("mds_use_philox", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    ("mds_get_noise", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("mds_set_check_interval", C.c_int, [C.c_void_p, C.c_int64]),
    ("mds_chunks_redone", C.c_int64, [C.c_void_p]),
    ("mds_mean_z_noalloc", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("mds_philox4x32_10", None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("mds_philox_normals", None, [C.c_uint64, C.c_int64, C.c_uint64, _dp]),
    ("mds_comp_interactions", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("mds_comp_demag_field", C.c_int, [C.c_void_p, _dp, _dp]),
    ("mds_rk_coeffs", C.c_int, [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp, _dp]),
    ("mds_step", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("mds_run", C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _dp]),
    ("mds_observe", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_run_observed", C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp]),
    ("mds_last_observables", C.c_int, [C.c_void_p, _dp]),
    ("mds_save_checkpoint", C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double]),
    ("mds_load_checkpoint", C.c_int, [C.c_void_p, C.c_char_p, _dp, _dp]),
    ("mds_synchronize", C.c_int, [C.c_void_p]),
    ("mds_enable_timing", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_kernel_class_count", C.c_int32, []),
    ("mds_kernel_class_name", C.c_char_p, [C.c_int32]),
    ("mds_kernel_stats", C.c_int, [C.c_void_p, _i64p, _dp]),
    ("mds_work_counts", C.c_int, [C.c_void_p, _i64p, _i64p, _i64p, _i64p]),
    ("mds_add_replica", C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_uint64, _dp, _dp,
                                  _dp, C.POINTER(C.c_int32)]),
    ("mds_add_replicas", C.c_int, [C.c_void_p, C.c_int32, _i64p, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint64), _dp,
                                   _dp, _dp, C.POINTER(C.c_int32)]),
    ("mds_remove_replica", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_replica_count", C.c_int32, [C.c_void_p]),
    ("mds_replica_set_ewald", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_int32]),
    ("mds_replica_info", C.c_int, [C.c_void_p, C.c_int32, _i64p, _i64p, _dp, _dp, _dp, C.POINTER(C.c_int32), _dp,
                                   C.POINTER(C.c_uint64)]),
    ("mds_replica_get_state", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp]),
    ("mds_replica_mz", C.c_int, [C.c_void_p, _dp]),
    ("mds_replica_stop_at", C.c_int, [C.c_void_p, C.c_int32, C.c_uint64]),
    ("mds_reserve", C.c_int, [C.c_void_p, C.c_int64]),
    ("mds_set_cu_mask", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_int32]),
    ("mds_device_cus", C.c_int32, [C.c_int]),
    ("mds_peak_probe", C.c_int, [C.c_int, _dp, _dp, _dp]),
    ("mds_set_graph", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_set_kspace_overlap", C.c_int, [C.c_void_p, C.c_int32]),
    ("mds_device_bytes", C.c_int64, [C.c_void_p]),
]

_lib = None


def lib():
    """Load libmdsim.so; raises if it is not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `make -C {HERE}` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.mds_config_sizeof() != C.sizeof(Config):
            raise ImportError("mds_config layout mismatch between libmdsim.so and capi.Config")
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def cm(a) -> np.ndarray:
    """(N,3) -> a new contiguous column-major flat buffer (a copy also for
    N = 1, where the transpose would already be contiguous)."""
    return np.array(np.asarray(a, dtype=np.float64).T, order="C", copy=True).reshape(-1)


def rm(flat, n) -> np.ndarray:
    return flat.reshape(3, n).T.copy()


def config_from_string(text: str) -> Config:
    c = Config()
    st = lib().mds_config_from_string(text.encode(), C.byref(c))
    if st:
        raise MdsError(st, "configuration text")
    return c


def config_from_ini(path: str) -> Config:
    c = Config()
    st = lib().mds_config_from_ini(path.encode(), C.byref(c))
    if st:
        raise MdsError(st, path)
    return c


def init_radii(cfg: Config, n: int, seed: int) -> np.ndarray:
    """Core radii of n particles (InitParticleProps' size distributions)."""
    r = np.zeros(n)
    st = lib().mds_init_radii(C.byref(cfg), n, int(seed) & 0xFFFFFFFFFFFFFFFF, _p(r))
    if st:
        raise MdsError(st, "init_radii")
    return r


def init_positions(cfg: Config, r_hydr, l_box: float, seed: int, periodic: bool = True) -> np.ndarray:
    """(N, 3) initial positions (InitCoordsFromSketch's translational part)."""
    rh = np.ascontiguousarray(r_hydr, dtype=np.float64)
    n = len(rh)
    p = np.zeros(3 * n)
    st = lib().mds_init_positions(C.byref(cfg), n, _p(rh), float(l_box), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                  1 if periodic else 0, _p(p))
    if st:
        raise MdsError(st, "init_positions")
    return rm(p, n)


def init_orientations(cfg: Config, n: int, seed: int):
    """(pos_mm, pos_ea), each (N, 3) (InitCoordsFromSketch's rotational part)."""
    mm, ea = np.zeros(3 * n), np.zeros(3 * n)
    st = lib().mds_init_orientations(C.byref(cfg), n, int(seed) & 0xFFFFFFFFFFFFFFFF, _p(mm), _p(ea))
    if st:
        raise MdsError(st, "init_orientations")
    return rm(mm, n), rm(ea, n)


def init_suspension_device(cfg: Config, n: int, seed: int, periodic: bool = True, l_box: float = 0.0,
                           device: int = 0):
    """(r_mag, L, pos, mm, ea) of an initial suspension drawn on the device
    (mds_init_suspension_device)."""
    L = C.c_double(l_box)
    r, pos, mm, ea = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
    st = lib().mds_init_suspension_device(C.byref(cfg), int(device), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                          1 if periodic else 0, C.byref(L), _p(r), _p(pos), _p(mm), _p(ea))
    if st:
        raise MdsError(st, "init_suspension_device")
    return r, L.value, rm(pos, n), rm(mm, n), rm(ea, n)


def init_coords_from_sim(path: str, n: int):
    """READ_FROM_FILE: (pos, mm, ea, t) from the last N x 10 block of a
    WriteCoords2TXT file (mds_init_coords_from_sim)."""
    pos, mm, ea, t = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n), C.c_double()
    st = lib().mds_init_coords_from_sim(os.fsencode(path), int(n), _p(pos), _p(mm), _p(ea), C.byref(t))
    if st != 0:
        raise MdsError(st, f"cannot read {n} x 10 coordinates from {path}")
    return rm(pos, n), rm(mm, n), rm(ea, n), t.value


def ewald_predicted_errors(cfg: Config, mag_mom, r_hydr, l_box: float, alpha: float, r_cut: float, nc: int):
    """(force, torque) rms errors predicted by the reference's error model,
    relative to the contact force and torque of the smallest particles."""
    mu = np.ascontiguousarray(mag_mom, dtype=np.float64)
    rh = np.ascontiguousarray(r_hydr, dtype=np.float64)
    f, t = C.c_double(), C.c_double()
    st = lib().mds_ewald_predicted_errors(C.byref(cfg), len(mu), _p(mu), _p(rh), float(l_box), float(alpha),
                                          float(r_cut), int(nc), C.byref(f), C.byref(t))
    if st:
        raise MdsError(st, "ewald_predicted_errors")
    return f.value, t.value


def philox4x32_10(ctr, key) -> tuple:
    """One Philox4x32-10 block of the host copy of the in-kernel generator."""
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().mds_philox4x32_10(c, k, o)
    return tuple(int(x) for x in o)


def philox_normals(seed: int, index: int, step: int) -> np.ndarray:
    """The nine unscaled normals (torque, field, force; xyz each) the noise
    kernel draws for particle `index` of a replica at its step `step`."""
    z = np.zeros(9)
    lib().mds_philox_normals(int(seed), int(index), int(step), _p(z))
    return z


def peak_probe(device: int = 0) -> dict:
    """Measured FP64 / FP32 vector FMA TFLOP/s and device copy GB/s."""
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    st = lib().mds_peak_probe(int(device), C.byref(a), C.byref(b), C.byref(c))
    if st != 0:
        raise MdsError(st, "peak probe")
    return {"fp64_vector_tflops": a.value, "fp32_vector_tflops": b.value, "copy_gbs": c.value}


def device_count() -> int:
    return lib().mds_device_count()


class Context:
    """One device context (mds_ctx)."""

    def __init__(self, cfg: Config, device: int = 0):
        self._L = lib()
        self.cfg = cfg
        h = C.c_void_p()
        st = self._L.mds_create(C.byref(cfg), device, C.byref(h))
        if st:
            raise MdsError(st, "mds_create")
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            self._L.mds_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, what):
        if st:
            raise MdsError(st, f"{what}: {self._L.mds_last_error(self.h).decode()}")

    def set_particles(self, r_mag, r_hydr=None, l_box=0.0):
        r_mag = np.ascontiguousarray(r_mag, dtype=np.float64)
        rh = None if r_hydr is None else np.ascontiguousarray(r_hydr, dtype=np.float64)
        self._chk(self._L.mds_set_particles(self.h, len(r_mag), _p(r_mag), _p(rh), float(l_box)), "set_particles")
        self.n = len(r_mag)

    def set_ewald(self, alpha, r_cut, nc):
        self._chk(self._L.mds_set_ewald(self.h, float(alpha), float(r_cut), int(nc)), "set_ewald")

    def set_tiling(self, cell_reach):
        self._chk(self._L.mds_set_tiling(self.h, int(cell_reach)), "set_tiling")

    def tiling(self):
        v = [C.c_int32() for _ in range(3)]
        self._chk(self._L.mds_get_tiling(self.h, *[C.byref(x) for x in v]), "get_tiling")
        return {"cell_reach": v[0].value, "ncell": v[1].value, "nstencil": v[2].value}

    def set_kspace_waves(self, waves):
        """Waves per SIMD the S(k) build may hold beside the pair kernels (0: no limit;
        -1, the default: one for the matrix-core build, no limit for the vector build)."""
        self._chk(self._L.mds_set_kspace_waves(self.h, int(waves)), "set_kspace_waves")

    def set_pair_list(self, skin, region_scale=1.0):
        """Pair lists with skin r_c * skin (0: search the cells every evaluation);
        region_scale sizes the list regions against the planner's estimate."""
        self._chk(self._L.mds_set_pair_list(self.h, float(skin), float(region_scale)), "set_pair_list")

    @staticmethod
    def pair_list_words(n, l_box, r_cut, skin=0.005):
        """Words of the list region the planner gives such a replica."""
        return int(lib().mds_pair_list_words(int(n), float(l_box), float(r_cut), float(skin)))

    def pair_list_stats(self, index=0):
        v = [C.c_int64() for _ in range(3)]
        self._chk(self._L.mds_pair_list_stats(self.h, int(index), *[C.byref(x) for x in v]), "pair_list_stats")
        return {"builds": v[0].value, "words": v[1].value, "capacity": v[2].value}

    def autotune(self, rc_fracs=(0.175, 0.2, 0.225, 0.25, 0.275, 0.3), reaches=(2, 3, 4), reps=2):
        """Time each (r_c, cell reach) candidate on the device; install the fastest."""
        rcs = np.ascontiguousarray(rc_fracs, dtype=np.float64)
        rea = np.ascontiguousarray(reaches, dtype=np.int32)
        brc, bre, bms = C.c_double(), C.c_int32(), C.c_double()
        self._chk(self._L.mds_autotune(self.h, _p(rcs), len(rcs), rea.ctypes.data_as(C.POINTER(C.c_int32)), len(rea),
                                       int(reps), C.byref(brc), C.byref(bre), C.byref(bms)), "autotune")
        return {"rc_frac": brc.value, "reach": bre.value, "eval_ms": bms.value}

    def params(self):
        v = [C.c_double() for _ in range(4)]
        nc = C.c_int32()
        self._chk(self._L.mds_get_params(self.h, *[C.byref(x) for x in v], C.byref(nc)), "get_params")
        return {"l_box": v[0].value, "r_cut_sr": v[1].value, "alpha": v[2].value, "r_cut": v[3].value,
                "nc": nc.value}

    def set_state(self, pos, mm, ea):
        a, b, c = cm(pos), cm(mm), cm(ea)
        self._chk(self._L.mds_set_state(self.h, _p(a), _p(b), _p(c)), "set_state")

    def get_state(self):
        n = self.n
        a, b, c = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        self._chk(self._L.mds_get_state(self.h, _p(a), _p(b), _p(c)), "get_state")
        return rm(a, n), rm(b, n), rm(c, n)

    def set_noise(self, n_t=None, n_b=None, n_f=None):
        arrs = [None if x is None else cm(x) for x in (n_t, n_b, n_f)]
        self._chk(self._L.mds_set_noise(self.h, *[_p(x) for x in arrs]), "set_noise")

    def use_philox(self, seed, counter=0):
        self._chk(self._L.mds_use_philox(self.h, int(seed), int(counter)), "use_philox")

    def get_noise(self):
        """(torque, field, force) thermal draws of the last step, N x 3."""
        n = self.n
        a, b, c = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        self._chk(self._L.mds_get_noise(self.h, _p(a), _p(b), _p(c)), "get_noise")
        return rm(a, n), rm(b, n), rm(c, n)

    def comp_interactions(self):
        n = self.n
        F, B, P, X = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3), np.zeros(3 * n)
        self._chk(self._L.mds_comp_interactions(self.h, _p(F), _p(B), _p(P), _p(X)), "comp_interactions")
        return rm(F, n), rm(B, n), P, rm(X, n)

    def comp_demag_field(self, mm):
        n = self.n
        m = cm(mm)
        B = np.zeros(3 * n)
        self._chk(self._L.mds_comp_demag_field(self.h, _p(m), _p(B)), "comp_demag_field")
        return rm(B, n)

    def rk_coeffs(self, t, mm, ea, field):
        n = self.n
        m, e, b = cm(mm), cm(ea), cm(field)
        km, ke = np.zeros(3 * n), np.zeros(3 * n)
        self._chk(self._L.mds_rk_coeffs(self.h, float(t), _p(m), _p(e), _p(b), _p(km), _p(ke)), "rk_coeffs")
        return rm(km, n), rm(ke, n)

    def step(self, t, dt, want_mz=True):
        tt, dd, du, mz = (C.c_double(t), C.c_double(dt), C.c_double(), C.c_double())
        self._chk(self._L.mds_step(self.h, C.byref(tt), C.byref(dd), C.byref(du),
                                   C.byref(mz) if want_mz else None), "step")
        return tt.value, dd.value, du.value, (mz.value if want_mz else None)

    def run(self, steps, t, dt, want_mz=False):
        """`steps` steps; with want_mz the <m_z> series (steps,) for one
        replica, (steps, replicas) for several."""
        tt, dd = C.c_double(t), C.c_double(dt)
        R = max(1, self.replica_count())
        mz = np.zeros(steps * R) if want_mz else None
        self._chk(self._L.mds_run(self.h, int(steps), C.byref(tt), C.byref(dd), _p(mz)), "run")
        if mz is not None and R > 1:
            mz = mz.reshape(steps, R)
        return tt.value, dd.value, mz

    # fields of the per-step transport observables (mds_observe)
    OBSERVABLES = ("t", "msd_x", "msd_y", "msd_z", "p_xy", "p_xz", "p_yz")

    def observe(self, on=True):
        """EvalTransCoeffs observables from now on, with the current
        positions as the reference (see OBSERVABLES)."""
        self._chk(self._L.mds_observe(self.h, 1 if on else 0), "observe")

    def run_observed(self, steps, t, dt):
        """run() that also returns <m_z> and the observables per step:
        mz (steps,) and obs (steps, 7) for one replica, (steps, R) and
        (steps, R, 7) for several."""
        tt, dd = C.c_double(t), C.c_double(dt)
        R = max(1, self.replica_count())
        mz = np.zeros(steps * R)
        obs = np.zeros(steps * R * 7)
        self._chk(self._L.mds_run_observed(self.h, int(steps), C.byref(tt), C.byref(dd), _p(mz), _p(obs)),
                  "run_observed")
        if R > 1:
            return tt.value, dd.value, mz.reshape(steps, R), obs.reshape(steps, R, 7)
        return tt.value, dd.value, mz, obs.reshape(steps, 7)

    def last_observables(self):
        R = max(1, self.replica_count())
        obs = np.zeros(R * 7)
        self._chk(self._L.mds_last_observables(self.h, _p(obs)), "last_observables")
        return obs.reshape(R, 7) if R > 1 else obs

    def save_checkpoint(self, path, t, dt):
        """Binary checkpoint of the replicas, state, clocks, noise counters
        and observables reference, with the caller's t and dt."""
        self._chk(self._L.mds_save_checkpoint(self.h, os.fsencode(path), float(t), float(dt)), "save_checkpoint")

    def load_checkpoint(self, path):
        """Replace this context's replicas by a checkpoint's; returns (t, dt)."""
        t, dt = C.c_double(), C.c_double()
        self._chk(self._L.mds_load_checkpoint(self.h, os.fsencode(path), C.byref(t), C.byref(dt)), "load_checkpoint")
        # the particle count of the arrays this wrapper allocates for the state
        self.n = sum(self.replica_info(r)["n"] for r in range(self.replica_count()))
        return t.value, dt.value

    def synchronize(self):
        self._chk(self._L.mds_synchronize(self.h), "synchronize")

    def enable_timing(self, on=True):
        self._chk(self._L.mds_enable_timing(self.h, 1 if on else 0), "enable_timing")

    def kernel_stats(self):
        k = self._L.mds_kernel_class_count()
        la = np.zeros(k, np.int64)
        ms = np.zeros(k)
        self._chk(self._L.mds_kernel_stats(self.h, la.ctypes.data_as(_i64p), _p(ms)), "kernel_stats")
        return {self._L.mds_kernel_class_name(i).decode(): (int(la[i]), float(ms[i])) for i in range(k)}

    def work_counts(self):
        v = [C.c_int64() for _ in range(4)]
        self._chk(self._L.mds_work_counts(self.h, *[C.byref(x) for x in v]), "work_counts")
        return {"pair_evals": v[0].value, "candidates": v[1].value, "kvecs": v[2].value,
                "overlap_events": v[3].value}

    # ------------------------------------------------------------ replicas
    def add_replica(self, r_mag, r_hydr=None, l_box=0.0, field_z=0.0, seed=0, pos=None, mm=None, ea=None):
        """Append a replica; returns its index."""
        r_mag = np.ascontiguousarray(r_mag, dtype=np.float64)
        rh = None if r_hydr is None else np.ascontiguousarray(r_hydr, dtype=np.float64)
        st = [None if x is None else cm(x) for x in (pos, mm, ea)]
        idx = C.c_int32()
        self._chk(self._L.mds_add_replica(self.h, len(r_mag), _p(r_mag), _p(rh), float(l_box), float(field_z),
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, *[_p(x) for x in st], C.byref(idx)),
                  "add_replica")
        self.n += len(r_mag)
        return idx.value

    def add_replicas(self, systems):
        """Append several replicas in one call; `systems` is a list of dicts
        with r_mag and optionally r_hydr, l_box, field_z, seed, pos, mm, ea.
        Returns the first new index."""
        ns = np.array([len(d["r_mag"]) for d in systems], np.int64)
        cat = lambda key: np.ascontiguousarray(np.concatenate([np.asarray(d[key], np.float64) for d in systems]))
        r = cat("r_mag")
        rh = cat("r_hydr") if all(d.get("r_hydr") is not None for d in systems) else None
        lb = np.array([float(d.get("l_box", 0.0)) for d in systems])
        fz = np.array([float(d.get("field_z", self.cfg.mag_flux_dens)) for d in systems])
        sd = np.array([int(d.get("seed", 0)) & 0xFFFFFFFFFFFFFFFF for d in systems], np.uint64)
        have = all(d.get("pos") is not None for d in systems)
        st = [cm(cat(k)) if have else None for k in ("pos", "mm", "ea")]
        first = C.c_int32()
        self._chk(self._L.mds_add_replicas(self.h, len(systems), ns.ctypes.data_as(_i64p), _p(r), _p(rh), _p(lb),
                                           _p(fz), sd.ctypes.data_as(C.POINTER(C.c_uint64)), *[_p(x) for x in st],
                                           C.byref(first)), "add_replicas")
        self.n += int(ns.sum())
        return first.value

    def remove_replica(self, index):
        n = self.replica_info(index)["n"]
        self._chk(self._L.mds_remove_replica(self.h, int(index)), "remove_replica")
        self.n -= n

    def replica_count(self):
        return int(self._L.mds_replica_count(self.h))

    def replica_set_ewald(self, index, alpha, r_cut, nc):
        self._chk(self._L.mds_replica_set_ewald(self.h, int(index), float(alpha), float(r_cut), int(nc)),
                  "replica_set_ewald")

    def replica_info(self, index, clock=False):
        n, off = C.c_int64(), C.c_int64()
        L, a, rc, t = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        nc = C.c_int32()
        steps = C.c_uint64()
        self._chk(self._L.mds_replica_info(self.h, int(index), C.byref(n), C.byref(off), C.byref(L), C.byref(a),
                                           C.byref(rc), C.byref(nc), C.byref(t) if clock else None,
                                           C.byref(steps) if clock else None), "replica_info")
        d = {"n": n.value, "offset": off.value, "l_box": L.value, "alpha": a.value, "r_cut": rc.value,
             "nc": nc.value}
        if clock:
            d.update(t=t.value, steps=steps.value)
        return d

    def replica_state(self, index):
        n = self.replica_info(index)["n"]
        a, b, c = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        self._chk(self._L.mds_replica_get_state(self.h, int(index), _p(a), _p(b), _p(c)), "replica_get_state")
        return rm(a, n), rm(b, n), rm(c, n)

    def replica_stop_at(self, index, step):
        """Replica `index` stops stepping at its step count `step`."""
        self._chk(self._L.mds_replica_stop_at(self.h, int(index), int(step)), "replica_stop_at")

    def replica_mz(self):
        mz = np.zeros(self.replica_count())
        self._chk(self._L.mds_replica_mz(self.h, _p(mz)), "replica_mz")
        return mz

    def reserve(self, n_particles):
        self._chk(self._L.mds_reserve(self.h, int(n_particles)), "reserve")

    def set_cu_mask(self, cus=None):
        """Run on the given compute-unit ids only (None: all)."""
        if not cus:
            self._chk(self._L.mds_set_cu_mask(self.h, None, 0), "set_cu_mask")
            return
        words = (max(cus) // 32) + 1
        m = np.zeros(words, np.uint32)
        for c in cus:
            m[c // 32] |= np.uint32(1 << (c % 32))
        self._chk(self._L.mds_set_cu_mask(self.h, m.ctypes.data_as(C.POINTER(C.c_uint32)), words), "set_cu_mask")

    def set_kspace_overlap(self, on=True):
        """S(k) on a second stream beside the pair kernels (default on)."""
        self._chk(self._L.mds_set_kspace_overlap(self.h, 1 if on else 0), "set_kspace_overlap")

    def set_check_interval(self, steps):
        """Graph-replayed steps per checked chunk (see mds_set_check_interval)."""
        self._chk(self._L.mds_set_check_interval(self.h, int(steps)), "set_check_interval")

    def chunks_redone(self):
        return int(self._L.mds_chunks_redone(self.h))

    def set_graph(self, on=True):
        self._chk(self._L.mds_set_graph(self.h, 1 if on else 0), "set_graph")

    def device_bytes(self):
        return int(self._L.mds_device_bytes(self.h))
