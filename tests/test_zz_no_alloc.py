"""The allocation-free <m_z> entry point (mds_mean_z_noalloc) against the
context's own <m_z> (mds_replica_mz) of the same moments, bit for bit, on two
calls in a row (the second with fewer particles, so stale chunk partials of
the first would show), plus its refusals. The test allocates the device
buffers it passes (through the HIP runtime the library links); the entry
point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds
from md_sim_wolfschwenger_amd import capi, workload

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def hip():
    capi.lib()  # loads the HIP runtime the library is linked against
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def upload(h, a: np.ndarray) -> C.c_void_p:
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def system_after_one_step(n: int, seed: int):
    cfg, r, rh, L, pos, mm, ea = workload.system("c1", seed=seed, n=n)
    mu = cfg.sat_mag * 4.0 / 3.0 * cfg.my_pi * r ** 3
    ctx = mds.Context(cfg)
    ctx.set_particles(r, rh, L)
    ctx.set_ewald(*workload.ewald_for(cfg, mu, rh, L, 0.3))
    ctx.set_state(pos, mm, ea)
    ctx.use_philox(seed, 0)
    ctx.step(0.0, 1e-12)
    mz = float(ctx.replica_mz()[0])
    _, mm1, _ = ctx.get_state()
    ctx.close()
    return mz, mm1


def test_mean_z_noalloc_matches_replica_mz_on_two_calls():
    h = hip()
    L = capi.lib()
    for n, seed in ((20000, 3), (3000, 4)):
        print(f"noalloc <m_z> n={n}", flush=True)
        mz, mm = system_after_one_step(n, seed)
        d_mm = upload(h, capi.cm(mm))
        d_out = upload(h, np.full(1, np.nan))
        assert L.mds_mean_z_noalloc(d_mm, n, d_out, None) == 0
        out = np.zeros(1)
        assert h.hipMemcpy(out.ctypes.data, d_out, 8, D2H) == 0  # null stream: after the kernels
        assert out[0] == mz
        assert h.hipFree(d_mm) == 0 and h.hipFree(d_out) == 0


def test_mean_z_noalloc_refuses_bad_sizes():
    h = hip()
    L = capi.lib()
    d = upload(h, np.zeros(3))
    assert L.mds_mean_z_noalloc(d, 0, d, None) == -1
    assert L.mds_mean_z_noalloc(None, 1, d, None) == -1
    assert L.mds_mean_z_noalloc(d, (1 << 27) + 1, d, None) == -8
    assert h.hipFree(d) == 0
