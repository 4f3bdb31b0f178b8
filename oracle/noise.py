"""Numpy restatement of the library's in-kernel thermal noise.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the noise
kernel; never by the product package.

The reference draws its thermal torque, field and force from EigenRand's
Mersenne twister (src/simulation.cpp:570, 584, 596); the library replaces
that stream with counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
SC'11) and Box-Muller. This module restates the generator independently of
the library's C code, vectorised over particles, so a test can compare the
device's draws with it exactly, and restates the reference's per-particle
scales sigma_T, sigma_B, sigma_F (src/simulation.cpp:567, 582, 594) so the
draws' distribution can be checked against them.
"""
from __future__ import annotations

import numpy as np

_MASK = np.uint64(0xFFFFFFFF)
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)


def philox4x32_10(ctr, key):
    """Philox4x32-10 on arrays: ctr is 4 arrays (or ints) of 32-bit words,
    key 2 of them; returns the 4 output words as uint64 arrays."""
    c0, c1, c2, c3 = [np.asarray(x, dtype=np.uint64) & _MASK for x in ctr]
    k0, k1 = [np.asarray(x, dtype=np.uint64) & _MASK for x in key]
    for _ in range(10):
        p0 = _M0 * c0  # < 2^64: both factors are below 2^32
        p1 = _M1 * c2
        c0, c1, c2, c3 = (p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & _MASK, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & _MASK
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def uniforms(seed: int, index, step: int, call: int):
    """(u1 in (0, 1], u2 in [0, 1)) of block `call` for particles `index`."""
    li = np.asarray(index, dtype=np.uint64)
    st = np.uint64(step)
    ctr = (li & _MASK, (li >> np.uint64(32)) ^ np.uint64(call << 24), st & _MASK, st >> np.uint64(32))
    key = (np.uint64(seed) & _MASK, np.uint64(seed) >> np.uint64(32))
    o0, o1, o2, o3 = philox4x32_10(ctr, key)
    a = (o0 << np.uint64(32)) | o1
    b = (o2 << np.uint64(32)) | o3
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return u1, u2


def normals(seed: int, index, step: int) -> np.ndarray:
    """(n, 9) unscaled normals: torque xyz, field xyz, force xyz."""
    out = []
    for call in range(5):
        u1, u2 = uniforms(seed, index, step, call)
        r = np.sqrt(-2.0 * np.log(u1))
        out += [r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)]
    return np.stack(out[:9], axis=1)


def sigmas(kB: float, temp: float, dt: float, mag_damp: float, gyro: float, sat_mag: float, zeta_rot, vol_mag,
           zeta_trans):
    """DefThermFluct's per-particle scales (src/simulation.cpp:567, 582, 594):
    sigma_T = sqrt(2 kB T zeta_rot / dt), sigma_B = sqrt(2 kB T alpha /
    (gamma Ms dt V_m)), sigma_F = sqrt(2 kB T zeta_trans / dt)."""
    s_t = np.sqrt((2.0 * kB * temp / dt) * np.asarray(zeta_rot))
    s_b = np.sqrt(2.0 * kB * temp * mag_damp / (gyro * sat_mag * dt) / np.asarray(vol_mag))
    s_f = np.sqrt((2.0 * kB * temp / dt) * np.asarray(zeta_trans))
    return s_t, s_b, s_f
