// Counter-based thermal noise shared by the noise kernel and the host.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// round constants 0xD2511F53 / 0xCD9E8D57 and the Weyl key bumps 0x9E3779B9 /
// 0xBB67AE85. The same code is compiled for gfx950 and for the host, so the
// host copy (exported as mds_philox4x32_10) is what the known-answer vectors
// of the published generator check.
//
// A replica's draw for particle li (index within the replica) at accepted
// step `step` takes five Philox blocks, call = 0..4, with counter
// (li lo, li hi ^ call << 24, step lo, step hi) and key (seed lo, seed hi);
// each block gives two 53-bit uniforms u1 in (0, 1], u2 in [0, 1) and
// Box-Muller turns them into two standard normals sqrt(-2 ln u1) (cos, sin)
// (2 pi u2). Normals 0-2 scale the torque, 3-5 the field and 6-8 the force
// (the order DefThermFluct draws them, src/simulation.cpp:563-603); the
// tenth is unused.
#ifndef MDSIM_PHILOX_H_
#define MDSIM_PHILOX_H_

#include <stdint.h>

#ifdef __HIPCC__
#define MDS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define MDS_HD inline
#endif

MDS_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// The two uniforms of one block: high and low 64-bit halves, top 53 bits.
MDS_HD void philox_uniforms(uint64_t seed, int64_t li, uint64_t step, int call, double* u1, double* u2) {
    uint32_t c[4] = {(uint32_t)li, (uint32_t)((uint64_t)li >> 32) ^ ((uint32_t)call << 24), (uint32_t)step,
                     (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)c[0] << 32) | c[1], b = ((uint64_t)c[2] << 32) | c[3];
    *u1 = (double)((a >> 11) + 1) * 0x1.0p-53;  // (0, 1]
    *u2 = (double)(b >> 11) * 0x1.0p-53;        // [0, 1)
}

// Nine standard normals of particle li at step `step` (z[9] unused slot kept
// out: the caller passes room for 10).
MDS_HD void philox_normals(uint64_t seed, int64_t li, uint64_t step, double z[10]) {
    for (int call = 0; call < 5; ++call) {
        double u1, u2;
        philox_uniforms(seed, li, step, call, &u1, &u2);
        const double rr = sqrt(-2.0 * log(u1));
        double s, co;
#ifdef __HIP_DEVICE_COMPILE__
        sincospi(2.0 * u2, &s, &co);
#else
        // 2 pi u2 in [0, 2 pi); the device's sincospi differs by an ulp or so
        s = sin(6.283185307179586476925 * u2);
        co = cos(6.283185307179586476925 * u2);
#endif
        z[2 * call] = rr * co;
        z[2 * call + 1] = rr * s;
    }
}

#undef MDS_HD
#endif  // MDSIM_PHILOX_H_
