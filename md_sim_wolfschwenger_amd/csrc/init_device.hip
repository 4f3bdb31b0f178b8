// init_device.hip - initial suspensions drawn on the device.
//
// InitParticleProps' radii (src/simulation.cpp:132-214) and
// InitCoordsFromSketch (src/simulation.cpp:220-351) as device kernels over
// the library's Philox4x32-10 stream (philox.h), for systems of up to 2e5
// particles that a serving loop sets up per request. The distributions are
// those of the host functions in init.cpp (mds_init_radii, _positions,
// _orientations); the numbers differ.
//
// Positions: the reference adds particle i at a uniform random position and
// rejects it while it is closer than r_i + r_j to an earlier particle
// (src/simulation.cpp:241-264), one particle after the other. Here the
// addition runs in rounds: every unplaced particle proposes a position; a
// proposal that overlaps no placed particle and no proposal of a lower index
// of the same round is accepted, the others propose again in the next round.
// Every accepted particle therefore keeps the reference's rule against all
// particles placed before it and all accepted with it, so no two particles
// overlap; proposals and overlaps are found through a cell grid of edge
// >= 2 r_max (the placed particles' grid, then the proposals').

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "mdsim_capi.h"
#include "philox.h"

namespace {

// independent streams of one seed for the radii, positions and orientations
uint64_t stream_seed(uint64_t seed, uint64_t salt) {
    uint64_t z = seed + salt * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// r = a (mode 0), a exp(b z) (mode 1: median a, sigma b of ln r) or
// exp(a + b z) (mode 2: mu, sigma of ln r), z standard normal
__global__ void k_init_radii(int64_t n, uint64_t seed, int32_t mode, double a, double b, double* __restrict__ r) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mode == 0) {
        r[i] = a;
        return;
    }
    double u1, u2;
    philox_uniforms(seed, i, 0, 0, &u1, &u2);
    const double z = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    r[i] = mode == 1 ? a * exp(b * z) : exp(a + b * z);
}

// easy axis uniform on the sphere by the reference's sampler: azimuth 2 pi u,
// elevation 2 asin sqrt(u') (src/simulation.cpp:296-312); moments by
// INIT_CONFIG_ROT (1: along z; 2: on the easy axes; 3: easy axes and
// moments along z)
__global__ void k_init_orient(int64_t n, uint64_t seed, int32_t rot, double two_pi, double* __restrict__ mm,
                              double* __restrict__ ea) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u1, u2;
    philox_uniforms(seed, i, 0, 0, &u1, &u2);
    const double az = two_pi * u2;
    const double el = 2.0 * asin(sqrt(u1));
    double e[3] = {sin(el) * cos(az), sin(el) * sin(az), cos(el)};
    if (rot == 3) e[0] = e[1] = 0.0, e[2] = 1.0;
    for (int k = 0; k < 3; ++k) {
        ea[k * n + i] = e[k];
        mm[k * n + i] = rot == 1 ? (k == 2 ? 1.0 : 0.0) : e[k];
    }
}

struct RsaGrid {
    int32_t nc;      // cells per dimension (1: a single cell holds everything)
    int32_t dump;    // key of particles outside the grid being built (-1: left out of the sort)
    int32_t periodic;
    double L, cell;
};

__device__ __forceinline__ int32_t rsa_coord(double x, const RsaGrid& g) {
    int32_t c = (int32_t)floor((x + 0.5 * g.L) / g.cell);
    return c < 0 ? 0 : (c >= g.nc ? g.nc - 1 : c);
}
__device__ __forceinline__ int32_t rsa_key(double x, double y, double z, const RsaGrid& g) {
    return (rsa_coord(z, g) * g.nc + rsa_coord(y, g)) * g.nc + rsa_coord(x, g);
}

// Whether position p of radius ri overlaps a particle listed in the grid
// (start/end/sorted) other than `self` and, with `below`, of an index below
// it: distance < r_i + r_j (by minimum image when periodic).
__device__ bool rsa_overlaps(const double p[3], double ri, const RsaGrid& g, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ end, const int32_t* __restrict__ sorted,
                             const double* __restrict__ xs, int64_t n, const double* __restrict__ r, int64_t self,
                             bool below) {
    const int32_t cx = rsa_coord(p[0], g), cy = rsa_coord(p[1], g), cz = rsa_coord(p[2], g);
    const int R = g.nc == 1 ? 0 : 1;
    for (int dz = -R; dz <= R; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                int32_t jx = cx + dx, jy = cy + dy, jz = cz + dz;
                if (g.periodic) {
                    jx = (jx + g.nc) % g.nc;
                    jy = (jy + g.nc) % g.nc;
                    jz = (jz + g.nc) % g.nc;
                } else if (jx < 0 || jy < 0 || jz < 0 || jx >= g.nc || jy >= g.nc || jz >= g.nc) {
                    continue;
                }
                const int32_t c = (jz * g.nc + jy) * g.nc + jx;
                for (int32_t k = start[c]; k < end[c]; ++k) {
                    const int32_t j = sorted[k];
                    if (j == self || (below && j > self)) continue;
                    double d2 = 0;
                    for (int d = 0; d < 3; ++d) {
                        double v = xs[d * n + j] - p[d];
                        if (g.periodic) v -= g.L * rint(v / g.L);
                        d2 += v * v;
                    }
                    const double lim = ri + r[j];
                    if (d2 < lim * lim) return true;
                }
            }
    return false;
}

// A proposal per unplaced particle: uniform in the box (periodic) or in
// [-(L/2 - r_i), L/2 - r_i)^3 (the reference's range); three Philox blocks
// per round, counter (particle, round, block).
__global__ void k_rsa_propose(int64_t n, uint64_t seed, uint64_t round, RsaGrid g, const double* __restrict__ r,
                              const int32_t* __restrict__ placed, double* __restrict__ cand) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || placed[i]) return;
    const double ext = g.periodic ? 0.5 * g.L : 0.5 * g.L - r[i];
    for (int d = 0; d < 3; ++d) {
        double u1, u2;
        philox_uniforms(seed, i, round, d, &u1, &u2);
        cand[d * n + i] = (2.0 * u2 - 1.0) * ext;  // u2 in [0, 1)
    }
}

// Proposals that overlap no placed particle enter the proposal grid.
__global__ void k_rsa_check(int64_t n, RsaGrid g, const double* __restrict__ r, const int32_t* __restrict__ placed,
                            const double* __restrict__ cand, const double* __restrict__ pos,
                            const int32_t* __restrict__ start, const int32_t* __restrict__ end,
                            const int32_t* __restrict__ sorted, int32_t* __restrict__ key_prop) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (placed[i]) {
        key_prop[i] = g.dump;
        return;
    }
    const double p[3] = {cand[i], cand[n + i], cand[2 * n + i]};
    const bool bad = rsa_overlaps(p, r[i], g, start, end, sorted, pos, n, r, -1, false);
    key_prop[i] = bad ? g.dump : rsa_key(p[0], p[1], p[2], g);
}

// A proposal in the grid is accepted unless it overlaps one of a lower index
// (also in the grid); accepted ones become placed.
__global__ void k_rsa_accept(int64_t n, RsaGrid g, const double* __restrict__ r, const int32_t* __restrict__ key_prop,
                             const double* __restrict__ cand, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ end, const int32_t* __restrict__ sorted,
                             double* __restrict__ pos, int32_t* __restrict__ placed, int32_t* __restrict__ key_placed,
                             int32_t* __restrict__ n_placed) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || key_prop[i] < 0) return;
    const double p[3] = {cand[i], cand[n + i], cand[2 * n + i]};
    if (rsa_overlaps(p, r[i], g, start, end, sorted, cand, n, r, i, true)) return;
    for (int d = 0; d < 3; ++d) pos[d * n + i] = p[d];
    placed[i] = 1;
    key_placed[i] = key_prop[i];
    atomicAdd(n_placed, 1);
}

// Device memory and stream of the initialisation, per device, kept until the
// process ends: a call may run while other host threads capture HIP graphs
// on their own streams, and freeing device memory (or the legacy default
// stream) would synchronise with those streams and break their captures. So
// the workspace only grows (a larger block is taken and the old one kept),
// and all work and copies go to a non-blocking stream of its own.
struct Workspace {
    std::mutex mu;
    hipStream_t stream = nullptr;
    char* base = nullptr;
    size_t cap = 0;
    std::vector<char*> old;  // outgrown blocks, never freed (see above)
};
Workspace g_ws[64];

// Carves aligned pieces out of the workspace block.
struct Carver {
    char* p;
    template <typename T>
    T* take(size_t count) {
        T* q = (T*)p;
        p += (std::max<size_t>(1, count) * sizeof(T) + 255) & ~(size_t)255;
        return q;
    }
};

#define DCHK(x)                                      \
    do {                                             \
        if ((x) != hipSuccess) return MDS_EDEVICE;   \
    } while (0)

}  // namespace

extern "C" {

mds_status mds_init_suspension_device(const mds_config* cfg, int device, int64_t n, uint64_t seed, int32_t periodic,
                                      double* l_box, double* r_mag, double* pos, double* pos_mm, double* pos_ea) {
    if (!cfg || n <= 0 || n >= ((int64_t)1 << 30) || !l_box || !pos) return MDS_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= 64) return MDS_EDEVICE;
    DCHK(hipSetDevice(device));
    Workspace& ws = g_ws[device];
    std::lock_guard<std::mutex> lk(ws.mu);
    if (!ws.stream) DCHK(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
    hipStream_t s = ws.stream;
    const int blocks = (int)((n + 255) / 256);
    // radii (mds_init_radii's modes)
    int32_t mode = 0;
    double a = cfg->r_mag_mean, b = 0;
    if (cfg->sigma_log > 0) {
        mode = 1;
        a = cfg->r_mag_median > 0 ? cfg->r_mag_median : cfg->r_mag_mean;
        b = cfg->sigma_log;
    } else if (cfg->size_dist == 1) {
        const double m = cfg->r_mag_mean, sg = cfg->sigma;
        if (!(m > 0) || !(sg > 0)) return MDS_ECONFIG;
        mode = 2;
        a = std::log(m * m / std::sqrt(m * m + sg * sg));  // src/simulation.cpp:180-181
        b = std::sqrt(std::log(1.0 + (sg / m) * (sg / m)));
    } else if (!(cfg->r_mag_mean > 0)) {
        return MDS_ECONFIG;
    }
    // the cell grid is fixed by the radii; the workspace is sized for the
    // largest grid that n particles can ask for
    const int64_t nc_cap = std::min<int64_t>(512, (int64_t)std::cbrt(8.0 * (double)n) + 1);
    const int64_t cells_cap = nc_cap * nc_cap * nc_cap;
    // 14 doubles and 4 int32 per particle, 4 int32 per cell, the scan tiles,
    // and 256 bytes of alignment for each of the 16 pieces
    const size_t need = (size_t)n * (14 * 8 + 4 * 4) + (size_t)cells_cap * 16 +
                        (size_t)(mds::scan_tiles(cells_cap) + 2) * 4 + 16 * 256;
    if (ws.cap < need) {
        char* p = nullptr;
        const size_t cap = need + need / 2;
        DCHK(hipMalloc((void**)&p, cap));
        if (ws.base) ws.old.push_back(ws.base);
        ws.base = p;
        ws.cap = cap;
    }
    Carver cv{ws.base};
    double* d_r = cv.take<double>(n);
    double* d_rh = cv.take<double>(n);
    double* d_mm = cv.take<double>(3 * n);
    double* d_ea = cv.take<double>(3 * n);
    double* d_pos = cv.take<double>(3 * n);
    double* d_cand = cv.take<double>(3 * n);
    int32_t* placed = cv.take<int32_t>(n);
    int32_t* key_prop = cv.take<int32_t>(n);
    int32_t* key_placed = cv.take<int32_t>(n);
    int32_t* sorted = cv.take<int32_t>(n);
    int32_t* cnt = cv.take<int32_t>(cells_cap);
    int32_t* start = cv.take<int32_t>(cells_cap);
    int32_t* end = cv.take<int32_t>(cells_cap);
    int32_t* cur = cv.take<int32_t>(cells_cap);
    int32_t* tmp = cv.take<int32_t>(mds::scan_tiles(cells_cap) + 1);
    int32_t* counter = cv.take<int32_t>(1);
    if ((size_t)(cv.p - ws.base) > ws.cap) return MDS_EINVAL;  // the pieces must fit the block sized above

    hipLaunchKernelGGL(k_init_radii, dim3(blocks), dim3(256), 0, s, n, stream_seed(seed, 1), mode, a, b, d_r);
    DCHK(hipGetLastError());
    std::vector<double> hr((size_t)n), hh((size_t)n);
    DCHK(hipMemcpyAsync(hr.data(), d_r, n * sizeof(double), hipMemcpyDeviceToHost, s));
    DCHK(hipStreamSynchronize(s));
    if (r_mag) std::copy(hr.begin(), hr.end(), r_mag);
    // hydrodynamic radii and box (InitParticleProps, src/simulation.cpp:56)
    mds_status st = mds_particle_props(cfg, n, hr.data(), hh.data(), nullptr, nullptr, nullptr, nullptr);
    if (st != MDS_OK) return st;
    if (!(*l_box > 0)) *l_box = mds_box_length(cfg, n, hh.data());
    const double L = *l_box;
    double rmax = 0;
    for (double v : hh) {
        if (!(v > 0) || !(v < L / 2)) return MDS_EINVAL;
        rmax = std::max(rmax, v);
    }
    // orientations
    hipLaunchKernelGGL(k_init_orient, dim3(blocks), dim3(256), 0, s, n, stream_seed(seed, 3), cfg->init_config_rot,
                       2.0 * cfg->my_pi, d_mm, d_ea);
    DCHK(hipGetLastError());
    if (pos_mm) DCHK(hipMemcpyAsync(pos_mm, d_mm, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    if (pos_ea) DCHK(hipMemcpyAsync(pos_ea, d_ea, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    DCHK(hipStreamSynchronize(s));
    if (cfg->init_config_trans == 2) return mds_init_positions(cfg, n, hh.data(), L, seed, periodic, pos);
    // positions: rounds of parallel random sequential addition
    RsaGrid g{};
    int64_t nc = (int64_t)std::floor(L / (2 * rmax));
    nc = std::max<int64_t>(1, std::min<int64_t>(nc, nc_cap));
    if (periodic && nc < 3) nc = 1;
    g.nc = (int32_t)nc;
    g.dump = -1;
    g.periodic = periodic ? 1 : 0;
    g.L = L;
    g.cell = L / (double)nc;
    const int32_t ncells = (int32_t)(nc * nc * nc);
    DCHK(hipMemcpyAsync(d_rh, hh.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    DCHK(mds::launch_fill_i32(s, placed, n, 0));
    DCHK(mds::launch_fill_i32(s, key_placed, n, g.dump));
    DCHK(mds::launch_fill_i32(s, counter, 1, 0));
    DCHK(mds::launch_cell_sort(s, n, key_placed, ncells, cnt, start, end, cur, sorted, tmp));
    const uint64_t ps = stream_seed(seed, 2);
    const int64_t max_rounds = 20000;
    int32_t done = 0;
    for (int64_t round = 0; done < n; ++round) {
        if (round >= max_rounds) return MDS_ECAPACITY;  // volume fraction too high for random addition
        hipLaunchKernelGGL(k_rsa_propose, dim3(blocks), dim3(256), 0, s, n, ps, (uint64_t)round, g, d_rh, placed,
                           d_cand);
        hipLaunchKernelGGL(k_rsa_check, dim3(blocks), dim3(256), 0, s, n, g, d_rh, placed, d_cand, d_pos, start, end,
                           sorted, key_prop);
        DCHK(hipGetLastError());
        DCHK(mds::launch_cell_sort(s, n, key_prop, ncells, cnt, start, end, cur, sorted, tmp));
        hipLaunchKernelGGL(k_rsa_accept, dim3(blocks), dim3(256), 0, s, n, g, d_rh, key_prop, d_cand, start, end,
                           sorted, d_pos, placed, key_placed, counter);
        DCHK(hipGetLastError());
        DCHK(mds::launch_cell_sort(s, n, key_placed, ncells, cnt, start, end, cur, sorted, tmp));
        DCHK(hipMemcpyAsync(&done, counter, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        DCHK(hipStreamSynchronize(s));
    }
    DCHK(hipMemcpyAsync(pos, d_pos, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    DCHK(hipStreamSynchronize(s));
    return MDS_OK;
}

}  // extern "C"
