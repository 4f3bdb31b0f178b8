// kernels.hip - gfx950 kernels of the Brownian/LLG step.
//
// Layout: every per-particle vector quantity is SoA (x[0..n), y[0..n),
// z[0..n)), the column-major layout of the reference's Eigen N x 3 arrays,
// so each component read or write by a wavefront is one coalesced 512 B line.
// Kernels are written for 64-lane wavefronts.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "erfcx_coef.h"
#include "kernels.h"
#include "philox.h"

namespace mds {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ double min_image(double v, double L) {
    // ApplyBoundaryCondVec (src/interaction.cpp:145-155): one shift each way
    if (v > L / 2) v -= L;
    if (v < -L / 2) v += L;
    return v;
}

// The same shift for |v| < L as v - L rint(v / L) (v - L and v + L are the
// same roundings as above; only |v| = L/2 exactly may differ, where the
// reference's choice is arbitrary too): three instructions, no selects.
__device__ __forceinline__ double min_image_r(double v, double L, double inv_L) {
    return fma(-rint(v * inv_L), L, v);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
// Wavefront index within the block, as a value the compiler knows to be
// uniform over the wavefront (so what it indexes stays in scalar registers).
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); }

// Ballot of a condition, taken from the compare's lane mask (HIP's __ballot
// takes an int and makes the condition a 0/1 value and back first).
__device__ __forceinline__ uint64_t ballot(bool c) { return __builtin_amdgcn_ballot_w64(c); }

__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ double readlane_d(double v, int l) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int cell_coord(double x, double L, double cell, int ncell) {
    int c = (int)floor((x + L / 2) / cell);
    return c < 0 ? 0 : (c >= ncell ? ncell - 1 : c);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// ------------------------------------------------------------------ fills
__global__ void k_fill_i32(int32_t* p, int64_t count, int32_t v, const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t count, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_scale(double* p, int64_t count, double f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] *= f;
}

int grid_for(int64_t n, int block, int cap = 65535 * 4) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

// ------------------------------------------------------------------ step setup
// pos0 <- pos (step-start positions: EvalSinCos and the pair scan read these),
// pos_cur <- pos (positions the overlap correction moves), cell keys in the
// replica's main and fine grids (global cell ids).
__device__ __forceinline__ int32_t main_key(const Rep& R, double x, double y, double z) {
    const double L = R.L;
    const int nc = R.ncell;
    return R.cell_off + (cell_coord(z, L, R.cell, nc) * nc + cell_coord(y, L, R.cell, nc)) * nc +
           cell_coord(x, L, R.cell, nc);
}

__global__ void k_prepare(int64_t n, const double* __restrict__ pos, double* __restrict__ pos0,
                          double* __restrict__ pos_cur, const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                          int32_t* __restrict__ key_main, int32_t* __restrict__ key_fine,
                          const double* __restrict__ pos_build, const int32_t* __restrict__ list_state,
                          unsigned long long* __restrict__ list_maxd2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool live = i < n;
    const int32_t r = live ? rep_of[i] : 0;
    double d2 = 0.0;
    if (live) {
        const Rep& R = reps[r];
        const double x = pos[i], y = pos[n + i], z = pos[2 * n + i];
        pos0[i] = x;
        pos0[n + i] = y;
        pos0[2 * n + i] = z;
        pos_cur[i] = x;
        pos_cur[n + i] = y;
        pos_cur[2 * n + i] = z;
        const double L = R.L;
        const int fn = R.fncell;
        if (!pos_build) {
            key_main[i] = main_key(R, x, y, z);
        } else if (list_state[r] != LIST_INVALID) {
            // displacement from the build position (minimum image): a
            // particle that crossed a face since the build is still near it
            const double iL = 1.0 / L;
            const double dx = min_image_r(x - pos_build[i], L, iL);
            const double dy = min_image_r(y - pos_build[n + i], L, iL);
            const double dz = min_image_r(z - pos_build[2 * n + i], L, iL);
            d2 = dx * dx + dy * dy + dz * dz;
        }
        key_fine[i] = R.fcell_off + (cell_coord(z, L, R.fcell, fn) * fn + cell_coord(y, L, R.fcell, fn)) * fn +
                      cell_coord(x, L, R.fcell, fn);
    }
    if (!pos_build) return;  // uniform
    // the replica's largest displacement (non-negative doubles order as
    // their bit patterns): one atomic per wavefront when its particles share
    // a replica, else one per particle
    const int32_t r0 = __builtin_amdgcn_readfirstlane(r);
    if (ballot(live && r != r0) == 0) {
        const double m = wave_max(d2);
        if (lane_id() == 0 && m > 0) atomicMax(&list_maxd2[r0], (unsigned long long)__double_as_longlong(m));
    } else if (live && d2 > 0) {
        atomicMax(&list_maxd2[r], (unsigned long long)__double_as_longlong(d2));
    }
}

// Pair list mode of each replica for this step.
// One block. sort_skip: 1 when every replica keeps its list, so that the
// main grid's keys, cell sort, order, work items and lists would come out as
// the last step left them (the kernels that make them return at once).
__global__ __launch_bounds__(256) void k_list_decide(int32_t nrep, const Rep* __restrict__ reps,
                                                     int32_t* __restrict__ list_state,
                                                     unsigned long long* __restrict__ list_maxd2,
                                                     int32_t* __restrict__ list_mode,
                                                     unsigned long long* __restrict__ list_builds,
                                                     int32_t* __restrict__ sort_skip) {
    int rebuild = 0;
    for (int32_t r = threadIdx.x; r < nrep; r += blockDim.x) {
        const int32_t st = list_state[r];
        const double d2 = __longlong_as_double((long long)list_maxd2[r]);
        int32_t m;
        if (st == LIST_INVALID || !(d2 <= reps[r].skin_half2)) m = LIST_FRESH;
        else m = st == LIST_STALE ? LIST_RESTORE : LIST_KEEP;
        list_mode[r] = m;
        if (m != LIST_KEEP) {
            list_builds[r] += 1;
            rebuild = 1;
        }
        list_state[r] = LIST_VALID;
        list_maxd2[r] = 0ull;
    }
    rebuild = __syncthreads_or(rebuild);
    if (threadIdx.x == 0) *sort_skip = rebuild ? 0 : 1;
}

// Build positions of the replicas built afresh; main-grid keys of every
// particle from its build position (so a kept replica's cells, order and
// work items come out as they were built).
__global__ void k_list_keys(int64_t n, const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                            const int32_t* __restrict__ list_mode, const double* __restrict__ pos0,
                            double* __restrict__ pos_build, int32_t* __restrict__ key_main,
                            const int32_t* __restrict__ skip) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || (skip && *skip)) return;
    const int32_t r = rep_of[i];
    double x, y, z;
    if (list_mode[r] == LIST_FRESH) {
        x = pos0[i];
        y = pos0[n + i];
        z = pos0[2 * n + i];
        pos_build[i] = x;
        pos_build[n + i] = y;
        pos_build[2 * n + i] = z;
    } else {
        x = pos_build[i];
        y = pos_build[n + i];
        z = pos_build[2 * n + i];
    }
    key_main[i] = main_key(reps[r], x, y, z);
}

// Counting sort by key; a negative key leaves the item out of the sort.
__global__ void k_count(int64_t n, const int32_t* __restrict__ key, int32_t* __restrict__ count,
                        const int32_t* __restrict__ skip) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (skip && *skip) return;
    if (i < n && key[i] >= 0) atomicAdd(&count[key[i]], 1);
}

// Exclusive scan of `count` in three passes over tiles of SCAN_TILE entries:
// tile sums, a scan of the tile sums (one block), and the scan within each
// tile plus its offset. Writes start, end (= start + count, if not null) and
// a cursor copy of start (if not null); total (if not null) gets the sum.
constexpr int SCAN_TILE = 2048;  // 256 threads x 8 entries

__device__ __forceinline__ int32_t block_excl_scan256(int32_t v, int32_t* sh, int32_t& block_total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int32_t u = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    block_total = sh[255];
    const int32_t r = sh[t] - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_scan_sums(const int32_t* __restrict__ count, int32_t m,
                                                   int32_t* __restrict__ tile_sum, const int32_t* __restrict__ skip) {
    __shared__ int32_t sh[256];
    if (skip && *skip) return;  // uniform over the block
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * 8;
    int32_t v = 0;
    for (int k = 0; k < 8; ++k)
        if (base + k < m) v += count[base + k];
    int32_t tot;
    block_excl_scan256(v, sh, tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t* __restrict__ tile_sum, int32_t nt,
                                                   int32_t* __restrict__ total, const int32_t* __restrict__ skip) {
    __shared__ int32_t part[1024];
    if (skip && *skip) return;  // uniform over the block
    const int t = threadIdx.x;
    const int64_t per = (nt + 1023) / 1024;
    const int64_t b = t * per, e = (b + per < nt) ? b + per : nt;
    int32_t s = 0;
    for (int64_t i = b; i < e; ++i) s += tile_sum[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - s;
    for (int64_t i = b; i < e; ++i) {
        const int32_t c = tile_sum[i];
        tile_sum[i] = run;
        run += c;
    }
    if (total && t == 1023) *total = part[1023];
}

__global__ __launch_bounds__(256) void k_scan_apply(const int32_t* __restrict__ count, int32_t m,
                                                    const int32_t* __restrict__ tile_off, int32_t* __restrict__ start,
                                                    int32_t* __restrict__ end, int32_t* __restrict__ cursor,
                                                    const int32_t* __restrict__ skip) {
    __shared__ int32_t sh[256];
    if (skip && *skip) return;  // uniform over the block
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * 8;
    int32_t c[8];
    int32_t v = 0;
    for (int k = 0; k < 8; ++k) {
        c[k] = base + k < m ? count[base + k] : 0;
        v += c[k];
    }
    int32_t tot;
    int32_t run = tile_off[blockIdx.x] + block_excl_scan256(v, sh, tot);
    for (int k = 0; k < 8; ++k) {
        const int64_t i = base + k;
        if (i < m) {
            start[i] = run;
            if (end) end[i] = run + c[k];
            if (cursor) cursor[i] = run;
        }
        run += c[k];
    }
}

__global__ void k_place(int64_t n, const int32_t* __restrict__ key, int32_t* __restrict__ cursor,
                        int32_t* __restrict__ sorted_idx, const int32_t* __restrict__ skip) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (skip && *skip) return;
    if (i < n && key[i] >= 0) sorted_idx[atomicAdd(&cursor[key[i]], 1)] = (int32_t)i;
}

// Cell-sorted records. The fourth component of the FP64 record is a tag,
// not a number: the particle's original index in the low 32 bits of its bit
// pattern (the overlap correction later sets bit 32 of the moved ones), so
// the pair kernels need no further gathers for either.
__global__ void k_records(int64_t n, const int32_t* __restrict__ sorted_idx, const double* __restrict__ pos0,
                          double4* __restrict__ rec, float4* __restrict__ rec32, int32_t* __restrict__ inv_perm,
                          const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                          const double* __restrict__ pos_build, const int32_t* __restrict__ list_mode,
                          float4* __restrict__ rec_build32, int32_t* __restrict__ lcount) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t i = sorted_idx[k];
    double x = pos0[i], y = pos0[n + i], z = pos0[2 * n + i];
    if (pos_build) {
        const int32_t r = rep_of[i];
        const double L = reps[r].L, iL = 1.0 / L;
        const double bx = pos_build[i], by = pos_build[n + i], bz = pos_build[2 * n + i];
        // the image of the step-start position nearest the build position
        // (no shift, and the same bits, unless the particle crossed a face)
        x = fma(rint((bx - x) * iL), L, x);
        y = fma(rint((by - y) * iL), L, y);
        z = fma(rint((bz - z) * iL), L, z);
        rec_build32[k] = make_float4((float)bx, (float)by, (float)bz, 0.0f);
        if (list_mode[r] != LIST_KEEP)
            for (int w = 0; w < PAIR_SLICES; ++w) lcount[PAIR_SLICES * k + w] = 0;
    }
    rec[k] = make_double4(x, y, z, __longlong_as_double((long long)(uint32_t)i));
    rec32[k] = make_float4((float)x, (float)y, (float)z, 0.0f);
    inv_perm[i] = (int32_t)k;
}

// Bit 32 of the tag of every particle moved by the overlap correction.
__global__ void k_mark_moved(const int32_t* __restrict__ n_moved, const int32_t* __restrict__ moved_list,
                             const int32_t* __restrict__ inv_perm, double4* __restrict__ rec) {
    const int32_t m = *n_moved;
    for (int32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        const int32_t k = inv_perm[moved_list[t]];
        const long long b = __double_as_longlong(rec[k].w) | (1ll << 32);
        rec[k].w = __longlong_as_double(b);
    }
}

// Work items of the pair kernel: one per 64 particles of a cell.
__global__ __launch_bounds__(1024) void k_work(int32_t ncells, const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ end, int32_t* __restrict__ work,
                                               int32_t* __restrict__ nwork, int32_t cap, int32_t* __restrict__ err,
                                               const int32_t* __restrict__ skip) {
    __shared__ int32_t part[1024];
    if (skip && *skip) return;  // uniform over the block
    const int t = threadIdx.x;
    const int64_t per = (ncells + 1023) / 1024;
    const int64_t b = t * per, e = (b + per < ncells) ? b + per : ncells;
    int32_t s = 0;
    for (int64_t c = b; c < e; ++c) s += (end[c] - start[c] + WAVE - 1) / WAVE;
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - s;
    for (int64_t c = b; c < e; ++c) {
        const int32_t nch = (end[c] - start[c] + WAVE - 1) / WAVE;
        if (nch > 256) {
            atomicOr(err, 1);
            continue;
        }
        for (int32_t k = 0; k < nch; ++k)
            if (run + k < cap) work[run + k] = ((int32_t)c << 8) | k;
        run += nch;
    }
    if (t == 1023) {
        *nwork = part[1023] < cap ? part[1023] : cap;
        if (part[1023] > cap) atomicOr(err, 2);
    }
}

// ------------------------------------------------------------------ contacts
// Candidate pairs (i < j) whose step-start distance is below the contact
// distance r_i + r_j + 1e-10 plus a search margin, listed in the reference's
// lexicographic (i, j) order.
// One search pass, half a wavefront per particle: lane t < 27 scans
// neighbour cell t (dz, dy, dx in increasing order), reading the cell's
// positions and radii from the fine-cell ordered records (contiguous per
// cell). Particles are taken in fine-cell order, so neighbouring half-waves
// share their cells (the lists stay indexed by the original index). The
// lanes' hits go to LDS in cell-scan order (the first MAXC kept, as a serial
// scan would keep them), are ranked by j and appended as the particle's run
// to cand_tmp, at a base taken with one atomic (cand_base[n] counts); the
// count goes to cand_count. After the scan of the counts, k_cand_compact
// moves each run to its particle's offset, so the list does not depend on
// the order in which the runs were appended.
constexpr int CONTACT_MAXC = 32;
__global__ void k_frec(int64_t n, const int32_t* __restrict__ fsorted, const double* __restrict__ pos0,
                       const double* __restrict__ r_mag, double4* __restrict__ frec, int32_t* __restrict__ cand_base) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s == 0) cand_base[n] = 0;  // the append counter of k_contacts
    if (s >= n) return;
    const int64_t i = fsorted[s];
    if ((uint64_t)i >= (uint64_t)n) return;  // every slot holds a particle; guard the reads anyway
    frec[s] = make_double4(pos0[i], pos0[n + i], pos0[2 * n + i], r_mag[i]);
}
__global__ __launch_bounds__(256) void k_contacts(int64_t n, const Rep* __restrict__ reps,
                                                  const int32_t* __restrict__ rep_of, const double* __restrict__ pos0,
                                                  const double* __restrict__ r_mag, const int32_t* __restrict__ fstart,
                                                  const int32_t* __restrict__ fend, const int32_t* __restrict__ fsorted,
                                                  const double4* __restrict__ frec,
                                                  int32_t* __restrict__ cand_count, int32_t* __restrict__ cand_base,
                                                  int32_t* __restrict__ cand_tmp, int32_t cap,
                                                  int32_t* __restrict__ err) {
    __shared__ int32_t lst[8][CONTACT_MAXC];
    const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
    const int64_t k = blockIdx.x * 8LL + g;
    const bool on = k < n;  // uniform over the half-wave
    int64_t i = 0;
    int32_t c = 0;
    double xi = 0, yi = 0, zi = 0, ri = 0, L = 1, extra = 0;
    bool act = false;
    if (on) {
        i = fsorted[k];
        const Rep& R = reps[rep_of[i]];
        L = R.L;
        extra = R.margin;
        const double fcell = R.fcell;
        const int32_t fncell = R.fncell, foff = R.fcell_off;
        xi = pos0[i];
        yi = pos0[n + i];
        zi = pos0[2 * n + i];
        ri = r_mag[i];
        const int cx = cell_coord(xi, L, fcell, fncell), cy = cell_coord(yi, L, fcell, fncell),
                  cz = cell_coord(zi, L, fcell, fncell);
        const int reach = fncell >= 3 ? 1 : 0;
        act = reach ? t < 27 : t == 0;
        const int dz = reach ? t / 9 - 1 : 0, dy = reach ? (t / 3) % 3 - 1 : 0, dx = reach ? t % 3 - 1 : 0;
        c = foff + (((cz + dz + fncell) % fncell) * fncell + (cy + dy + fncell) % fncell) * fncell +
            (cx + dx + fncell) % fncell;
    }
    int32_t s0 = 0, s1 = 0;
    if (act) {
        s0 = fstart[c];
        s1 = fend[c];
    }
    auto hit = [&](int32_t s) {
        const double4 q = frec[s];
        const double vx = min_image(q.x - xi, L), vy = min_image(q.y - yi, L), vz = min_image(q.z - zi, L);
        const double lim = ri + q.w + 1e-10 + extra;
        return vx * vx + vy * vy + vz * vz < lim * lim;
    };
    int cl = 0;
    for (int32_t s = s0; s < s1; ++s) {
        const int32_t j = fsorted[s];
        if (j > i && hit(s)) ++cl;
    }
    int tot = cl;
    for (int o = 16; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 32);
    // exclusive prefix of the lanes' counts, in cell-scan order
    int pre = cl;
    for (int o = 1; o < 32; o <<= 1) {
        const int v = __shfl_up(pre, o, 32);
        if (t >= o) pre += v;
    }
    pre -= cl;
    if (on && t == 0 && tot > CONTACT_MAXC) atomicOr(err, 4);
    if (tot > 0)
        for (int32_t s = s0, m = pre; s < s1 && m < CONTACT_MAXC; ++s) {
            const int32_t j = fsorted[s];
            if (j > i && hit(s)) lst[g][m++] = j;
        }
    wave_sync();
    const int cn = tot < CONTACT_MAXC ? tot : CONTACT_MAXC;
    int32_t base = 0;
    if (on && t == 0) {
        if (cn > 0) base = atomicAdd(&cand_base[n], cn);
        cand_count[i] = tot;
        cand_base[i] = base;
    }
    base = __shfl(base, 0, 32);
    if (on && t < cn) {
        const int32_t e = lst[g][t];
        int rank = 0;
        for (int u = 0; u < cn; ++u) rank += lst[g][u] < e;
        if (base + rank < cap)
            cand_tmp[base + rank] = e;
        else
            atomicOr(err, 8);
    }
}

// Each particle's run of candidates from the append area to its offset (the
// exclusive scan of the counts): partner j, and owner i in the second half.
__global__ void k_cand_compact(int64_t n, const int32_t* __restrict__ cand_count, const int32_t* __restrict__ cand_off,
                               const int32_t* __restrict__ cand_base, const int32_t* __restrict__ cand_tmp,
                               int32_t* __restrict__ cand_i, int32_t* __restrict__ cand_j, int32_t cap,
                               int32_t* __restrict__ err) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t cn = min(cand_count[i], CONTACT_MAXC);
    if (cn <= 0) return;
    const int32_t b = cand_base[i], o = cand_off[i];
    for (int32_t m = 0; m < cn; ++m) {
        if (b + m >= cap) break;  // not written (flagged by k_contacts)
        if (o + m >= cap) {
            atomicOr(err, 8);
            break;
        }
        cand_i[o + m] = (int32_t)i;
        cand_j[o + m] = cand_tmp[b + m];
    }
}

// Sequential overlap correction of CompInteractVecs (src/interaction.cpp:120-129)
// over the candidate pairs in lexicographic order: a pair closer than contact
// moves both particles half the overlap along the pre-move axis, and later
// pairs see the moved positions. Every move is recorded as a version
// (pair key, new position) of that particle, so the pair kernels can give each
// pair the positions the reference's scan would have used. The scan is
// sequential within a replica; one thread per replica.
__global__ void k_resolve(int32_t nrep, int64_t n, const Rep* __restrict__ reps, const double* __restrict__ r_mag,
                          const int32_t* __restrict__ cand_off, const int32_t* __restrict__ cand_i,
                          const int32_t* __restrict__ cand_j, double* __restrict__ cur, const double* __restrict__ pos0,
                          int32_t* __restrict__ ev_head, int32_t* __restrict__ ev_tail, int32_t* __restrict__ ev_next,
                          uint64_t* __restrict__ ev_key, double* __restrict__ ev_pos, int32_t ev_cap,
                          int32_t* __restrict__ n_events, double* __restrict__ max_disp, double disp_limit,
                          int32_t* __restrict__ moved_list, int32_t* __restrict__ n_moved, int32_t* __restrict__ err) {
    const int32_t rid = blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= nrep) return;
    const Rep& R = reps[rid];
    // candidates beyond the list capacity were never written (flagged by k_contacts)
    const int32_t cap = ev_cap / 2;
    const int32_t kb = min(cand_off[R.off], cap), ke = min(cand_off[R.off + R.n], cap);
    const double L = R.L, rc = R.rc;
    double md = 0;
    for (int32_t k = kb; k < ke; ++k) {
        const int32_t i = cand_i[k], j = cand_j[k];
        double v[3];
        for (int d = 0; d < 3; ++d) v[d] = min_image(cur[d * n + j] - cur[d * n + i], L);
        const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(r < rc)) continue;
        const double dCont = r_mag[i] + r_mag[j] + 1e-10;
        if (!(r < dCont)) continue;
        const double dd = dCont - r;
        const uint64_t key = (uint64_t)i * (uint64_t)n + (uint64_t)j;
        const int32_t pp[2] = {i, j};
        for (int d = 0; d < 3; ++d) {
            const double u = v[d] / r;
            cur[d * n + i] -= 0.5 * dd * u;
            cur[d * n + j] += 0.5 * dd * u;
        }
        for (int s = 0; s < 2; ++s) {
            const int32_t p = pp[s];
            const int32_t e = atomicAdd(n_events, 1);
            if (e >= ev_cap) {
                atomicOr(err, 16);
                break;
            }
            ev_key[e] = key;
            double dsq = 0;
            for (int d = 0; d < 3; ++d) {
                ev_pos[3 * (int64_t)e + d] = cur[d * n + p];
                const double dl = cur[d * n + p] - pos0[d * n + p];
                dsq += dl * dl;
            }
            md = fmax(md, sqrt(dsq));
            ev_next[e] = -1;
            if (ev_tail[p] < 0) {
                ev_head[p] = e;
                moved_list[atomicAdd(n_moved, 1)] = p;
            } else
                ev_next[ev_tail[p]] = e;
            ev_tail[p] = e;
        }
    }
    // non-negative doubles order as their bit patterns
    atomicMax((unsigned long long*)max_disp, (unsigned long long)__double_as_longlong(md));
    if (md > disp_limit) atomicOr(err, 32);
}

// ------------------------------------------------------------------ moments
template <int BS>
__device__ __forceinline__ void block_sum3(double a, double b, double c, double* out3) {
    __shared__ double sh[3][BS / WAVE];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x / WAVE;
    if (lane_id() == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
        sh[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < BS / WAVE; ++k) {
            s0 += sh[0][k];
            s1 += sh[1][k];
            s2 += sh[2][k];
        }
        out3[0] = s0;
        out3[1] = s1;
        out3[2] = s2;
    }
}

// magMomVec = posMm .* magMom (src/simulation.cpp:638), in original and
// cell-sorted order.
__global__ __launch_bounds__(256) void k_moments(int64_t n, const double* __restrict__ mm,
                                                 const double* __restrict__ mag_mom, const int32_t* __restrict__ inv_perm,
                                                 const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                                                 double* __restrict__ mom, double4* __restrict__ rec_mom,
                                                 float4* __restrict__ rec_mom32) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = mag_mom[i];
    const double mx = mm[i] * a, my = mm[n + i] * a, mz = mm[2 * n + i] * a;
    if (mom) {
        mom[i] = mx;
        mom[n + i] = my;
        mom[2 * n + i] = mz;
    }
    if (!inv_perm) return;
    const int32_t k = inv_perm[i];
    rec_mom[k] = make_double4(mx, my, mz, 0.0);
    // FP32 copy in units of the replica's largest moment (mixed pair path)
    const double s = reps[rep_of[i]].inv_mref;
    rec_mom32[k] = make_float4((float)(mx * s), (float)(my * s), (float)(mz * s), 0.0f);
}

// Per-replica sums in a fixed order: chunks of RED_CHUNK particles (never
// straddling a replica) are summed by one block each, then each replica adds
// its chunks' sums in order. Columns col0 .. col0 + ncol - 1 of an SoA
// N x 3 array.
constexpr int RED_CHUNK = RED_CHUNK_PARTICLES;

__global__ __launch_bounds__(256) void k_red_part(const int2* __restrict__ chunks, const Rep* __restrict__ reps,
                                                  int64_t n, const double* __restrict__ v, int32_t col0, int32_t ncol,
                                                  double* __restrict__ part) {
    const int2 ch = chunks[blockIdx.x];
    const Rep& R = reps[ch.x];
    const int64_t b = ch.y, e = min((int64_t)ch.y + RED_CHUNK, R.off + R.n);
    double a[3] = {0, 0, 0};
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x)
        for (int c = 0; c < ncol; ++c) a[c] += v[(int64_t)(col0 + c) * n + i];
    __shared__ double r3[3];
    block_sum3<256>(a[0], a[1], a[2], r3);
    __syncthreads();
    if (threadIdx.x < 3) part[3 * (int64_t)blockIdx.x + threadIdx.x] = r3[threadIdx.x];
}

// magMomSum (src/simulation.cpp:639) per replica
__global__ void k_rep_sum3(const Rep* __restrict__ reps, int32_t nrep, const double* __restrict__ part,
                           double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const Rep& R = reps[r];
    double a = 0, b = 0, c = 0;
    for (int k = R.red_off; k < R.red_off + R.nred; ++k) {
        a += part[3 * (int64_t)k];
        b += part[3 * (int64_t)k + 1];
        c += part[3 * (int64_t)k + 2];
    }
    out[3 * r] = a;
    out[3 * r + 1] = b;
    out[3 * r + 2] = c;
}

// <m_z> per replica (src/simulation.cpp:515); with a series buffer, also
// stored at the running step index.
__global__ void k_rep_mean_z(const Rep* __restrict__ reps, int32_t nrep, const double* __restrict__ part,
                             double* __restrict__ out, const int32_t* __restrict__ counter,
                             double* __restrict__ series, int32_t series_cap) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const Rep& R = reps[r];
    double a = 0;
    for (int k = R.red_off; k < R.red_off + R.nred; ++k) a += part[3 * (int64_t)k];
    const double m = a / (double)R.n;
    out[r] = m;
    if (series) {
        const int32_t k = *counter;
        if (k < series_cap) series[(int64_t)k * nrep + r] = m;
    }
}

// ------------------------------------------------------------------ allocation-free <m_z>
// Scratch of the allocation-free <m_z> entry point: one partial sum per chunk
// of RED_CHUNK particles (512 KiB). Every device holds its own copy for the
// life of the process; host code reaches it only through the address that
// hipGetSymbolAddress returns with that device current.
constexpr int64_t NOALLOC_MAX_CHUNKS = NOALLOC_MAX_PARTICLES / RED_CHUNK_PARTICLES;
__device__ __align__(256) double scratch[NOALLOC_MAX_CHUNKS];

// The z column of chunk blockIdx.x, summed as k_red_part sums it.
__global__ __launch_bounds__(256) void k_red_part_z(int64_t n, const double* __restrict__ mm,
                                                    double* __restrict__ part) {
    const int64_t b = (int64_t)blockIdx.x * RED_CHUNK, e = min(b + RED_CHUNK, n);
    double a = 0;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) a += mm[2 * n + i];
    __shared__ double r3[3];
    block_sum3<256>(a, 0.0, 0.0, r3);
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = r3[0];
}

// The chunks' sums in order, over n (as k_rep_mean_z).
__global__ void k_mean_z_final(int64_t nch, int64_t n, const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0;
    for (int64_t k = 0; k < nch; ++k) a += part[k];
    out[0] = a / (double)n;
}

// ------------------------------------------------------------------ real space
// Real-space Ewald pair terms (CompDipInteractEwaldR, src/interaction.cpp:383-422;
// field only: CompDipFieldEwaldR, 555-582), pressure (EvalPressTensLR,
// src/postprocessing.cpp:195-200) and, with the coating potential, the
// short-range forces (src/interaction.cpp:766-832).
//
// One block of PAIR_SLICES wavefronts takes up to 64 particles of one cell
// (the i particles, one per lane; the FP64 kernels share their positions and
// moments in LDS, the mixed kernel reads them through L1) and each wavefront one interleaved slice of the half stencil: the cell
// itself (pairs j > i in cell-sorted order) and the neighbour offsets that
// are positive in (z, y, x) order, so every pair within r_c is met exactly
// once (in a box that is one cell, the wavefronts share that cell's windows,
// taken as a ring: SliceWindows). Each candidate j is tested by all lanes (a
// cheap FP32 distance test);
// lanes that hit push (lane, j) into the wavefront's LDS queue, and queued
// pairs are evaluated densely, one pair per lane, so the expensive erfc/exp
// path never runs with idle lanes. The queue is read with a stride of the
// number of batches it fills, which spreads the lanes that share a j (one
// candidate hits many lanes) over the batches and keeps LDS atomic conflicts
// low. Each evaluation updates both particles: i in the wavefront's
// accumulator, j in an accumulator over the current 64-particle block of the
// neighbour cell (the window), which goes to the cell-sorted fixed-point sums
// when the window is done; the wavefronts' i sums are added in a fixed order
// at the end. Every pair is evaluated in the orientation of its lower
// original index (the reference's i < j), so both particles see the
// reference's geometry.
// A wavefront evaluates its queue once it holds this many slots (whole
// quads), and at the end of each window; one candidate adds at most 64 slots
// more. Longer queues give more batches per flush (loads of one batch overlap
// the arithmetic of others) at the cost of LDS; the sizes are the fastest of
// 128-640 measured at C3 for each kernel, the mixed force pass's with the
// k-space chain beside it (a longer queue runs it alone as fast but leaves
// less LDS to the chain's blocks).
template <bool FORCE, bool MIXED>
constexpr int pair_qflush() {
    return MIXED ? (FORCE ? 320 : 384) : (FORCE ? 512 : 640);
}
template <bool MIXED, bool COATING>
struct IChunk {
    double4 ipos[WAVE];      // positions; .w is the record's tag
    typename std::conditional<MIXED && !COATING, float4, double4>::type imom[WAVE];
};
// A wavefront's queue and FP64 accumulators (LDS atomics: `ds_add_f64`; FP32
// accumulators with `ds_add_f32` ran the mixed kernel 4x slower).
template <bool FORCE, int QF>
struct IWave {
    static_assert(QF % WAVE == 0, "the flush reads whole 64-slot batches of the queue");
    static constexpr int NA = FORCE ? 6 : 3;    // (force,) field components
    static constexpr int QFLUSH = QF;
    static constexpr int QCAP = QF + WAVE;
    int32_t q[QCAP];
    double acc[NA][WAVE];    // this wavefront's sums of the i particles
    double jacc[NA][WAVE];   // sums of the current window's j particles
};
// The periodic image of the current window's cell: j + (x, y, z) is the image
// nearest the i cell, which for a pair within r_c < L/2 is the minimum image,
// so the separation needs one add per component. With one cell per box side
// (`wrap`) every image is in the one cell and the minimum image is taken per
// pair.
struct WinImage {
    double x, y, z;
    bool wrap;
};

// One cell per box side (every pair is within one cell): the cell's 64-
// particle chunks form a ring, and chunk k takes the windows at ring
// distance d = 0 .. D from itself, D = nch / 2 (for an even nch, distance
// nch / 2 only from the lower half), so every chunk pair is met once and the
// work items of the cell hold about the same number of windows (against
// nch - k from chunk k on); the wavefronts of the block take them in turn.
// Other grids: each wavefront takes every PAIR_SLICES-th stencil cell, with
// all of its windows. Windows of slice si for wavefront wv: count, and the
// start of window wi.
struct SliceWindows {
    int32_t first, step, count, base, nch, chunk;
    __device__ int32_t start(int32_t wi) const {
        if (nch == 0) return base + wi * WAVE;
        int32_t kk = chunk + wi;
        if (kk >= nch) kk -= nch;
        return base + kk * WAVE;
    }
};
__device__ __forceinline__ SliceWindows slice_windows(bool allpairs, int wv, int32_t cs, int32_t ce, int32_t cell_s,
                                                      int32_t chunk) {
    SliceWindows sw;
    if (allpairs) {
        const int32_t nch = (ce - cell_s + WAVE - 1) / WAVE;
        sw.first = wv;
        sw.step = PAIR_SLICES;
        sw.count = nch / 2 + 1 - ((nch % 2 == 0 && chunk >= nch / 2) ? 1 : 0);
        sw.base = cell_s;
        sw.nch = nch;
        sw.chunk = chunk;
    } else {
        sw.first = 0;
        sw.step = 1;
        sw.count = (ce - cs + WAVE - 1) / WAVE;
        sw.base = cs;
        sw.nch = 0;
        sw.chunk = 0;
    }
    return sw;
}

// Positions of lo and hi as the reference's sequential scan sees them when it
// reaches pair (lo, hi): the last overlap move of each with a smaller
// lexicographic key; and whether this pair itself was moved.
__device__ __forceinline__ void version_pair(const Geometry& g, int32_t lo, int32_t hi, double& xl, double& yl,
                                             double& zl, double& xh, double& yh, double& zh, int& event) {
    const uint64_t key = (uint64_t)lo * (uint64_t)g.n + (uint64_t)hi;
    int32_t e = g.ev_head[lo];
    while (e >= 0 && g.ev_key[e] < key) {
        xl = g.ev_pos[3 * (int64_t)e];
        yl = g.ev_pos[3 * (int64_t)e + 1];
        zl = g.ev_pos[3 * (int64_t)e + 2];
        e = g.ev_next[e];
    }
    event = (e >= 0 && g.ev_key[e] == key) ? 1 : 0;
    e = g.ev_head[hi];
    while (e >= 0 && g.ev_key[e] < key) {
        xh = g.ev_pos[3 * (int64_t)e];
        yh = g.ev_pos[3 * (int64_t)e + 1];
        zh = g.ev_pos[3 * (int64_t)e + 2];
        e = g.ev_next[e];
    }
}

// 1/d by the hardware reciprocal and two Newton steps (d > 0, normal)
__device__ __forceinline__ double rcp64(double d) {
    double y = __builtin_amdgcn_rcp(d);
    y = fma(y, fma(-d, y, 1.0), y);
    y = fma(y, fma(-d, y, 1.0), y);
    return y;
}

// 1/sqrt(q) by the hardware estimate and two Newton steps (q > 0, normal)
__device__ __forceinline__ double rsqrt64(double q) {
    double y = __builtin_amdgcn_rsq(q);
    const double h = 0.5 * q;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// exp(-y) for 0 <= y < 700: Cody-Waite reduction by ln 2 (ln 2 split into a
// 32-bit head and its remainder, so k * head is exact) to |f| <= ln2 / 2 and
// the degree-12 Taylor polynomial of e^f (truncation < 2e-16 relative).
__device__ __forceinline__ double exp_neg(double y) {
    constexpr double kLog2e = 1.4426950408889634;
    constexpr double kLn2Hi = 0x1.62e42ff000000p-1;
    constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;
    const double k = rint(y * kLog2e);
    double f = fma(k, kLn2Hi, -y);  // k ln2 - y
    f = fma(k, kLn2Lo, f);
    double p = 2.08767569878681e-09;
    p = fma(p, f, 2.505210838544172e-08);
    p = fma(p, f, 2.755731922398589e-07);
    p = fma(p, f, 2.7557319223985893e-06);
    p = fma(p, f, 2.48015873015873e-05);
    p = fma(p, f, 0.0001984126984126984);
    p = fma(p, f, 0.001388888888888889);
    p = fma(p, f, 0.008333333333333333);
    p = fma(p, f, 0.041666666666666664);
    p = fma(p, f, 0.16666666666666666);
    p = fma(p, f, 0.5);
    p = fma(p, f, 1.0);
    p = fma(p, f, 1.0);
    return ldexp(p, -(int)k);
}

__device__ __forceinline__ double erfcx64(double x) {
    const double t = fma(-2.0 * kErfcxK64, rcp64(x + kErfcxK64), 1.0);
    double p = kErfcx64[kErfcxDeg64];
#pragma unroll
    for (int k = kErfcxDeg64 - 1; k >= 0; --k) p = fma(p, t, kErfcx64[k]);
    return p;
}
__device__ __forceinline__ float erfcx32(float x) {
    // hardware reciprocal (1 ulp): far below the 8.7e-7 error of the fit
    const float t = 1.0f - 2.0f * kErfcxK32 * __builtin_amdgcn_rcpf(x + kErfcxK32);
    float p = kErfcx32[kErfcxDeg32];
#pragma unroll
    for (int k = kErfcxDeg32 - 1; k >= 0; --k) p = fmaf(p, t, kErfcx32[k]);
    return p;
}

// FP32 terms of one pair in units of ell (lengths) and mref (moments): field
// on i from j and (FORCE) the force on j, from the separation v = r u from i
// to j (|v| = rs, irs = 1 / rs). With the reference's B, C, D (a1, a2, a3
// here) and u = v / r: C (mj.u) r^2 u = C (mj.v) v, and the force's
// (C mi.mj - D r^2 (mi.u)(mj.u)) r u + C r ((mj.u) mi + (mi.u) mj)
// = (C mi.mj - D (mi.v)(mj.v)) v + C ((mj.v) mi + (mi.v) mj), so u is never
// formed.
// With G = e^-x^2 / r^2 and c = 2 alpha / sqrt(pi) (t = c G):
// a1 = G (erfcx(x) / r + c), a2 = 3 a1 / r^2 + 2 alpha^2 c G,
// a3 = 5 a2 / r^2 + 4 alpha^4 c G. e^-x^2 is 2^(r^2 (-alpha^2 log2 e)).
// `on` false gives exact zeros (a1 = a2 = a3 = 0) whatever r is.
template <bool FORCE>
__device__ __forceinline__ void pair_terms32(const PairParams& pp, float r2, float rs, float irs, float vx, float vy,
                                             float vz, float mix, float miy, float miz, float mjx, float mjy,
                                             float mjz, float B[3], float Bo[3], float F[3], bool on = true) {
    const float irs2 = irs * irs;
    const float x = pp.al_s * rs;
    const float c = 1.1283791670955126f * pp.al_s;
    const float G = __builtin_amdgcn_exp2f(r2 * (-1.4426950408889634f * pp.al2_s)) * irs2;
    float a1 = G * fmaf(erfcx32(x), irs, c);
    float a2 = fmaf(a1, 3.0f * irs2, (2.0f * pp.al2_s * c) * G);
    if (!on) {
        a1 = 0.0f;
        a2 = 0.0f;
    }
    const float dj = mjx * vx + mjy * vy + mjz * vz;
    const float di = mix * vx + miy * vy + miz * vz;
    const float g = a2 * dj, h = a2 * di;
    B[0] = g * vx - a1 * mjx;
    B[1] = g * vy - a1 * mjy;
    B[2] = g * vz - a1 * mjz;
    Bo[0] = h * vx - a1 * mix;
    Bo[1] = h * vy - a1 * miy;
    Bo[2] = h * vz - a1 * miz;
    if (FORCE) {
        const float a3 = on ? fmaf(a2, 5.0f * irs2, (4.0f * pp.al4_s * c) * G) : 0.0f;
        const float mimj = mix * mjx + miy * mjy + miz * mjz;
        const float c0 = a2 * mimj - a3 * di * dj;
        F[0] = c0 * vx + (g * mix + h * mjx);
        F[1] = c0 * vy + (g * miy + h * mjy);
        F[2] = c0 * vz + (g * miz + h * mjz);
    }
}

// Terms of one pair (CompDipInteractEwaldR, src/interaction.cpp:396-417):
// field on i from j, and (FORCE) the force on j (applied -F to i), from the
// separation r, its reciprocal, the unit vector u from i to j and both
// moments. FP64 path in SI units without the facEwald factor; FP32 path in
// units of ell and mref (see PairParams). erfc(alpha r) is
// exp(-(alpha r)^2) erfcx(alpha r), sharing the Gaussian of the t term.
template <bool FORCE, bool MIXED>
__device__ __forceinline__ void pair_terms(const PairParams& pp, double r, double ir, double ux, double uy, double uz,
                                           const double4& mi, const double4& mj, double B[3], double Bo[3],
                                           double F[3]) {
    if (!MIXED) {
        // B = a1, C = a2, D = a3 of the reference with the products regrouped
        // so that r^2 and the uniform constants are formed once
        const double ir2 = ir * ir;
        const double r2 = r * r;
        const double x = pp.alpha * r;
        const double gau = exp_neg(x * x);
        const double t = (2.0 * pp.alpha * pp.ir_pi) * (gau * ir2);
        const double a1 = gau * erfcx64(x) * (ir2 * ir) + t;
        const double a2 = 3.0 * a1 * ir2 + 2.0 * pp.alpha2 * t;
        const double dj = mj.x * ux + mj.y * uy + mj.z * uz;
        const double di = mi.x * ux + mi.y * uy + mi.z * uz;
        const double a2r2 = a2 * r2;
        const double g2 = a2r2 * dj;
        B[0] = g2 * ux - a1 * mj.x;
        B[1] = g2 * uy - a1 * mj.y;
        B[2] = g2 * uz - a1 * mj.z;
        const double h2 = a2r2 * di;
        Bo[0] = h2 * ux - a1 * mi.x;
        Bo[1] = h2 * uy - a1 * mi.y;
        Bo[2] = h2 * uz - a1 * mi.z;
        if (FORCE) {
            const double a3 = 5.0 * a2 * ir2 + 4.0 * pp.alpha4 * t;
            const double mimj = mi.x * mj.x + mi.y * mj.y + mi.z * mj.z;
            const double c0 = (a2 * mimj - (a3 * r2) * (di * dj)) * r;
            const double c1 = a2 * r;
            F[0] = c0 * ux + c1 * (dj * mi.x + di * mj.x);
            F[1] = c0 * uy + c1 * (dj * mi.y + di * mj.y);
            F[2] = c0 * uz + c1 * (dj * mi.z + di * mj.z);
        }
    } else {
        const double s = pp.inv_mref;
        const double rl = r * pp.inv_ell;
        float b[3], bo[3], f[3];
        pair_terms32<FORCE>(pp, (float)(rl * rl), (float)rl, (float)(ir * pp.ell), (float)(ux * rl), (float)(uy * rl),
                            (float)(uz * rl),
                            (float)(mi.x * s), (float)(mi.y * s), (float)(mi.z * s), (float)(mj.x * s),
                            (float)(mj.y * s), (float)(mj.z * s), b, bo, f);
        for (int d = 0; d < 3; ++d) {
            B[d] = b[d];
            Bo[d] = bo[d];
            if (FORCE) F[d] = f[d];
        }
    }
}

// Short-range forces (CompSterRep, CompElectrostatRep, CompForceVdW,
// src/interaction.cpp:766-832) of a pair at stored length r, as the
// magnitude along u(lo -> hi) of the force on hi; the steric shell diameter
// uses r_lo only, as the reference does.
__device__ __forceinline__ double short_range(const PairParams& pp, double r, double rl, double rh) {
    const double dDelta = 2 * (rl + pp.eff_len);
    double sf = 0;
    if (r <= dDelta) sf = (2 * rh / (rl + rh)) * pp.ster_fac * rl * rl * log(dDelta / r);
    sf += (rl * rh / (rl + rh)) * pp.el_fac * exp(-pp.kappa * (r - rl - rh));
    const double q1 = (r - rl - rh) * (r + rl + rh);
    const double q2 = r * r - (rl - rh) * (rl - rh);
    sf += -pp.hamaker * r * pow(4 * rl * rh, 3.0) / (6 * (q1 * q1) * (q2 * q2));
    return sf;
}

// The replica's Ewald constants and FP32 units in a PairParams copy.
__device__ __forceinline__ void apply_rep(PairParams& p, const Rep& R) {
    p.alpha = R.alpha;
    p.alpha2 = R.alpha2;
    p.alpha4 = R.alpha4;
    p.rc = R.rc;
    p.rc_sr = R.rc_sr;
    p.ell = R.ell;
    p.inv_ell = R.inv_ell;
    p.inv_mref = R.inv_mref;
    p.bscale = R.bscale;
    p.fscale = R.fscale;
    p.pscale = R.pscale;
    p.al_s = R.al_s;
    p.al2_s = R.al2_s;
    p.al4_s = R.al4_s;
}

// Pair sums cross wavefronts as two 64-bit fixed-point words,
// x = H s1 + L s2 with s1 = 2^e (per replica and quantity) and s2 = s1 2^-40:
// integer atomics, so the totals do not depend on the order in which the
// wavefronts add them. Both words are read off the binary64 pattern of
// v + 1.5 2^52, which for |v| < 2^51 holds round(v) in its low bits; the
// residual x - H s1 is exact (Sterbenz). Returns false when |x| / s1 is not
// below 2^50 (or x is not finite).
constexpr double FX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
constexpr double FX_LO = 1099511627776.0;        // 2^40 = s1 / s2
__device__ __forceinline__ bool fx_add(unsigned long long* __restrict__ hi, unsigned long long* __restrict__ lo,
                                       double x, double s1, double is1) {
    const double t = fma(x, is1, FX_MAGIC);
    const double h = t - FX_MAGIC;
    const double r = fma(-h, s1, x);
    const double t2 = fma(r, is1 * FX_LO, FX_MAGIC);
    const long long H = __double_as_longlong(t) - __double_as_longlong(FX_MAGIC);
    const long long L = __double_as_longlong(t2) - __double_as_longlong(FX_MAGIC);
    if (H) atomicAdd(hi, (unsigned long long)H);
    if (L) atomicAdd(lo, (unsigned long long)L);
    return fabs(h) < 0x1p50;
}
__device__ __forceinline__ double fx_value(unsigned long long hi, unsigned long long lo, double s1) {
    return fma((double)(long long)hi, s1, (double)(long long)lo * (s1 * (1.0 / FX_LO)));
}

// Add component c (force 0..2, field 3..5) of a pair sum for cell-sorted
// particle k.
__device__ __forceinline__ bool fx_add_comp(const PairOut& out, const Rep& R, int64_t n, int c, int64_t k, double x) {
    const int q = c < 3 ? 1 : 0;  // force, field
    return fx_add(&out.acc[(int64_t)(2 * c) * n + k], &out.acc[(int64_t)(2 * c + 1) * n + k], x, R.fx_s[q],
                  R.fx_is[q]);
}

// Sum over the four lanes of a quad (DPP quad permutations, fixed order:
// every lane gets (v0 + v1) + (v2 + v3) of its quad).
__device__ __forceinline__ float quad_sum(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
    return v;
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double quad_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    return v;
}

// Queue entries: (j << 6) | lane of i; PAIR_PAD marks a padding slot.
constexpr int32_t PAIR_PAD = (int32_t)0x80000000u;

// A quad leader's running j sums: with the strided queue order, a lane group
// meets the consecutive quads of one candidate in consecutive batches, so
// the leader adds their quad sums in registers and writes the window
// accumulator once per candidate (and at the end of the flush) instead of
// once per batch.
// The mixed kernel keeps the run in FP32: it spans the batches of one flush
// (at most 16 quad sums of FP32 terms: one candidate's hits).
template <typename T>
struct JRun {
    int32_t js = -1;  // window slot of the candidate being summed, -1: none
    T s[6] = {0, 0, 0, 0, 0, 0};
};
template <typename IW, typename T>
__device__ __forceinline__ void jrun_flush(IW& W, JRun<T>& run) {
    constexpr int NA = IW::NA;
    if (run.js >= 0)
        for (int d = 0; d < NA; ++d) atomicAdd(&W.jacc[d][run.js], (double)run.s[d]);
    run.js = -1;
    for (int d = 0; d < NA; ++d) run.s[d] = 0;
}

// One queued pair (lane il of the chunk, sorted index j in the window that
// starts at jb, whose cell lies at image im), evaluated once for both particles. The four lanes of a quad
// share j (a candidate's hits are queued in whole quads); j's terms are summed
// over the quad and added once by its first lane (`leader`). Every load is
// issued first (one memory round trip); padding and inactive lanes read
// entry 0's records, which exist, and are masked. Pairs with an endpoint
// moved by the overlap correction are left to k_pairs_moved.
template <bool FORCE, bool COATING, bool MIXED, typename RT, typename WT>
__device__ __forceinline__ void pair_dense(const Geometry& g, const double4* __restrict__ rec_mom, const PairParams& pp,
                                           const IChunk<MIXED, COATING>& C, WT& W, int32_t entry,
                                           bool moved_any, int32_t jb, const WinImage& im, double press[3],
                                           float press32[3], int64_t& npairs, uint32_t s0, JRun<RT>& run) {
    constexpr int FB = FORCE ? 3 : 0;  // first field component of the accumulators
    const bool active = entry >= 0;
    // a quad's first slot holds a hit whenever the quad holds one
    const bool leader = active && (lane_id() & 3) == 0;
    const uint32_t il = entry & 63;
    const uint32_t j = (entry >> 6) & 0x1FFFFFF;  // padding: 0
    const int js = (int)j - jb;
    // the mixed kernel reads its i records through L1 (the chunk's 64 records
    // stay cached), which frees the LDS for the accumulators; the FP64
    // kernels, ALU-bound, keep them in LDS
    constexpr bool kIGlobal = MIXED && !COATING;
    const double4 pi = kIGlobal ? (g.rec_pos + s0)[il] : C.ipos[il];  // chunk base + lane: 32-bit offsets
    const double4 pj = g.rec_pos[j];
    const auto mi = [&] {
        if constexpr (kIGlobal) return (g.rec_mom32 + s0)[il];
        else return C.imom[il];
    }();
    // separation from i to the minimum image of j
    auto sep = [&](double& vx, double& vy, double& vz) {
        if (im.wrap) {
            const double L = g.L, iL = g.inv_L;
            vx = min_image_r(pj.x - pi.x, L, iL);
            vy = min_image_r(pj.y - pi.y, L, iL);
            vz = min_image_r(pj.z - pi.z, L, iL);
        } else {
            vx = (pj.x + im.x) - pi.x;
            vy = (pj.y + im.y) - pi.y;
            vz = (pj.z + im.z) - pi.z;
        }
    };
    const auto mj = [&] {
        if constexpr (MIXED && !COATING) return g.rec_mom32[j];
        else return rec_mom[j];
    }();
    const long long ti = __double_as_longlong(pi.w), tj = __double_as_longlong(pj.w);
    // not if an endpoint moved (flag in the tag's high word)
    bool valid = active && !(moved_any && ((uint32_t)(ti >> 32) | (uint32_t)(tj >> 32)));
    const bool i_lo = (int32_t)ti < (int32_t)tj;
    if constexpr (MIXED && !COATING) {
        // separation in FP64 (positions are absolute), everything after it in
        // FP32 in units of ell = r_c, moments in units of mref (FP32 copies);
        // v from i to j
        double dx, dy, dz;
        sep(dx, dy, dz);
        const float vx = (float)(dx * pp.inv_ell);
        const float vy = (float)(dy * pp.inv_ell);
        const float vz = (float)(dz * pp.inv_ell);
        const float r2 = vx * vx + vy * vy + vz * vz;
        valid = valid && r2 < 1.0f;
        // the pair count from compare masks (a ballot of the combined
        // condition would first turn it into a 0/1 value)
        npairs += __popcll(ballot(active) & ballot(r2 < 1.0f) &
                           (moved_any ? ballot(((uint32_t)(ti >> 32) | (uint32_t)(tj >> 32)) == 0) : ~0ull));
        const float irs = __builtin_amdgcn_rsqf(r2);
        const float rs = r2 * irs;
        // every lane evaluates (no branch); the terms of lanes without a pair
        // are exact zeros, so the quad sums need no selects
        float Bi[3], Bj[3], F[3];
        pair_terms32<FORCE>(pp, r2, rs, irs, vx, vy, vz, mi.x, mi.y, mi.z, mj.x, mj.y, mj.z, Bi, Bj, F, valid);
        if (FORCE) {
            // the window's FP32 partials (a few pairs per lane) join the
            // FP64 sums when the window is done
            press32[0] += vx * F[1];
            press32[1] += vx * F[2];
            press32[2] += vy * F[2];
        }
        if (valid) {
            for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[FB + d][il], (double)Bi[d]);
            if (FORCE)
                for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[d][il], -(double)F[d]);
        }
        // j's quad sums, in every lane, added to every lane's run before the
        // leader's branch (so each DPP step is one add); when a leader's
        // candidate changes, the run up to this batch goes to the window
        // accumulator and the new run starts with this batch's sum (only
        // the leaders' runs are ever written)
        constexpr int NA = FORCE ? 6 : 3;
        float qs[6];
        for (int d = 0; d < 3; ++d) qs[FB + d] = quad_sum(Bj[d]);
        if (FORCE)
            for (int d = 0; d < 3; ++d) qs[d] = quad_sum(F[d]);
        // a leader whose candidate changes writes its run and starts a new
        // one with this batch's sum; the run is updated by selects, so the
        // compiler keeps it in place (no copies of a branch's result)
        const bool change = leader && js != run.js;
        if (change && run.js >= 0)
            for (int d = 0; d < NA; ++d) atomicAdd(&W.jacc[d][run.js], (double)run.s[d]);
        for (int d = 0; d < NA; ++d) run.s[d] = change ? qs[d] : run.s[d] + qs[d];
        run.js = change ? js : run.js;
    } else if constexpr (!COATING) {
        // FP64 without coating: u from i to j. The reference orients every
        // pair from the lower to the higher index (src/interaction.cpp:99-104);
        // the separation and each term here are odd or even in u, so the other
        // orientation changes them only by a sign (and by rounding where a sum
        // of two products is contracted the other way round), and no select
        // on the orientation is needed.
        double vx, vy, vz;
        sep(vx, vy, vz);
        const double r2 = vx * vx + vy * vy + vz * vz;
        const double ir = rsqrt64(r2);
        const double r = r2 * ir;
        valid = valid && r < pp.rc;
        npairs += __popcll(ballot(valid));
        const double ux = vx * ir, uy = vy * ir, uz = vz * ir;
        double Bi[3], Bj[3], F[3];  // field on i, on j; force on j
        pair_terms<FORCE, MIXED>(pp, r, ir, ux, uy, uz, mi, mj, Bi, Bj, F);
        if (valid) {
            for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[FB + d][il], Bi[d]);
            if (FORCE) {
                const double rr = MIXED ? r * pp.inv_ell : r;
                press[0] += rr * ux * F[1];
                press[1] += rr * ux * F[2];
                press[2] += rr * uy * F[2];
                for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[d][il], -F[d]);
            }
        }
        double b[3], f[3];
        for (int d = 0; d < 3; ++d) b[d] = quad_sum(valid ? Bj[d] : 0.0);
        if (FORCE)
            for (int d = 0; d < 3; ++d) f[d] = quad_sum(valid ? F[d] : 0.0);
        // the same run update by selects as the mixed path's
        constexpr int NA = FORCE ? 6 : 3;
        const bool change = leader && js != run.js;
        if (change && run.js >= 0)
            for (int d = 0; d < NA; ++d) atomicAdd(&W.jacc[d][run.js], run.s[d]);
        for (int d = 0; d < 3; ++d) run.s[FB + d] = change ? b[d] : run.s[FB + d] + b[d];
        if (FORCE)
            for (int d = 0; d < 3; ++d) run.s[d] = change ? f[d] : run.s[d] + f[d];
        run.js = change ? js : run.js;
    } else {
        // canonical lo -> hi orientation of CompInteractVecs (src/interaction.cpp:99-104):
        // the short-range terms take the lo particle's radius first
        const double L = g.L, iL = g.inv_L;
        const double vx = min_image_r(i_lo ? pj.x - pi.x : pi.x - pj.x, L, iL);
        const double vy = min_image_r(i_lo ? pj.y - pi.y : pi.y - pj.y, L, iL);
        const double vz = min_image_r(i_lo ? pj.z - pi.z : pi.z - pj.z, L, iL);
        const double r2 = vx * vx + vy * vy + vz * vz;
        const double ir = rsqrt64(r2);
        const double r = r2 * ir;
        valid = valid && r < pp.rc;
        npairs += __popcll(ballot(valid));
        const double ux = vx * ir, uy = vy * ir, uz = vz * ir;  // u from lo to hi
        double Bl[3], Bh[3], F[3];  // field on lo, on hi; force on hi
        pair_terms<FORCE, MIXED>(pp, r, ir, ux, uy, uz, i_lo ? mi : mj, i_lo ? mj : mi, Bl, Bh, F);
        if (valid) {
            for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[FB + d][il], i_lo ? Bl[d] : Bh[d]);
            if (FORCE) {
                const double rr = MIXED ? r * pp.inv_ell : r;
                press[0] += rr * ux * F[1];
                press[1] += rr * ux * F[2];
                press[2] += rr * uy * F[2];
                if (COATING && r < pp.rc_sr) {
                    const double rml = pp.r_mag[i_lo ? (int32_t)ti : (int32_t)tj];
                    const double rmh = pp.r_mag[i_lo ? (int32_t)tj : (int32_t)ti];
                    const double sf = short_range(pp, r, rml, rmh) / pp.fscale;
                    F[0] += sf * ux;
                    F[1] += sf * uy;
                    F[2] += sf * uz;
                    press[0] += rr * ux * (sf * uy);
                    press[1] += rr * ux * (sf * uz);
                    press[2] += rr * uy * (sf * uz);
                }
                for (int d = 0; d < 3; ++d) atomicAdd(&W.acc[d][il], i_lo ? -F[d] : F[d]);
            }
        }
        for (int d = 0; d < 3; ++d) {
            const double b = quad_sum(valid ? (i_lo ? Bh[d] : Bl[d]) : 0.0);
            if (leader) atomicAdd(&W.jacc[FB + d][js], b);
        }
        if (FORCE)
            for (int d = 0; d < 3; ++d) {
                const double f = quad_sum(valid ? (i_lo ? F[d] : -F[d]) : 0.0);
                if (leader) atomicAdd(&W.jacc[d][js], f);
            }
    }
}

// Evaluate the queued pairs (qn slots, whole quads) in ceil(qn / 64)
// batches; batch b takes quads b, b + nb, b + 2 nb, ... Afterwards the used
// slots are padding again.
template <bool FORCE, bool COATING, bool MIXED>
__device__ __forceinline__ void pair_flush(const Geometry& g, const double4* __restrict__ rec_mom,
                                           const PairParams& pp, const IChunk<MIXED, COATING>& C,
                                           IWave<FORCE, pair_qflush<FORCE, MIXED>()>& W, int qn, bool moved_any,
                                           int32_t jb,
                                           const WinImage& im, double press[3], float press32[3], int64_t& npairs,
                                           int32_t s0) {
    wave_sync();
    const int lane = lane_id();
    const int nq = qn >> 2;
    const int nb = (nq + WAVE / 4 - 1) / (WAVE / 4);
    JRun<typename std::conditional<MIXED && !COATING, float, double>::type> run;
    // qn <= QCAP, a multiple of 64, so nb <= QCAP / 64 and every
    // slot read, 4 qi + 3 <= 64 nb - 1, is inside the queue; the slots from
    // qn on hold padding
    for (int b = 0; b < nb; ++b) {
        const int qi = (lane >> 2) * nb + b;
        pair_dense<FORCE, COATING, MIXED>(g, rec_mom, pp, C, W, W.q[4 * qi + (lane & 3)], moved_any, jb, im, press,
                                          press32, npairs, (uint32_t)s0, run);
    }
    jrun_flush(W, run);
    wave_sync();
    for (int p = lane; p < qn; p += WAVE) W.q[p] = PAIR_PAD;
    wave_sync();
}

template <bool FORCE, bool COATING, bool MIXED>
__device__ __forceinline__ void pairs_block(Geometry g, const double4* __restrict__ rec_mom, PairParams pp,
                                            PairOut out, const int4* __restrict__ stencil_all,
                                            unsigned long long* __restrict__ cnt_pairs) {
    constexpr bool F32 = MIXED && !COATING;
    using IW = IWave<FORCE, pair_qflush<FORCE, MIXED>()>;
    __shared__ IChunk<MIXED, COATING> C;
    __shared__ IW waves[PAIR_SLICES];
    const int wv = wave_id(), lane = lane_id();
    IW& W = waves[wv];
    const int32_t w = blockIdx.x;
    const int32_t nwork = *g.nwork;
    if (w >= nwork) return;  // uniform over the block
    const int32_t item = g.work[w];
    const int32_t cell = item >> 8, chunk = item & 255;
    // the replica of this cell: its box, grid, stencil and Ewald constants
    const Rep& R = g.reps[g.cell_rep[cell]];
    g.L = R.L;
    g.inv_L = 1.0 / R.L;
    apply_rep(pp, R);
    const int4* stencil = stencil_all + R.sten_off;  // [0] is the cell itself, then the positive half
    const int32_t nhalf = R.nhalf;
    const int64_t n = g.n;
    const int32_t s0 = g.cell_start[cell] + chunk * WAVE;
    const int32_t e0 = min(s0 + WAVE, g.cell_end[cell]);
    const int32_t k = s0 + lane;
    const bool vi = k < e0;
    double4 pi = make_double4(0, 0, 0, 0);
    if (vi) {
        pi = g.rec_pos[k];
        if constexpr (!(MIXED && !COATING)) {  // the mixed kernel reads them through L1
            if (wv == 0) {
                C.ipos[lane] = pi;
                C.imom[lane] = rec_mom[k];
            }
        }
    }
    for (int d = 0; d < IW::NA; ++d) {
        W.acc[d][lane] = 0.0;
        W.jacc[d][lane] = 0.0;
    }
    for (int p = lane; p < IW::QCAP; p += WAVE) W.q[p] = PAIR_PAD;
    __syncthreads();
    const bool moved_any = *g.n_events > 0;
    const uint64_t vmask = ballot(vi);
    const double L = g.L;
    const int nc = R.ncell;
    const bool allpairs = nc == 1;
    const int lc = cell - R.cell_off;
    const int cx = lc % nc, cy = (lc / nc) % nc, cz = lc / (nc * nc);
    // candidate radius: r_c with a relative slack covering the rounding of
    // the FP32 distance test
    const float rcf = (float)(pp.rc * (1.0 + 1e-5));
    const float rr2f = rcf * rcf;
    const float Lf = (float)L;
    const float pxf = (float)pi.x, pyf = (float)pi.y, pzf = (float)pi.z;
    // bounding box of the chunk's particles: a j farther than the candidate
    // radius from the box is farther from every i, so the per-i tests skip it.
    // Componentwise |fl(q - p)| >= the box gap for every p in the box
    // (rounding is monotone); the slack on the squared radius covers the
    // rounding of the sums of squares.
    // The bounds are uniform over the wave and kept in scalar registers.
    const float big = 3.0e38f;
    auto wmax = [](float v) {
        return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, wave_max(v))));
    };
    const float bx0 = -wmax(vi ? -pxf : -big), bx1 = wmax(vi ? pxf : -big);
    const float by0 = -wmax(vi ? -pyf : -big), by1 = wmax(vi ? pyf : -big);
    const float bz0 = -wmax(vi ? -pzf : -big), bz1 = wmax(vi ? pzf : -big);
    const float rb2f = rr2f * 1.0001f;
    double press[3] = {0, 0, 0};
    float press32[3] = {0.0f, 0.0f, 0.0f};  // F32: the current window's pressure partials
    int64_t npairs = 0;
    int qn = 0;
    bool ok = true;
    // with one cell per box side every wavefront takes the cell itself (the
    // only slice) and a share of its windows
    for (int si = allpairs ? 0 : wv; si < nhalf; si += PAIR_SLICES) {
        const bool self = si == 0;
        const int4 o = stencil[si];
        int jx = cx + o.x, jy = cy + o.y, jz = cz + o.z;
        // periodic image of the neighbour cell nearest this cell
        double sx = 0, sy = 0, sz = 0;
        if (jx < 0) { jx += nc; sx = -L; } else if (jx >= nc) { jx -= nc; sx = L; }
        if (jy < 0) { jy += nc; sy = -L; } else if (jy >= nc) { jy -= nc; sy = L; }
        if (jz < 0) { jz += nc; sz = -L; } else if (jz >= nc) { jz -= nc; sz = L; }
        const int c = R.cell_off + (jz * nc + jy) * nc + jx;
        const WinImage im{sx, sy, sz, allpairs};
        const int32_t ce = g.cell_end[c];
        // in the cell itself only j > i: the blocks from this chunk on
        const int32_t cs = self ? s0 : g.cell_start[c];
        const SliceWindows sw = slice_windows(allpairs, wv, cs, ce, g.cell_start[c], chunk);
        for (int32_t wi = sw.first; wi < sw.count; wi += sw.step) {
            const int32_t jb = sw.start(wi);
            // the chunk's own window takes i < j only (later windows of the
            // cell take every lane, earlier ones are not met)
            const bool selfw = self && jb == s0;
            const int32_t jl = jb + lane;
            // j's image nearest this cell (exact: every j within r_c of an i
            // of this cell lies in exactly one stencil cell), in FP32 for the
            // distance test; the exact r < r_c decision is made on the
            // evaluated pair
            float qx = 0, qy = 0, qz = 0;
            bool near = false;
            if (jl < ce) {
                const float4 pj = g.rec_pos32[jl];
                qx = pj.x + (float)sx;
                qy = pj.y + (float)sy;
                qz = pj.z + (float)sz;
                const float ex = fmaxf(fmaxf(bx0 - qx, qx - bx1), 0.0f);
                const float ey = fmaxf(fmaxf(by0 - qy, qy - by1), 0.0f);
                const float ez = fmaxf(fmaxf(bz0 - qz, qz - bz1), 0.0f);
                near = allpairs || ex * ex + ey * ey + ez * ez < rb2f;
            }
            const int cnt = min(WAVE, ce - jb);
            bool any = false;
            // the candidates near the box, in increasing order (uniform
            // mask); the cell itself (one stencil cell of a wavefront's
            // dozens) takes only i < j and has its own copy of the loop
            auto scan = [&](auto self_c) {
                constexpr bool SELF = decltype(self_c)::value;
                for (uint64_t cm = ballot(near); cm; cm &= cm - 1) {
                    const int t = __builtin_ctzll(cm);
                    float dx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qx), t)) - pxf;
                    float dy = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qy), t)) - pyf;
                    float dz = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qz), t)) - pzf;
                    if (allpairs) {
                        dx = dx > Lf / 2 ? dx - Lf : (dx < -Lf / 2 ? dx + Lf : dx);
                        dy = dy > Lf / 2 ? dy - Lf : (dy < -Lf / 2 ? dy + Lf : dy);
                        dz = dz > Lf / 2 ? dz - Lf : (dz < -Lf / 2 ? dz + Lf : dz);
                    }
                    // lanes that take the pair: valid i, inside, and in the
                    // cell itself only i < j (lane < jb + t - s0); the masks
                    // are combined as scalars
                    const bool inside = dx * dx + dy * dy + dz * dz < rr2f;
                    uint64_t m = ballot(inside) & vmask;
                    bool hit = inside && vi;  // the lanes of m
                    if constexpr (SELF) {
                        const int32_t lo = jb + t - s0;
                        m &= lo >= WAVE ? ~0ull : (lo <= 0 ? 0ull : (1ull << lo) - 1);
                        hit = hit && lane < lo;
                    }
                    if (m) {
                        // the hits of one candidate fill whole quads
                        any = true;
                        if (hit) W.q[qn + mbcnt(m)] = ((jb + t) << 6) | lane;
                        qn += (__popcll(m) + 3) & ~3;
                        if (qn >= IW::QFLUSH) {
                            pair_flush<FORCE, COATING, MIXED>(g, rec_mom, pp, C, W, qn, moved_any, jb, im, press,
                                                              press32, npairs, s0);
                            qn = 0;
                        }
                    }
                }
            };
            if (selfw)
                scan(std::true_type{});
            else
                scan(std::false_type{});
            if (any) {
                // window done: its j sums go to the global sums
                if (qn > 0)
                    pair_flush<FORCE, COATING, MIXED>(g, rec_mom, pp, C, W, qn, moved_any, jb, im, press, press32,
                                                      npairs, s0);
                qn = 0;
                if (lane < cnt)
                    for (int d = 0; d < IW::NA; ++d)
                        ok &= fx_add_comp(out, R, n, FORCE ? d : d + 3, jl, W.jacc[d][lane]);
                for (int d = 0; d < IW::NA; ++d) W.jacc[d][lane] = 0.0;
                if (F32 && FORCE)
                    for (int d = 0; d < 3; ++d) {
                        press[d] += (double)press32[d];
                        press32[d] = 0.0f;
                    }
                wave_sync();
            }
        }
    }
    if (FORCE) {
        const double p0 = wave_sum(press[0]), p1 = wave_sum(press[1]), p2 = wave_sum(press[2]);
        if (lane < 3)
            out.press_partial[3 * ((int64_t)w * PAIR_SLICES + wv) + lane] =
                (lane == 0 ? p0 : lane == 1 ? p1 : p2) * pp.pscale;
    }
    if (cnt_pairs && lane == 0) atomicAdd(cnt_pairs, (unsigned long long)(2 * npairs));  // ordered pairs (npairs is the wave's)
    // the wavefronts' i sums, in a fixed order
    __syncthreads();
    if (wv == 0 && vi)
        for (int d = 0; d < IW::NA; ++d) {
            const double x = (waves[0].acc[d][lane] + waves[1].acc[d][lane]) +
                              (waves[2].acc[d][lane] + waves[3].acc[d][lane]);
            ok &= fx_add_comp(out, R, n, FORCE ? d : d + 3, k, x);
        }
    if (!ok) atomicOr(out.err, 64);
}

// Register budgets: the FP64 field pass is held to 5 waves per SIMD and the
// FP64 force pass to 4 (128 VGPRs with a 16-byte spill, from 134 and 3
// waves: 3.27 -> 2.94 ms at C3); the mixed kernels fit 5 waves unforced.
template <bool FORCE, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED ? 1 : (FORCE ? 4 : 5)))) void k_pairs(Geometry g, const double4* __restrict__ rec_mom, PairParams pp,
                                               PairOut out, const int4* __restrict__ stencil_all,
                                               unsigned long long* __restrict__ cnt_pairs) {
    pairs_block<FORCE, false, MIXED>(g, rec_mom, pp, out, stencil_all, cnt_pairs);
}

// With the coating potential (short-range forces, force pass only).
template <bool MIXED>
__global__ __launch_bounds__(256) void k_pairs_coat(Geometry g, const double4* __restrict__ rec_mom, PairParams pp,
                                                    PairOut out, const int4* __restrict__ stencil_all,
                                                    unsigned long long* __restrict__ cnt_pairs) {
    pairs_block<true, true, MIXED>(g, rec_mom, pp, out, stencil_all, cnt_pairs);
}

// ------------------------------------------------------------------ pair lists
// A pair list replaces the cell search of every evaluation by one search per
// build: the list of a replica holds, for each work item and wavefront, the
// queue the pair kernel's search would produce over the build positions with
// the list radius r_l = r_c + skin, as a stream of windows (two header words
// and the window's queue entries). While no particle is farther than skin / 2
// from its build position, every pair within r_c is in the list; the
// evaluation keeps the exact r < r_c test on the current records, so the extra
// pairs give exact zeros. The build runs the same search as the pair kernel
// (work items, half stencil split over the wavefronts, window order), twice:
// pass 0 counts each stream's words, pass 1 writes them at the offsets
// k_list_scan computes per replica.
constexpr int LIST_QFLUSH = 512;
constexpr int LIST_QCAP = LIST_QFLUSH + WAVE;
// batches per flush of the evaluation from the lists
constexpr int LIST_NB = 6;

__device__ __forceinline__ double image_shift(int code, double L) { return code == 1 ? L : (code == 2 ? -L : 0.0); }

template <bool FILL>
__global__ __launch_bounds__(256) void k_list_pass(Geometry g, const int4* __restrict__ stencil_all,
                                                   int32_t* __restrict__ lcount, const int32_t* __restrict__ loff,
                                                   int32_t* __restrict__ list) {
    __shared__ int32_t qs[PAIR_SLICES][LIST_QCAP];
    const int wv = wave_id(), lane = lane_id();
    const int32_t w = blockIdx.x;
    if (w >= *g.nwork) return;  // uniform over the block
    const int32_t item = g.work[w];
    const int32_t cell = item >> 8, chunk = item & 255;
    const int32_t rep = g.cell_rep[cell];
    const int32_t mode = g.list_mode[rep];
    if (mode != LIST_FRESH && mode != LIST_RESTORE) return;
    const Rep& R = g.reps[rep];
    const int4* stencil = stencil_all + R.sten_off;
    const int32_t nhalf = R.nhalf;
    const int32_t s0 = g.cell_start[cell] + chunk * WAVE;
    const int32_t e0 = min(s0 + WAVE, g.cell_end[cell]);
    const bool vi = s0 + lane < e0;
    int32_t* q = qs[wv];
    for (int p = lane; p < LIST_QCAP; p += WAVE) q[p] = PAIR_PAD;
    wave_sync();
    const float4 pb = vi ? g.rec_build32[s0 + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint64_t vmask = ballot(vi);
    const double L = R.L;
    const int nc = R.ncell;
    const bool allpairs = nc == 1;
    const int lc = cell - R.cell_off;
    const int cx = lc % nc, cy = (lc / nc) % nc, cz = lc / (nc * nc);
    // list radius with the pair kernel's FP32 slack
    const float rcf = (float)(R.rlist * (1.0 + 1e-5));
    const float rr2f = rcf * rcf;
    const float Lf = (float)L;
    const float pxf = pb.x, pyf = pb.y, pzf = pb.z;
    const float big = 3.0e38f;
    auto wmax = [](float v) {
        return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, wave_max(v))));
    };
    const float bx0 = -wmax(vi ? -pxf : -big), bx1 = wmax(vi ? pxf : -big);
    const float by0 = -wmax(vi ? -pyf : -big), by1 = wmax(vi ? pyf : -big);
    const float bz0 = -wmax(vi ? -pzf : -big), bz1 = wmax(vi ? pzf : -big);
    const float rb2f = rr2f * 1.0001f;
    const int64_t sidx = (int64_t)PAIR_SLICES * s0 + wv;
    // absolute word index of the stream (FILL), or words so far (count)
    int64_t cur = FILL ? R.list_base + loff[sidx] : 0;
    const int64_t wend = FILL ? cur + lcount[sidx] : 0;
    int64_t whdr = -1;  // the current window's header words, once it has entries
    int qn = 0;
    auto emit = [&]() {
        wave_sync();
        if (whdr < 0) {
            whdr = cur;
            cur += LIST_HDR;
        }
        if (FILL)
            for (int p = lane; p < qn; p += WAVE)
                if (cur + p < wend) list[cur + p] = q[p];
        cur += qn;
        wave_sync();
        for (int p = lane; p < qn; p += WAVE) q[p] = PAIR_PAD;
        wave_sync();
        qn = 0;
    };
    for (int si = allpairs ? 0 : wv; si < nhalf; si += PAIR_SLICES) {
        const bool self = si == 0;
        const int4 o = stencil[si];
        int jx = cx + o.x, jy = cy + o.y, jz = cz + o.z;
        int ix = 0, iy = 0, iz = 0;  // image codes: 1 = +L, 2 = -L
        if (jx < 0) { jx += nc; ix = 2; } else if (jx >= nc) { jx -= nc; ix = 1; }
        if (jy < 0) { jy += nc; iy = 2; } else if (jy >= nc) { jy -= nc; iy = 1; }
        if (jz < 0) { jz += nc; iz = 2; } else if (jz >= nc) { jz -= nc; iz = 1; }
        const float sx = (float)image_shift(ix, L), sy = (float)image_shift(iy, L), sz = (float)image_shift(iz, L);
        const int img = ix | (iy << 2) | (iz << 4);
        const int c = R.cell_off + (jz * nc + jy) * nc + jx;
        const int32_t ce = g.cell_end[c];
        const int32_t cs = self ? s0 : g.cell_start[c];
        const SliceWindows sw = slice_windows(allpairs, wv, cs, ce, g.cell_start[c], chunk);
        for (int32_t wi = sw.first; wi < sw.count; wi += sw.step) {
            const int32_t jb = sw.start(wi);
            const bool selfw = self && jb == s0;
            const int32_t jl = jb + lane;
            float qx = 0, qy = 0, qz = 0;
            bool near = false;
            if (jl < ce) {
                const float4 pj = g.rec_build32[jl];
                qx = pj.x + sx;
                qy = pj.y + sy;
                qz = pj.z + sz;
                const float ex = fmaxf(fmaxf(bx0 - qx, qx - bx1), 0.0f);
                const float ey = fmaxf(fmaxf(by0 - qy, qy - by1), 0.0f);
                const float ez = fmaxf(fmaxf(bz0 - qz, qz - bz1), 0.0f);
                near = allpairs || ex * ex + ey * ey + ez * ez < rb2f;
            }
            const int cnt = min(WAVE, ce - jb);
            bool any = false;
            for (uint64_t cm = ballot(near); cm; cm &= cm - 1) {
                const int t = __builtin_ctzll(cm);
                float dx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qx), t)) - pxf;
                float dy = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qy), t)) - pyf;
                float dz = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, qz), t)) - pzf;
                if (allpairs) {
                    dx = dx > Lf / 2 ? dx - Lf : (dx < -Lf / 2 ? dx + Lf : dx);
                    dy = dy > Lf / 2 ? dy - Lf : (dy < -Lf / 2 ? dy + Lf : dy);
                    dz = dz > Lf / 2 ? dz - Lf : (dz < -Lf / 2 ? dz + Lf : dz);
                }
                const bool inside = dx * dx + dy * dy + dz * dz < rr2f;
                uint64_t m = ballot(inside) & vmask;
                bool hit = inside && vi;
                if (selfw) {
                    // in the chunk's own window only i < j
                    const int32_t lo = jb + t - s0;
                    m &= lo >= WAVE ? ~0ull : (lo <= 0 ? 0ull : (1ull << lo) - 1);
                    hit = hit && lane < lo;
                }
                if (m) {
                    any = true;
                    if (hit) q[qn + mbcnt(m)] = ((jb + t) << 6) | lane;
                    qn += (__popcll(m) + 3) & ~3;
                    if (qn >= LIST_QFLUSH) emit();
                }
            }
            if (any) {
                if (qn > 0) emit();
                if (FILL && lane == 0 && whdr + 1 < wend) {
                    list[whdr] = jb;
                    list[whdr + 1] = (int32_t)(((cur - whdr - LIST_HDR) << 13) | ((int64_t)cnt << 6) | img);
                }
                whdr = -1;
            }
        }
    }
    if (!FILL && lane == 0) lcount[sidx] = (int32_t)cur;
}

// Offsets of the streams of each replica being built (one block per
// replica, over its sorted slots x wavefronts), its word count, and the
// region check.
__global__ __launch_bounds__(1024) void k_list_scan(const Rep* __restrict__ reps, int32_t* __restrict__ list_mode,
                                                    int32_t* __restrict__ list_state,
                                                    const int32_t* __restrict__ lcount, int32_t* __restrict__ loff,
                                                    long long* __restrict__ list_used, int32_t* __restrict__ err) {
    __shared__ long long part[1024];
    const int32_t r = blockIdx.x;
    const int32_t mode = list_mode[r];
    if (mode != LIST_FRESH && mode != LIST_RESTORE) return;  // uniform over the block
    const Rep& R = reps[r];
    const int64_t b0 = PAIR_SLICES * R.off, m = PAIR_SLICES * R.n;
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    const int64_t b = t * per < m ? t * per : m, e = b + per < m ? b + per : m;
    long long sum = 0;
    for (int64_t i = b; i < e; ++i) sum += lcount[b0 + i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - sum;
    for (int64_t i = b; i < e; ++i) {
        loff[b0 + i] = (int32_t)(run < 0x7fffffff ? run : 0x7fffffff);
        run += lcount[b0 + i];
    }
    if (t == 1023) {
        const long long total = part[1023];
        list_used[r] = total;
        if (total > R.list_cap) {
            list_mode[r] = LIST_BAD;
            list_state[r] = LIST_INVALID;
            atomicOr(err, 128);
        }
    }
}

// Real-space pass from the pair lists: the pair kernel's block (64 i
// particles of a work item, four wavefronts, the i sums added in a fixed
// order at the end) with each wavefront's candidate search replaced by its
// stream. A window's entries are read in batches of up to the kernel's queue
// length, in the pair kernel's strided order (batch b of a flush takes quads
// b, b + nb, ...), and the window's j sums go to the fixed-point sums at its
// end, as in the pair kernel.
template <bool FORCE>
struct LWave {
    static constexpr int NA = FORCE ? 6 : 3;
    double acc[NA][WAVE];
    double jacc[NA][WAVE];
};

template <bool FORCE, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED ? 1 : (FORCE ? 4 : 5)))) void k_pairs_list(
    Geometry g, const double4* __restrict__ rec_mom, PairParams pp, PairOut out,
    unsigned long long* __restrict__ cnt_pairs) {
    constexpr bool F32 = MIXED;
    using LW = LWave<FORCE>;
    __shared__ IChunk<MIXED, false> C;
    __shared__ LW waves[PAIR_SLICES];
    const int wv = wave_id(), lane = lane_id();
    LW& W = waves[wv];
    const int32_t w = blockIdx.x;
    if (w >= *g.nwork) return;  // uniform over the block
    const int32_t item = g.work[w];
    const int32_t cell = item >> 8, chunk = item & 255;
    const int32_t rep = g.cell_rep[cell];
    if (g.list_mode[rep] == LIST_BAD) return;  // flagged: no list this step
    const Rep& R = g.reps[rep];
    g.L = R.L;
    g.inv_L = 1.0 / R.L;
    apply_rep(pp, R);
    const int64_t n = g.n;
    const int32_t s0 = g.cell_start[cell] + chunk * WAVE;
    const int32_t e0 = min(s0 + WAVE, g.cell_end[cell]);
    const int32_t k = s0 + lane;
    const bool vi = k < e0;
    if constexpr (!MIXED) {  // the mixed kernel reads the i records through L1
        if (vi && wv == 0) {
            C.ipos[lane] = g.rec_pos[k];
            C.imom[lane] = rec_mom[k];
        }
    }
    for (int d = 0; d < LW::NA; ++d) {
        W.acc[d][lane] = 0.0;
        W.jacc[d][lane] = 0.0;
    }
    __syncthreads();
    const bool moved_any = *g.n_events > 0;
    const bool allpairs = R.ncell == 1;
    const double L = R.L;
    double press[3] = {0, 0, 0};
    float press32[3] = {0.0f, 0.0f, 0.0f};
    int64_t npairs = 0;
    bool ok = true;
    using RT = typename std::conditional<MIXED, float, double>::type;
    const int64_t sidx = (int64_t)PAIR_SLICES * s0 + wv;
    const int32_t len = __builtin_amdgcn_readfirstlane(g.lcount[sidx]);
    const int32_t* st = g.list + (R.list_base + __builtin_amdgcn_readfirstlane(g.loff[sidx]));
    // entries of a flush (up to LIST_NB batches of 64, in the strided order)
    // and its batch count; the entries of the next flush (the next window's
    // first one included, its header read ahead) are loaded while this one
    // is evaluated, so a batch waits only for its records
    auto flush_shape = [&](int32_t Ew, int32_t fw, int& nq, int& nb) {
        nq = min(LIST_NB * WAVE, Ew - fw) >> 2;
        nb = (nq + WAVE / 4 - 1) / (WAVE / 4);
    };
    auto fetch = [&](int32_t pw, int32_t Ew, int32_t fw, int32_t (&dst)[LIST_NB]) {
        int nq, nb;
        flush_shape(Ew, fw, nq, nb);
        const int g4 = (lane >> 2) * nb;
        const int32_t base = pw + LIST_HDR + fw;
#pragma unroll
        for (int b = 0; b < LIST_NB; ++b) {
            const int qb = g4 + b;
            dst[b] = st[(b < nb && qb < nq) ? base + 4 * qb + (lane & 3) : base];
        }
    };
    // a header that does not fit the stream ends it (never written so)
    auto window_len = [&](int32_t pw, int32_t w) {
        const int32_t Ew = (int32_t)((uint32_t)w >> 13);
        return (Ew >= 4 && Ew <= len - pw - LIST_HDR) ? Ew : -1;
    };
    if (len >= LIST_HDR + 4) {
        int32_t p = 0, f = 0;
        int32_t jb = __builtin_amdgcn_readfirstlane(st[0]), w1 = __builtin_amdgcn_readfirstlane(st[1]);
        int32_t E = window_len(0, w1);
        int32_t pn = 0, jbn = 0, w1n = 0;
        auto read_ahead = [&]() {
            pn = p + LIST_HDR + E;
            if (pn + LIST_HDR <= len) {
                jbn = __builtin_amdgcn_readfirstlane(st[pn]);
                w1n = __builtin_amdgcn_readfirstlane(st[pn + 1]);
            }
        };
        int32_t cur[LIST_NB], nxt[LIST_NB];
        if (E > 0) {
            fetch(0, E, 0, cur);
            read_ahead();
        }
        while (E > 0) {
            // the next flush: the rest of this window, or the next window
            const bool more = f + LIST_NB * WAVE < E;
            const int32_t En = more ? E : (pn + LIST_HDR <= len ? window_len(pn, w1n) : -1);
            // one unconditional fetch (the loads stay in flight through this
            // flush); past the stream's end it reads this flush again
            const bool nw = !more && En > 0;
            fetch(nw ? pn : p, nw ? En : E, more ? f + LIST_NB * WAVE : (nw ? 0 : f), nxt);
            const WinImage im{image_shift(w1 & 3, L), image_shift((w1 >> 2) & 3, L),
                              image_shift((w1 >> 4) & 3, L), allpairs};
            int nq, nb;
            flush_shape(E, f, nq, nb);
            const int g4 = (lane >> 2) * nb;
            JRun<RT> run;
#pragma unroll
            for (int b = 0; b < LIST_NB; ++b)
                if (b < nb)
                    pair_dense<FORCE, false, MIXED>(g, rec_mom, pp, C, W, g4 + b < nq ? cur[b] : PAIR_PAD, moved_any,
                                                    jb, im, press, press32, npairs, (uint32_t)s0, run);
            jrun_flush(W, run);
            if (!more) {
                wave_sync();
                // window done: its j sums go to the global sums
                const int cnt = (w1 >> 6) & 127;
                if (lane < cnt)
                    for (int d = 0; d < LW::NA; ++d)
                        ok &= fx_add_comp(out, R, n, FORCE ? d : d + 3, jb + lane, W.jacc[d][lane]);
                for (int d = 0; d < LW::NA; ++d) W.jacc[d][lane] = 0.0;
                if (F32 && FORCE)
                    for (int d = 0; d < 3; ++d) {
                        press[d] += (double)press32[d];
                        press32[d] = 0.0f;
                    }
                wave_sync();
            }
#pragma unroll
            for (int b = 0; b < LIST_NB; ++b) cur[b] = nxt[b];
            if (more) {
                f += LIST_NB * WAVE;
            } else {
                p = pn;
                jb = jbn;
                w1 = w1n;
                E = En;
                f = 0;
                if (E > 0) read_ahead();
            }
        }
    }
    if (FORCE) {
        const double p0 = wave_sum(press[0]), p1 = wave_sum(press[1]), p2 = wave_sum(press[2]);
        if (lane < 3)
            out.press_partial[3 * ((int64_t)w * PAIR_SLICES + wv) + lane] =
                (lane == 0 ? p0 : lane == 1 ? p1 : p2) * pp.pscale;
    }
    if (cnt_pairs && lane == 0) atomicAdd(cnt_pairs, (unsigned long long)(2 * npairs));
    // the wavefronts' i sums, in a fixed order
    __syncthreads();
    if (wv == 0 && vi)
        for (int d = 0; d < LW::NA; ++d) {
            const double x = (waves[0].acc[d][lane] + waves[1].acc[d][lane]) +
                             (waves[2].acc[d][lane] + waves[3].acc[d][lane]);
            ok &= fx_add_comp(out, R, n, FORCE ? d : d + 3, k, x);
        }
    if (!ok) atomicOr(out.err, 64);
}

// Cell-sorted fixed-point pair sums -> original order, in SI units.
__global__ void k_pairs_out(int64_t n, const int32_t* __restrict__ sorted_idx, const int32_t* __restrict__ rep_of,
                            const Rep* __restrict__ reps, const unsigned long long* __restrict__ acc,
                            double* __restrict__ force, double* __restrict__ field) {
    // the rounded pair sum is added to the k-space value (one rounding of the
    // sum of the two, whichever was written first), so no contraction
#pragma clang fp contract(off)
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t o = sorted_idx[k];
    const Rep& R = reps[rep_of[o]];
    for (int d = 0; d < 3; ++d) {
        const double v = fx_value(acc[(6 + 2 * d) * n + k], acc[(7 + 2 * d) * n + k], R.fx_s[0]) * R.bscale;
        field[d * n + o] += v;
    }
    if (force)
        for (int d = 0; d < 3; ++d) {
            const double v = fx_value(acc[(2 * d) * n + k], acc[(2 * d + 1) * n + k], R.fx_s[1]) * R.fscale;
            force[d * n + o] += v;
        }
}

// Within each cell, the particle list ordered by original index (k_place
// fills `placed` in atomic order; the ordered list goes to sorted_idx). Ranks
// by counting. Blocks [0, cell_blocks) take the cells of at most one
// wavefront of particles, one cell per wave, from registers. The other
// blocks take 64-position chunks of `placed` that lie in larger cells (a
// small box is one cell holding its whole replica): the block's ORDER_SPLIT
// wavefronts each count against one share of the cell, and the shares' counts
// are summed in LDS, so a cell of m particles costs m / (64 ORDER_SPLIT)
// steps on the critical path instead of m^2 / 64.
constexpr int ORDER_SPLIT = 4;

__global__ __launch_bounds__(WAVE * ORDER_SPLIT) void k_cell_order(int64_t n, int32_t ncells, int32_t cell_blocks,
                                                                   const int32_t* __restrict__ key,
                                                                   const int32_t* __restrict__ start,
                                                                   const int32_t* __restrict__ end,
                                                                   const int32_t* __restrict__ placed,
                                                                   int32_t* __restrict__ sorted_idx,
                                                                   const int32_t* __restrict__ skip) {
    if (skip && *skip) return;  // uniform over the block
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    if ((int32_t)blockIdx.x < cell_blocks) {
        for (int32_t c = blockIdx.x * ORDER_SPLIT + wv; c < ncells; c += cell_blocks * ORDER_SPLIT) {
            const int32_t s = start[c], m = end[c] - s;
            if (m <= 0 || m > WAVE) continue;
            const int32_t v = lane < m ? placed[s + lane] : INT32_MAX;
            int r = 0;
            for (int t = 0; t < m; ++t) r += __builtin_amdgcn_readlane(v, t) < v;
            if (lane < m) sorted_idx[s + r] = v;
        }
        return;
    }
    __shared__ int32_t part[ORDER_SPLIT][WAVE];
    const int64_t chunks = (n + WAVE - 1) / WAVE;
    for (int64_t ch = blockIdx.x - cell_blocks; ch < chunks; ch += gridDim.x - cell_blocks) {
        // the cell of position p, from the particle placed there; positions
        // past the last cell hold stale entries and match no cell's range
        const int64_t p = ch * WAVE + lane;
        int32_t v = INT32_MAX, s = 0, m = 0;
        if (p < n) {
            const int32_t x = placed[p];
            if (x >= 0 && x < n) {
                const int32_t k = key[x];
                if (k >= 0 && k < ncells) {
                    const int32_t s1 = start[k], m1 = end[k] - s1;
                    if (p >= s1 && p < (int64_t)s1 + m1) {
                        v = x;
                        s = s1;
                        m = m1;
                    }
                }
            }
        }
        const bool big = m > WAVE;
        uint64_t todo = ballot(big);
        if (!todo) continue;  // the same for every wave of the block
        int32_t r = 0;
        while (todo) {
            // the lanes of one large cell at a time
            const int l = __builtin_ctzll(todo);
            const int32_t sc = __builtin_amdgcn_readlane(s, l), mc = __builtin_amdgcn_readlane(m, l);
            const bool mine = big && s == sc;
            const int32_t share = (mc + ORDER_SPLIT * WAVE - 1) / (ORDER_SPLIT * WAVE) * WAVE;
            const int32_t b0 = wv * share, b1 = min(mc, b0 + share);
            int32_t rr = 0;
            for (int32_t base = b0; base < b1; base += WAVE) {
                const int32_t u = base + lane < b1 ? placed[sc + base + lane] : INT32_MAX;
#pragma unroll 16
                for (int t = 0; t < WAVE; ++t) rr += __builtin_amdgcn_readlane(u, t) < v;
            }
            if (mine) r = rr;
            todo &= ~ballot(mine);
        }
        part[wv][lane] = r;
        __syncthreads();
        if (wv == 0 && big) {
            int32_t tot = 0;
            for (int q = 0; q < ORDER_SPLIT; ++q) tot += part[q][lane];
            sorted_idx[s + tot] = v;
        }
        __syncthreads();
    }
}

// Pairs with at least one endpoint moved by the overlap correction, in FP64
// with the positions the reference's sequential scan gives each pair
// (CompInteractVecs, src/interaction.cpp:99-129): one wavefront per moved
// particle p, lanes over the stencil's particles; a pair of two moved
// particles is done by the lower index's wavefront. Results go into the same
// fixed-point sums as k_pairs' (before k_pairs_out).
template <bool FORCE, bool COATING>
__global__ __launch_bounds__(64) void k_pairs_moved(Geometry g, const double4* __restrict__ rec_mom, PairParams pp,
                                                    PairOut out, const int4* __restrict__ stencil_all,
                                                    int32_t press_off, unsigned long long* __restrict__ cnt_pairs) {
    const int32_t nmv = *g.n_moved;
    const int lane = lane_id();
    const int64_t n = g.n;
    bool ok = true;
    for (int32_t mi = blockIdx.x; mi < nmv; mi += gridDim.x) {
        const int32_t p = g.moved_list[mi];
        const Rep& R = g.reps[g.rep_of[p]];
        g.L = R.L;
        apply_rep(pp, R);
        const double L = R.L;
        const int nc = R.ncell;
        const int4* stencil = stencil_all + R.sten_off;
        const int32_t nsten = R.nsten;
        const int32_t kp = g.inv_perm[p];
        const double4 pp0 = g.rec_pos[kp];
        const double4 mp = rec_mom[kp];
        const int cx = cell_coord(pp0.x, L, R.cell, nc), cy = cell_coord(pp0.y, L, R.cell, nc),
                  cz = cell_coord(pp0.z, L, R.cell, nc);
        double press[3] = {0, 0, 0};
        int64_t npairs = 0;
        for (int si = 0; si < nsten; ++si) {
            const int4 o = stencil[si];
            const int c = R.cell_off + (((cz + o.z + nc) % nc) * nc + (cy + o.y + nc) % nc) * nc + (cx + o.x + nc) % nc;
            const int32_t cs = g.cell_start[c], ce = g.cell_end[c];
            for (int32_t j = cs + lane; j < ce; j += WAVE) {
                if (j == kp) continue;
                const int32_t oj = g.sorted_idx[j];
                if (g.ev_head[oj] >= 0 && oj < p) continue;  // done by oj's wavefront
                const double4 pj = g.rec_pos[j];
                const bool p_lo = p < oj;
                const int32_t lo = p_lo ? p : oj, hi = p_lo ? oj : p;
                double xl = p_lo ? pp0.x : pj.x, yl = p_lo ? pp0.y : pj.y, zl = p_lo ? pp0.z : pj.z;
                double xh = p_lo ? pj.x : pp0.x, yh = p_lo ? pj.y : pp0.y, zh = p_lo ? pj.z : pp0.z;
                int event = 0;
                version_pair(g, lo, hi, xl, yl, zl, xh, yh, zh, event);
                const double vx = min_image(xh - xl, L), vy = min_image(yh - yl, L), vz = min_image(zh - zl, L);
                const double r0 = sqrt(vx * vx + vy * vy + vz * vz);
                if (!(r0 < pp.rc)) continue;
                ++npairs;
                const double rmp = pp.r_mag[p], rmj = pp.r_mag[oj];
                const double r = event ? rmp + rmj + 1e-10 : r0;
                const double ir0 = 1.0 / r0;
                const double ir = event ? 1.0 / r : ir0;
                // u from lo to hi, taken from the pre-correction separation
                const double ux = vx * ir0, uy = vy * ir0, uz = vz * ir0;
                const double4 mlo = p_lo ? mp : rec_mom[j], mhi = p_lo ? rec_mom[j] : mp;
                double Bl[3], Bh[3], Fh[3];
                // fields on lo and on hi, force on hi (u from lo to hi)
                pair_terms<FORCE, false>(pp, r, ir, ux, uy, uz, mlo, mhi, Bl, Bh, Fh);
                // into the fixed-point sums (units of the pair kernel's sums)
                const double f = pp.fac;
                const int32_t klo = p_lo ? kp : j, khi = p_lo ? j : kp;
                const double cb = f / pp.bscale;
                for (int d = 0; d < 3; ++d) {
                    ok &= fx_add_comp(out, R, n, 3 + d, klo, cb * Bl[d]);
                    ok &= fx_add_comp(out, R, n, 3 + d, khi, cb * Bh[d]);
                }
                if (FORCE) {
                    double sf = 0;
                    if (COATING && r0 < pp.rc_sr) sf = short_range(pp, r, p_lo ? rmp : rmj, p_lo ? rmj : rmp);
                    const double u3[3] = {ux, uy, uz};
                    const double cf = 1.0 / pp.fscale;
                    for (int d = 0; d < 3; ++d) {
                        const double Fd = f * Fh[d] + sf * u3[d];
                        ok &= fx_add_comp(out, R, n, d, klo, -Fd * cf);
                        ok &= fx_add_comp(out, R, n, d, khi, Fd * cf);
                    }
                    press[0] += r * ux * (f * Fh[1] + sf * uy);
                    press[1] += r * ux * (f * Fh[2] + sf * uz);
                    press[2] += r * uy * (f * Fh[2] + sf * uz);
                }
            }
        }
        if (FORCE) {
            const double p0 = wave_sum(press[0]), p1 = wave_sum(press[1]), p2 = wave_sum(press[2]);
            if (lane == 0) {
                // fixed-point words (hi, lo) per component: order-independent
                unsigned long long* w =
                    reinterpret_cast<unsigned long long*>(out.press_partial) + 3 * (int64_t)press_off + 6 * g.rep_of[p];
                ok &= fx_add(w, w + 1, p0, R.px_s, R.px_is);
                ok &= fx_add(w + 2, w + 3, p1, R.px_s, R.px_is);
                ok &= fx_add(w + 4, w + 5, p2, R.px_s, R.px_is);
            }
        }
        if (cnt_pairs) {
            const int64_t np = wave_sum(npairs);
            if (lane == 0) atomicAdd(cnt_pairs, (unsigned long long)(2 * np));
        }
    }
    if (!ok) atomicOr(out.err, 64);
}

// ------------------------------------------------------------------ k space
// cos/sin(2 pi n x_d / L), n = 0..nc, by the Rapaport recurrence of EvalSinCos
// (src/interaction.cpp:521-548). Layout per particle: [d][n][cos, sin], in
// FP64 for a replica with a vector k-space kernel and rounded to FP32 (each
// axis zero padded to tw4_of(nc) floats) for one with a matrix-core kernel.
__global__ void k_sincos_tab(int64_t n, const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                             const double* __restrict__ pos0, double* __restrict__ tab_all,
                             float* __restrict__ tab32_all) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rep& R = reps[rep_of[i]];
    const int32_t nc = R.nc;
    const bool f64 = R.kmode != 3, f32 = R.kmode != 0;
    const int64_t stride = 3 * (int64_t)(nc + 1) * 2;
    const int tw4 = tw4_of(nc);
    double* t = tab_all + R.tab_off + (i - R.off) * stride;
    float* t32 = tab32_all + R.tab32_off + (i - R.off) * 3 * (int64_t)tw4;
    for (int d = 0; d < 3; ++d) {
        const double tt = R.two_pi_over_L * pos0[d * n + i];
        double* td = t + d * (nc + 1) * 2;
        float* td32 = t32 + d * tw4;
        auto put = [&](int k, double c, double s) {
            if (f64) {
                td[2 * k] = c;
                td[2 * k + 1] = s;
            }
            if (f32) {
                td32[2 * k] = (float)c;
                td32[2 * k + 1] = (float)s;
            }
        };
        double s1, c1;
        sincos(tt, &s1, &c1);
        put(0, 1.0, 0.0);
        if (nc >= 1) put(1, c1, s1);
        const double u = 2.0 * c1;
        double cm2 = 1, sm2 = 0, cm1 = c1, sm1 = s1;
        for (int k = 2; k <= nc; ++k) {
            const double ck = u * cm1 - cm2, sk = u * sm1 - sm2;
            put(k, ck, sk);
            cm2 = cm1;
            sm2 = sm1;
            cm1 = ck;
            sm1 = sk;
        }
        if (f32)
            for (int k = 2 * (nc + 1); k < tw4; ++k) td32[k] = 0.0f;
    }
}

// Structure factors S(k) = sum_j (n . m_j) e^{i k.r_j} (the first pass of
// CompDipInteractEwaldF, src/interaction.cpp:459-483), over the half space of
// n vectors. One lane per row segment (ny, nz, |nx| in [alo, ahi]); the
// particles of a chunk are walked in step by the whole wavefront (broadcast
// loads), and the segment's nx values are accumulated in registers through
// the recurrence e^{i(a+1)x} = e^{iax} e^{ix}. Each chunk writes its partial
// sums; k_kpack adds them in a fixed order.
// Replica of item b of a concatenated per-replica range (WHICH: 0 S(k)
// build items, 1 k vectors, 2 packed scatter entries): the last replica whose
// first item is <= b.
template <int WHICH>
__device__ __forceinline__ int find_rep(const Rep* __restrict__ reps, int32_t nrep, int64_t b) {
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int32_t first = WHICH == 0   ? reps[mid].kb_off
                              : WHICH == 1 ? reps[mid].k_off
                              : WHICH == 2 ? reps[mid].pk_off
                              : WHICH == 3 ? reps[mid].kbm_off
                                           : reps[mid].ksm_off;
        if (first <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int SEG>
__global__ __launch_bounds__(64) void k_kbuild(int64_t n_all, const Rep* __restrict__ reps, int32_t nrep,
                                               int32_t kb_items, const KSeg* __restrict__ segs_all,
                                               const double* __restrict__ tab_all, const double* __restrict__ mom,
                                               double* __restrict__ partial_all) {
    // work item = (replica, particle chunk, wavefront of row segments); the
    // segment waves of one chunk are adjacent blocks. (An XCD-contiguous
    // block order and segment classes of 8/4/2 values were both measured
    // slower.) A grid smaller than the items walks them in strides.
  for (int32_t item = blockIdx.x; item < kb_items; item += gridDim.x) {
    const Rep& R = reps[find_rep<0>(reps, nrep, item)];
    if (R.kmode & 1) continue;  // this replica's S(k) is built on the matrix cores
    const int32_t li = item - R.kb_off;
    const int32_t segw = li % R.nsegw, ch = li / R.nsegw;
    const KSeg* segs = segs_all + R.seg_off;
    const int32_t nseg = R.nseg;
    const int32_t nc = R.nc, chunk = R.kchunk, nk = R.nk;
    const double* tab = tab_all + R.tab_off;
    const int64_t n = n_all;
    const int r = segw * WAVE + threadIdx.x;
    const bool vr = r < nseg;
    const KSeg sg = vr ? segs[r] : KSeg{0, 0, 1, 0, 0, 0};
    const int ay = sg.ny < 0 ? -sg.ny : sg.ny;
    const double ysg = sg.ny < 0 ? -1.0 : 1.0;
    const bool half = (sg.ny == 0 && sg.nz == 0);
    const int64_t stride = 3 * (int64_t)(nc + 1) * 2;
    double pc_[SEG], ps_[SEG], mc_[SEG], ms_[SEG];
#pragma unroll
    for (int a = 0; a < SEG; ++a) pc_[a] = ps_[a] = mc_[a] = ms_[a] = 0.0;
    const int64_t jb = (int64_t)ch * chunk;
    const int64_t je = jb + chunk < R.n ? jb + chunk : R.n;
    for (int64_t jr = jb; jr < je; ++jr) {
        const double* t = tab + jr * stride;
        const int64_t j = R.off + jr;
        const double ex = t[2], sx = t[3];
        const double* ty = t + (nc + 1) * 2;
        const double* tz = t + 2 * (nc + 1) * 2;
        const double yc = ty[2 * ay], ys = ysg * ty[2 * ay + 1];
        const double zc = tz[2 * sg.nz], zs = tz[2 * sg.nz + 1];
        const double wc = yc * zc - ys * zs, ws = ys * zc + yc * zs;
        const double mx = mom[j], my = mom[n + j], mz = mom[2 * n + j];
        const double A = sg.ny * my + sg.nz * mz;
        double xc = t[2 * sg.alo], xs = t[2 * sg.alo + 1];
#pragma unroll
        for (int q = 0; q < SEG; ++q) {
            const int a = sg.alo + q;
            if (a <= sg.ahi) {
                const double p1 = xc * wc, p2 = xs * ws, p3 = xc * ws, p4 = xs * wc;
                const double qp = a * mx + A, qm = -a * mx + A;
                pc_[q] += qp * (p1 - p2);
                ps_[q] += qp * (p3 + p4);
                mc_[q] += qm * (p1 + p2);
                ms_[q] += qm * (p3 - p4);
            }
            const double nxc = xc * ex - xs * sx, nxs = xs * ex + xc * sx;
            xc = nxc;
            xs = nxs;
        }
    }
    if (!vr) continue;
    double* out = partial_all + R.part_off + (int64_t)ch * nk * 2;
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        const int a = sg.alo + q;
        if (a > sg.ahi) break;
        const int kp = half ? sg.koff + a - 1 : sg.koff + a + sg.nxmax;
        out[2 * kp] = pc_[q];
        out[2 * kp + 1] = ps_[q];
        if (a >= 1 && !half) {
            const int km = sg.koff - a + sg.nxmax;
            out[2 * km] = mc_[q];
            out[2 * km + 1] = ms_[q];
        }
    }
  }
}

// The same S(k) partials from FP32 matrix-core products (mixed-precision
// path). Per row (ny, nz) and |nx| = a, with X = (c, s) = e^{i a x_j},
// W = e^{i (ny y_j + nz z_j)}, A = ny my_j + nz mz_j, P = W A and Q = W mx_j:
//   S(+a) = (aP - bP) + a (aQ - bQ) + i [(cP + dP) + a (cQ + dQ)],
//   S(-a) = (aP + bP) - a (aQ + bQ) + i [(cP - dP) - a (cQ - dQ)],
// aP = sum c P.re, bP = sum s P.im, cP = sum c P.im, dP = sum s P.re (Q
// alike), which is the first pass of CompDipInteractEwaldF
// (src/interaction.cpp:459-483) with e^{i k.r} = X W. The four real sums are
// products (a, c|s) x particles times particles x rows, taken with
// v_mfma_f32_16x16x4f32: M = 8 values of a times (c, s), N = 16 rows, K = 4
// particles; four products (P.re, P.im, Q.re, Q.im) per M tile. In that
// instruction lane l gives A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// D[4 (l / 16) + r][l % 16], r = 0..3, so a lane ends with all eight sums of
// two values of a and one row. Accumulation is FP32 within a particle chunk;
// the chunks are added in FP64 by k_kpack.
// Block: four wavefronts on four groups of 16 rows, one particle chunk and
// one pass of KBM_NX values of a; the chunk's FP32 table rows are copied to
// LDS, KBM_SUB particles at a time (n_c <= KBM_NCMAX: the LDS).
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int KBM_SUB = 64, KBM_WAVES = 4;

__global__ __launch_bounds__(256) void k_kbuild_mfma(int64_t n_all, const Rep* __restrict__ reps, int32_t nrep,
                                                     int32_t items, const KRow* __restrict__ rows_all,
                                                     const float* __restrict__ tab32_all,
                                                     const double* __restrict__ mom,
                                                     double* __restrict__ partial_all) {
    // dynamic LDS: [KBM_SUB][2 KBM_NX] x phases, [KBM_SUB][4] moments,
    // [KBM_SUB][2 tw4] y and z tables (tw4 of this replica's n_c)
    extern __shared__ float kbm_smem[];
    float(*sx)[2 * KBM_NX] = (float(*)[2 * KBM_NX])kbm_smem;
    float(*sm)[4] = (float(*)[4])(kbm_smem + KBM_SUB * 2 * KBM_NX);
    float* syz = kbm_smem + KBM_SUB * (2 * KBM_NX + 4);
    // a grid smaller than the items walks them in strides (the grid is a
    // multiple of eight, so an item keeps its XCD); every skip below is
    // uniform over the block, and the staging loop's first barrier keeps one
    // item's LDS from the next
  for (int32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const Rep& R = reps[find_rep<3>(reps, nrep, item)];
    if (!(R.kmode & 1)) continue;  // this replica is built by k_kbuild
    // XCD-aware order: blocks are dealt to the eight XCDs by index mod 8, so
    // all row blocks of chunk ch get indices = ch (mod 8) and the chunk's
    // tables are fetched into one XCD's L2 (chunks in groups of eight; the
    // last group's missing chunks return at once)
    const int32_t li = item - R.kbm_off;
    const int32_t nblk = (R.nrowg + KBM_WAVES - 1) / KBM_WAVES, nb = nblk * R.kbm_pass;
    const int32_t grp = li / (8 * nb), rr = li % (8 * nb);
    const int32_t ch = 8 * grp + (rr & 7), b = rr >> 3;
    if (ch >= R.nchunks) continue;
    const int32_t blk = b % nblk, pass = b / nblk;
    const int wv = wave_id(), lane = lane_id(), tid = threadIdx.x;
    const int kk = lane >> 4, col = lane & 15;
    const int nc = R.nc, tw4 = tw4_of(nc);
    const int64_t stride = 3 * (int64_t)tw4;
    const float* tab = tab32_all + R.tab32_off;
    const int64_t n = n_all;
    const int a0 = pass * KBM_NX;
    // this lane's row (B side)
    const int r = (blk * KBM_WAVES + wv) * KBM_ROWS + col;
    const bool vrow = r < R.nrows;
    const KRow row = vrow ? rows_all[R.row_off + r] : KRow{0, 0, -1, 0, 0};
    const int yo = 2 * (row.ny < 0 ? -row.ny : row.ny), zo = tw4 + 2 * row.nz;
    const float ysg = row.ny < 0 ? -1.0f : 1.0f;
    const float fny = (float)row.ny, fnz = (float)row.nz;
    // M tiles (8 values of a each) this wavefront needs: its rows come by
    // decreasing |nx| range, so short rows share waves that skip the upper
    // tiles
    int amax = row.nxmax;
    for (int o = 32; o > 0; o >>= 1) amax = max(amax, __shfl_xor(amax, o));
    amax = __builtin_amdgcn_readfirstlane(amax);
    const int ntile = amax < a0 ? 0 : min(3, (amax - a0) / 8 + 1);
    f32x4 acc[3][4];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[t][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t jb = (int64_t)ch * R.kchunk;
    const int64_t je = jb + R.kchunk < R.n ? jb + R.kchunk : R.n;
    for (int64_t s0 = jb; s0 < je; s0 += KBM_SUB) {
        // stage KBM_SUB particles, four floats at a time: X for a in
        // [a0, a0 + KBM_NX), the y and z tables, the moment; zeros past the
        // chunk's end and past n_c (the tables' own padding up to tw4)
        // Every load reads a valid address (out-of-range slots read the
        // chunk's first particle and are replaced by zeros), so a group of
        // them is issued before the first wait instead of one round trip each.
        __syncthreads();
        {
            constexpr int NXQ = 2 * KBM_NX / 4, TOT = KBM_SUB * NXQ;  // 3 per thread
            static_assert(TOT % 256 == 0, "whole rounds of the block");
            float4 v[TOT / 256];
            bool ok[TOT / 256];
#pragma unroll
            for (int u = 0; u < TOT / 256; ++u) {
                const int e = tid + 256 * u, p = e / NXQ, q = e % NXQ;
                const int64_t jr = s0 + p;
                const int o = 2 * a0 + 4 * q;
                ok[u] = jr < je && o < tw4;
                v[u] = *(const float4*)(tab + (ok[u] ? jr : jb) * stride + (ok[u] ? o : 0));
            }
#pragma unroll
            for (int u = 0; u < TOT / 256; ++u) {
                const int e = tid + 256 * u;
                ((float4*)sx[e / NXQ])[e % NXQ] = ok[u] ? v[u] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        {
            const int yq = tw4 / 2, tot = KBM_SUB * yq;
            for (int e0 = tid; e0 < tot; e0 += 4 * 256) {
                float4 v[4];
                bool ok[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = e0 + 256 * u, p = e / yq, q = e % yq;
                    const int64_t jr = s0 + p;
                    ok[u] = e < tot && jr < je;
                    v[u] = *(const float4*)(tab + (ok[u] ? jr : jb) * stride + tw4 + (ok[u] ? 4 * q : 0));
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (e0 + 256 * u < tot)
                        ((float4*)syz)[e0 + 256 * u] = ok[u] ? v[u] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        if (tid < KBM_SUB * 3) {
            const int p = tid / 3, d = tid % 3;
            const int64_t jr = s0 + p;
            sm[p][d] = jr < je ? (float)mom[d * n + R.off + jr] : 0.0f;
        }
        __syncthreads();
        const int steps = (int)((je - s0 < KBM_SUB ? je - s0 : KBM_SUB) + 3) / 4;
        for (int st = 0; st < steps; ++st) {
            const int p = 4 * st + kk;
            float bv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (vrow) {
                const float* yz = syz + p * 2 * tw4;
                const float yc = yz[yo], ys = ysg * yz[yo + 1];
                const float zc = yz[zo], zs = yz[zo + 1];
                const float wc = yc * zc - ys * zs, ws = ys * zc + yc * zs;
                const float mx = sm[p][0];
                const float A = fny * sm[p][1] + fnz * sm[p][2];
                bv[0] = wc * A;
                bv[1] = ws * A;
                bv[2] = wc * mx;
                bv[3] = ws * mx;
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t >= ntile) break;  // uniform
                const float xa = sx[p][16 * t + col];
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, bv[b], acc[t][b], 0, 0, 0);
            }
        }
    }
    if (!vrow) continue;
    double* out = partial_all + R.part_off + (int64_t)ch * R.nk * 2;
    const bool half = row.ny == 0 && row.nz == 0;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int qq = 0; qq < 2; ++qq) {
            const int a = a0 + 8 * t + 2 * kk + qq;
            if (a > row.nxmax || (half && a == 0)) continue;
            const double aP = acc[t][0][2 * qq], dP = acc[t][0][2 * qq + 1];
            const double cP = acc[t][1][2 * qq], bP = acc[t][1][2 * qq + 1];
            const double aQ = acc[t][2][2 * qq], dQ = acc[t][2][2 * qq + 1];
            const double cQ = acc[t][3][2 * qq], bQ = acc[t][3][2 * qq + 1];
            const double fa = (double)a;
            const int kp = half ? row.koff + a - 1 : row.koff + a + row.nxmax;
            out[2 * kp] = (aP - bP) + fa * (aQ - bQ);
            out[2 * kp + 1] = (cP + dP) + fa * (cQ + dQ);
            if (a >= 1 && !half) {
                const int km = row.koff - a + row.nxmax;
                out[2 * km] = (aP + bP) - fa * (aQ + bQ);
                out[2 * km + 1] = (cP - dP) - fa * (cQ - dQ);
            }
        }
  }
}

// Sum the S(k) partials of both k vectors of a packed entry over the
// particle chunks and store (U+ + U-, U+ - U-, V+ + V-, V+ - V-),
// U = e Re S, V = e Im S. A block takes 64 entries; its four wavefronts sum
// a quarter of the chunks each and the quarters are added in a fixed order.
__global__ __launch_bounds__(256) void k_kpack(int32_t npk_all, const Rep* __restrict__ reps, int32_t nrep,
                                               const KPack* __restrict__ desc, const double* __restrict__ partial_all,
                                               double4* __restrict__ pack, float4* __restrict__ pack32) {
    __shared__ double4 part[4][WAVE];
    const int wv = wave_id(), lane = lane_id();
    const int j = blockIdx.x * WAVE + lane;
    double cp = 0, sp = 0, cm = 0, sm = 0;
    KPack d{-1, -1, 0.0, -1, 0};
    int64_t d32 = -1;  // slot in the FP32 tile-major copy
    double im = 0.0;
    if (j < npk_all) {
        const Rep& R = reps[find_rep<2>(reps, nrep, j)];
        d = desc[j];
        if (R.kmode & 2) {
            d32 = R.pk32_off + d.t32;
            im = R.inv_mref;
        }
        const int nk = R.nk;
        const double* partial = partial_all + R.part_off;
        const int c0 = (R.nchunks * wv) / 4, c1 = (R.nchunks * (wv + 1)) / 4;
        // both k vectors' (re, im) as one 16-byte load each, unconditional
        // (an absent one reads slot 0 and is not added), four chunks in
        // flight; the chunks are still added in order
        const bool hp = d.kp >= 0, hm = d.km >= 0;
        const int64_t op = hp ? 2 * (int64_t)d.kp : 0, om = hm ? 2 * (int64_t)d.km : 0;
#pragma unroll 4
        for (int ch = c0; ch < c1; ++ch) {
            const double* q = partial + (int64_t)ch * nk * 2;
            const double2 vp = *(const double2*)(q + op);
            const double2 vm = *(const double2*)(q + om);
            if (hp) {
                cp += vp.x;
                sp += vp.y;
            }
            if (hm) {
                cm += vm.x;
                sm += vm.y;
            }
        }
    }
    part[wv][lane] = make_double4(cp, sp, cm, sm);
    __syncthreads();
    if (wv != 0 || j >= npk_all) return;
    const double4 a = part[0][lane], b = part[1][lane], c = part[2][lane], e = part[3][lane];
    cp = ((a.x + b.x) + c.x) + e.x;
    sp = ((a.y + b.y) + c.y) + e.y;
    cm = ((a.z + b.z) + c.z) + e.z;
    sm = ((a.w + b.w) + c.w) + e.w;
    const double up = d.e * cp, vp = d.e * sp, um = d.e * cm, vm = d.e * sm;
    const double4 v = make_double4(up + um, up - um, vp + vm, vp - vm);
    pack[j] = v;
    if (d32 >= 0) pack32[d32] = make_float4((float)(v.x * im), (float)(v.y * im), (float)(v.z * im), (float)(v.w * im));
}

// Rows [r0, r1) of the k-space scatter (k_kscatter) for particle i of
// replica R, into out[0..5] = (B0, B1, B2, F0, F1, F2) before constants. Per row
// (ny, nz) and |nx| = a the pair of k vectors (+-a, ny, nz) is folded into
// four products with the particle's phase: with G1 = SU wc + SV ws,
// G2 = DV wc - DU ws, G3 = DU wc + DV ws, G4 = SV wc - SU ws (S/D: sum and
// difference of the two vectors' e S(k), w = e^{i(ny y + nz z)}) and
// x = e^{i a x}:
//   sum over the pair of e Re(S conj(phase))        = xc G1 + xs G2
//   its nx-weighted difference (the nx component)   = xc G3 + xs G4
//   sum of e Im(S conj(phase)) (force), difference  = xs G3 - xc G4, xs G1 - xc G2
// The ny and nz components share the row's sums; constants are applied once.
template <bool FORCE>
__device__ __forceinline__ void kscatter_rows(int64_t n, int64_t i, const Rep& R, int32_t r0, int32_t r1,
                                              const KRow* __restrict__ rows_all, const double4* __restrict__ pack_all,
                                              const double* __restrict__ tab_all, const double* __restrict__ mom,
                                              double out[6]) {
    const int32_t nc = R.nc;
    const KRow* rows = rows_all + R.row_off;
    const double4* pack = pack_all + R.pk_off;
    const int64_t stride = 3 * (int64_t)(nc + 1) * 2;
    const double* t = tab_all + R.tab_off + (i - R.off) * stride;
    const double* ty = t + (nc + 1) * 2;
    const double* tz = t + 2 * (nc + 1) * 2;
    const double ex = t[2], sx = t[3], ey = ty[2], sy = ty[3];
    const double mx = mom[i], my = mom[n + i], mz = mom[2 * n + i];
    double F0 = 0, F1 = 0, F2 = 0, B0 = 0, B1 = 0, B2 = 0;
    double yc = 1, ys = 0, zc = 1, zs = 0;
    int cur_nz = -1;
    for (int r = r0; r < r1; ++r) {
        const KRow row = rows[r];
        if (row.nz != cur_nz) {
            cur_nz = row.nz;
            zc = tz[2 * row.nz];
            zs = tz[2 * row.nz + 1];
            const int ay = row.ny < 0 ? -row.ny : row.ny;
            yc = ty[2 * ay];
            ys = row.ny < 0 ? -ty[2 * ay + 1] : ty[2 * ay + 1];
        } else {
            const double nyc = yc * ey - ys * sy, nys = ys * ey + yc * sy;
            yc = nyc;
            ys = nys;
        }
        const double wc = yc * zc - ys * zs, ws = ys * zc + yc * zs;
        const double A = row.ny * my + row.nz * mz;
        const double4* P = pack + row.poff;
        double xc = 1.0, xs = 0.0, ad = 0.0, am = 0.0;
        double rb = 0.0, rs = 0.0;
        // two entries per iteration (rows are padded with zero entries to an
        // even length), loaded together so one wait covers both; the field
        // pass (fewer registers) takes two such pairs per wait
#pragma unroll(FORCE ? 1 : 2)
        for (int a = 0; a <= row.nxmax; a += 2) {
            const double4 p2[2] = {P[a], P[a + 1]};
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const double4 p = p2[u];
                const double G1 = p.x * wc + p.z * ws, G2 = p.w * wc - p.y * ws;
                const double G3 = p.y * wc + p.w * ws, G4 = p.z * wc - p.x * ws;
                rb += xc * G1 + xs * G2;
                B0 += ad * (xc * G3 + xs * G4);
                if (FORCE) {
                    const double ts = xs * G3 - xc * G4, td = xs * G1 - xc * G2;
                    rs += A * ts + am * td;
                    F0 += ad * (A * td + am * ts);
                    am += mx;
                }
                ad += 1.0;
                const double nxc = xc * ex - xs * sx, nxs = xs * ex + xc * sx;
                xc = nxc;
                xs = nxs;
            }
        }
        B1 += row.ny * rb;
        B2 += row.nz * rb;
        if (FORCE) {
            F1 += row.ny * rs;
            F2 += row.nz * rs;
        }
    }
    out[0] = B0;
    out[1] = B1;
    out[2] = B2;
    out[3] = F0;
    out[4] = F1;
    out[5] = F2;
}

// Second pass of CompDipInteractEwaldF / CompDipFieldEwaldF
// (src/interaction.cpp:485-504, 644-662) plus the self and vacuum-surface
// terms (510-513). A block holds 64 particles (one per lane) in four
// wavefronts; wavefront w sums the rows [ksplit[w], ksplit[w+1]) (quarters of
// the table by entries), so that 1e5 particles make ~6 k waves rather than
// 1.6 k; the four partial sums are added in a fixed order in LDS. Phases
// advance by the recurrences in y (row to row) and x (within a row). A
// wavefront whose particles share one replica (all but those straddling a
// replica boundary) reads the table through scalar loads.
template <bool FORCE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_kscatter(int64_t n, const Rep* __restrict__ reps,
                                                  const int32_t* __restrict__ rep_of, const KRow* __restrict__ rows_all,
                                                  const double4* __restrict__ pack_all,
                                                  const double* __restrict__ tab_all, const double* __restrict__ mom,
                                                  const double* __restrict__ msum_all,
                                                  const double* __restrict__ add_force, double* __restrict__ force,
                                                  double* __restrict__ field) {
    __shared__ double part[4][6][WAVE];
    const int wv = wave_id(), lane = lane_id();
    const int64_t i = blockIdx.x * (int64_t)WAVE + lane;
    const int64_t ic = i < n ? i : n - 1;
    const int32_t rid = rep_of[ic];
    const int32_t rid0 = __builtin_amdgcn_readfirstlane(rid);
    const bool uniform = __ballot(rid != rid0) == 0;
    // particles of replicas scattered on the matrix cores are left to k_kscatter_mfma
    const bool mine = !(reps[rid].kmode & 2);
    double o[6] = {0, 0, 0, 0, 0, 0};
    if (i < n && mine) {
        if (uniform) {
            const Rep& R = reps[rid0];
            const int32_t r0 = R.ksplit[wv], r1 = wv < 3 ? R.ksplit[wv + 1] : R.nrows;
            kscatter_rows<FORCE>(n, i, R, r0, r1, rows_all, pack_all, tab_all, mom, o);
        } else {
            const Rep& R = reps[rid];
            const int32_t r0 = R.ksplit[wv], r1 = wv < 3 ? R.ksplit[wv + 1] : R.nrows;
            kscatter_rows<FORCE>(n, i, R, r0, r1, rows_all, pack_all, tab_all, mom, o);
        }
    }
    for (int d = 0; d < 6; ++d) part[wv][d][lane] = o[d];
    __syncthreads();
    if (wv != 0 || i >= n || !mine) return;
    double a[6];
    for (int d = 0; d < 6; ++d) a[d] = ((part[0][d][lane] + part[1][d][lane]) + part[2][d][lane]) + part[3][d][lane];
    const Rep& R = reps[rid];
    const double* msum = msum_all + 3 * rid;
    const double mx = mom[i], my = mom[n + i], mz = mom[2 * n + i];
    const double cb = R.c_flux, cf = R.c_force, c_self = R.c_self, c_surf = R.c_surf;
    // written here; the real-space pair sums are added after (k_pairs_out)
    field[i] = -cb * a[0] + c_self * mx - c_surf * msum[0];
    field[n + i] = -cb * a[1] + c_self * my - c_surf * msum[1];
    field[2 * n + i] = -cb * a[2] + c_self * mz - c_surf * msum[2];
    if (FORCE) {
        force[i] = cf * a[3] + (add_force ? add_force[i] : 0.0);
        force[n + i] = cf * a[4] + (add_force ? add_force[n + i] : 0.0);
        force[2 * n + i] = cf * a[5] + (add_force ? add_force[2 * n + i] : 0.0);
    }
}

// The scatter from FP32 matrix-core products (mixed-precision path). With
// the packed entries (SU, DU, SV, DV) of a row, x = (c, s) = e^{i a x_i} and
// w = e^{i (ny y_i + nz z_i)}, the sums over |nx| = a of k_kscatter's
// per-entry terms are, per particle and row,
//   E1 = sum (c SU + s DV),     E2 = sum (c SV - s DU),
//   O1 = sum a (c DU + s SV),   O2 = sum a (c DV - s SU),
//   E1'', E2'': E1, E2 with weight a^2,
//   field: B0 += wc O1 + ws O2, (B1, B2) += (ny, nz) (wc E1 + ws E2),
//   force: with A = ny my + nz mz, rs = A (ws E1 - wc E2) + mx (ws O1 - wc O2),
//          F0 += A (ws O1 - wc O2) + mx (ws E1'' - wc E2''), (F1, F2) += (ny, nz) rs.
// The E and O sums are products particles x (a, c|s) times (a, c|s) x rows:
// v_mfma_f32_16x16x4f32 with M = 16 particles, N = 16 rows, K = two values
// of a times (c, s); a lane then holds them for four particles and one row
// and adds the row's terms in FP64. A block takes 64 particles of one
// replica (16 per wavefront) and walks the rows 16 at a time; the particles'
// FP32 phase tables and each row tile's FP32 entries (k_kpack's tile-major
// copy, in units of the replica's largest moment) are copied to LDS.
template <bool FORCE>
__global__ __launch_bounds__(256) void k_kscatter_mfma(int64_t n, const Rep* __restrict__ reps, int32_t nrep,
                                                       int32_t items, const KRow* __restrict__ rows_all,
                                                       const float4* __restrict__ pack32_all,
                                                       const float* __restrict__ tab32_all,
                                                       const double* __restrict__ mom,
                                                       const double* __restrict__ msum_all,
                                                       const double* __restrict__ add_force,
                                                       double* __restrict__ force, double* __restrict__ field) {
    extern __shared__ float4 smem4[];
    const int32_t item = blockIdx.x;
    if (item >= items) return;  // uniform over the block
    const int rid = find_rep<4>(reps, nrep, item);
    const Rep& R = reps[rid];
    if (!(R.kmode & 2)) return;  // uniform over the block: this replica is scattered by k_kscatter
    const int64_t pb = (int64_t)(item - R.ksm_off) * WAVE;  // replica-local first particle
    const int wv = wave_id(), lane = lane_id(), tid = threadIdx.x;
    const int kk = lane >> 4, col = lane & 15;
    const int nc = R.nc, tw4 = tw4_of(nc), xw = (2 * (nc + 1) + 2 + 3) & ~3, nent = nc + 2;
    const int64_t stride = 3 * (int64_t)tw4;
    const float* tab = tab32_all + R.tab32_off;
    const float4* pk32 = pack32_all + R.pk32_off;
    const double im = R.inv_mref;
    float* sx = (float*)smem4;                   // [64][xw] x phases (c, s) per a, zero padded to a + 1 = n_c + 2
    float* syz = sx + WAVE * xw;                 // [64][2 tw4] y, then z tables
    // (values, not a select between a table address and a local zero,
    // which would put the zero in scratch memory)
    for (int e = tid; e < WAVE * (xw / 4); e += 256) {
        const int p = e / (xw / 4), q = e % (xw / 4);
        const int64_t jr = pb + p;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (jr < R.n && 4 * q < tw4) v = *(const float4*)(tab + jr * stride + 4 * q);
        ((float4*)sx)[e] = v;
    }
    for (int e = tid; e < WAVE * (tw4 / 2); e += 256) {
        const int p = e / (tw4 / 2), q = e % (tw4 / 2);
        const int64_t jr = pb + p;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (jr < R.n) v = *(const float4*)(tab + jr * stride + tw4 + 4 * q);
        ((float4*)syz)[e] = v;
    }
    __syncthreads();  // the tables are staged
    // this lane's particles (rows of the products): 16 wv + 4 kk + r
    // (loads from a valid index for every lane, all issued before the first
    // wait; lanes past the replica's end take zeros)
    float mxs[4], mys[4], mzs[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t jr = pb + 16 * wv + 4 * kk + r;
        const int64_t j = R.off + (jr < R.n ? jr : 0);
        const double x = mom[j], y = mom[n + j], z = mom[2 * n + j];
        mxs[r] = jr < R.n ? (float)(x * im) : 0.0f;
        mys[r] = jr < R.n ? (float)(y * im) : 0.0f;
        mzs[r] = jr < R.n ? (float)(z * im) : 0.0f;
    }
    constexpr int NA = FORCE ? 6 : 3;
    double acc[4][NA];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int d = 0; d < NA; ++d) acc[r][d] = 0.0;
    const KRow* rows = rows_all + R.row_off;
    const float* xrow = sx + (16 * wv + col) * xw + kk;
    for (int r0 = 0; r0 < R.nrows; r0 += 16) {
        const int nt = R.nrows - r0 < 16 ? R.nrows - r0 : 16;
        // the rows come by decreasing |nx| range: the tile's first is its longest
        const int amax = rows[r0].nxmax;
        const int kn = (amax + 2) / 2;  // a in [0, 2 kn), within [0, n_c + 1]
        // the tile's entries (zero where a row has none), read through the
        // caches (the tiles of a replica take a few hundred KB)
        const float4* tile = pk32 + (int64_t)(r0 / 16) * nent * 16;
        f32x4 c[NA == 6 ? 6 : 4];
#pragma unroll
        for (int u = 0; u < (NA == 6 ? 6 : 4); ++u) c[u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const bool xs_ = kk & 1;
        // the next a's entry and phase are loaded before this a's products
        // (the entry comes from L2: one wait per a otherwise)
        float4 p_next = tile[(kk >> 1) * 16 + col];
        float x_next = xrow[0];
        for (int ks = 0; ks < kn; ++ks) {
            const int a = 2 * ks + (kk >> 1);
            const float xa = x_next;
            const float4 p = p_next;
            if (ks + 1 < kn) {  // uniform over the wavefront
                p_next = tile[(a + 2) * 16 + col];
                x_next = xrow[4 * (ks + 1)];
            }
            const float fa = (float)a;
            const float e1 = xs_ ? p.w : p.x;          // SU | DV
            const float e2 = xs_ ? -p.y : p.z;         // SV | -DU
            const float o1 = fa * (xs_ ? p.z : p.y);   // a (DU | SV)
            const float o2 = fa * (xs_ ? -p.x : p.w);  // a (DV | -SU)
            c[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, e1, c[0], 0, 0, 0);
            c[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, e2, c[1], 0, 0, 0);
            c[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, o1, c[2], 0, 0, 0);
            c[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, o2, c[3], 0, 0, 0);
            if constexpr (FORCE) {
                const float a2 = fa * fa;
                c[4] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, a2 * e1, c[4], 0, 0, 0);
                c[5] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, a2 * e2, c[5], 0, 0, 0);
            }
        }
        if (col < nt) {
            const KRow rw = rows[r0 + col];
            const int yo = 2 * (rw.ny < 0 ? -rw.ny : rw.ny), zo = tw4 + 2 * rw.nz;
            const float ysg = rw.ny < 0 ? -1.0f : 1.0f, fny = (float)rw.ny, fnz = (float)rw.nz;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float* yz = syz + (16 * wv + 4 * kk + r) * 2 * tw4;
                const float yc = yz[yo], ys = ysg * yz[yo + 1], zc = yz[zo], zs = yz[zo + 1];
                const float wc = yc * zc - ys * zs, ws = ys * zc + yc * zs;
                const float E1 = c[0][r], E2 = c[1][r], O1 = c[2][r], O2 = c[3][r];
                const float rb = wc * E1 + ws * E2;
                acc[r][0] += (double)(wc * O1 + ws * O2);
                acc[r][1] += (double)(fny * rb);
                acc[r][2] += (double)(fnz * rb);
                if constexpr (FORCE) {
                    const float A = fny * mys[r] + fnz * mzs[r];
                    const float td1 = ws * O1 - wc * O2;
                    const float rs = A * (ws * E1 - wc * E2) + mxs[r] * td1;
                    acc[r][3] += (double)(A * td1 + mxs[r] * (ws * c[4][r] - wc * c[5][r]));
                    acc[r][4] += (double)(fny * rs);
                    acc[r][5] += (double)(fnz * rs);
                }
            }
        }
    }
    // totals over the 16 row lanes of each group (the shuffles stay in it)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int d = 0; d < NA; ++d) {
            double v = acc[r][d];
            for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o);
            acc[r][d] = v;
        }
    if (col >= 4) return;
    const int64_t jr = pb + 16 * wv + 4 * kk + col;
    if (jr >= R.n) return;
    double a[NA];
#pragma unroll
    for (int d = 0; d < NA; ++d) a[d] = col == 0 ? acc[0][d] : col == 1 ? acc[1][d] : col == 2 ? acc[2][d] : acc[3][d];
    const int64_t i = R.off + jr;
    const double* msum = msum_all + 3 * rid;
    const double mref = 1.0 / im;
    const double mx = mom[i], my = mom[n + i], mz = mom[2 * n + i];
    const double cb = R.c_flux * mref, cf = R.c_force * mref * mref, c_self = R.c_self, c_surf = R.c_surf;
    // written here; the real-space pair sums are added after (k_pairs_out)
    field[i] = -cb * a[0] + c_self * mx - c_surf * msum[0];
    field[n + i] = -cb * a[1] + c_self * my - c_surf * msum[1];
    field[2 * n + i] = -cb * a[2] + c_self * mz - c_surf * msum[2];
    if constexpr (FORCE) {
        force[i] = cf * a[3] + (add_force ? add_force[i] : 0.0);
        force[n + i] = cf * a[4] + (add_force ? add_force[n + i] : 0.0);
        force[2 * n + i] = cf * a[5] + (add_force ? add_force[2 * n + i] : 0.0);
    }
}

// ------------------------------------------------------------------ integrator
struct Coef {
    double v[8];
};

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// One Runge-Kutta stage of IntegrationRot (src/simulation.cpp:707-868): the
// right-hand side RKCoeffs (876-902) at the stage point, then either the next
// stage point y + h sum_r a_r k_r (with its moments for the next field
// evaluation) or, at the last stage, the error norm |h sum_r e_r kMm_r| and
// the accepted, renormalised state y + h sum_r b_r k_r (746-747, 866-867).
__global__ __launch_bounds__(256) void k_stage(StageParams sp, int32_t stage, int32_t stages,
                                               const double* __restrict__ mm0, const double* __restrict__ ea0,
                                               const double* __restrict__ ymm, const double* __restrict__ yea,
                                               const double* __restrict__ field, const double* __restrict__ n_t,
                                               const double* __restrict__ n_b, double* __restrict__ kmm,
                                               double* __restrict__ kea, Coef a_next, Coef b_acc, Coef e_err,
                                               double* __restrict__ ymm_out, double* __restrict__ yea_out,
                                               double* __restrict__ mom, double4* __restrict__ rec_mom,
                                               double* __restrict__ err_partial, double* __restrict__ new_mm,
                                               double* __restrict__ new_ea) {
    const int64_t n = sp.n;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool v = i < n;
    double mo[3] = {0, 0, 0};
    double errn = 0;
    if (v) {
        // DefExtField (src/simulation.cpp:546-556) at the step's start time
        const int32_t ri = sp.rep_of[i];
        const double bext = (sp.rep_t[ri] < sp.t_mag) ? sp.reps[ri].field_z : 0.0;
        double m[3], e[3], B[3], T[3], Bt[3];
        for (int d = 0; d < 3; ++d) {
            m[d] = ymm[d * n + i];
            e[d] = yea[d * n + i];
            B[d] = field[d * n + i];
            T[d] = n_t[d * n + i];
            Bt[d] = n_b[d * n + i];
        }
        const double dt = m[0] * e[0] + m[1] * e[1] + m[2] * e[2];
        double Be[3];
        for (int d = 0; d < 3; ++d) Be[d] = (d == 2 ? bext : 0.0) + B[d] + sp.anis_const * (e[d] * dt) + Bt[d];
        double ke[3] = {0, 0, 0};
        if (sp.mobil) {
            const double c1 = sp.c1[i], c2 = sp.c2[i];
            double Te[3];
            cross3(T, e, Te);
            for (int d = 0; d < 3; ++d) ke[d] = (m[d] - e[d] * dt) * (c1 * dt) + Te[d] * c2;
        }
        double b2[3], b3[3], km[3];
        cross3(m, Be, b2);
        cross3(m, b2, b3);
        for (int d = 0; d < 3; ++d) km[d] = -sp.vel_mm_const * (b2[d] + sp.mag_damp * b3[d]);
        const int64_t so = (int64_t)stage * 3 * n;
        for (int d = 0; d < 3; ++d) {
            kmm[so + d * n + i] = km[d];
            kea[so + d * n + i] = ke[d];
        }
        if (stage + 1 < stages) {
            for (int d = 0; d < 3; ++d) {
                double sm = a_next.v[stage] * km[d], se = a_next.v[stage] * ke[d];
                for (int r = 0; r < stage; ++r) {
                    sm += a_next.v[r] * kmm[(int64_t)r * 3 * n + d * n + i];
                    se += a_next.v[r] * kea[(int64_t)r * 3 * n + d * n + i];
                }
                const double ym = mm0[d * n + i] + sp.h * sm;
                ymm_out[d * n + i] = ym;
                yea_out[d * n + i] = ea0[d * n + i] + sp.h * se;
                mo[d] = ym * sp.mag_mom[i];
                mom[d * n + i] = mo[d];
            }
            const int32_t k = sp.inv_perm[i];
            rec_mom[k] = make_double4(mo[0], mo[1], mo[2], 0.0);
            const double s = sp.reps[sp.rep_of[i]].inv_mref;
            sp.rec_mom32[k] = make_float4((float)(mo[0] * s), (float)(mo[1] * s), (float)(mo[2] * s), 0.0f);
        } else {
            double nm[3], ne[3], ev[3];
            for (int d = 0; d < 3; ++d) {
                double sm = b_acc.v[stage] * km[d], se = b_acc.v[stage] * ke[d], er = e_err.v[stage] * km[d];
                for (int r = 0; r < stage; ++r) {
                    const double kr = kmm[(int64_t)r * 3 * n + d * n + i];
                    sm += b_acc.v[r] * kr;
                    er += e_err.v[r] * kr;
                    se += b_acc.v[r] * kea[(int64_t)r * 3 * n + d * n + i];
                }
                nm[d] = mm0[d * n + i] + sp.h * sm;
                ne[d] = ea0[d * n + i] + sp.h * se;
                ev[d] = er;
            }
            errn = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
            const double lm = sqrt(nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2]);
            const double le = sqrt(ne[0] * ne[0] + ne[1] * ne[1] + ne[2] * ne[2]);
            for (int d = 0; d < 3; ++d) {
                new_mm[d * n + i] = nm[d] / lm;
                new_ea[d * n + i] = ne[d] / le;
            }
        }
    }
    (void)mo;
    if (stage + 1 == stages) {
        __shared__ double sh[256 / WAVE];
        errn = wave_max(errn);
        if (lane_id() == 0) sh[threadIdx.x / WAVE] = errn;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0;
            for (int k = 0; k < 256 / WAVE; ++k) m = fmax(m, sh[k]);
            err_partial[blockIdx.x] = m;
        }
    }
}

__global__ __launch_bounds__(256) void k_reduce_max(const double* __restrict__ partial, int32_t count, double scale,
                                                    double* __restrict__ out) {
    __shared__ double sh[256 / WAVE];
    double m = 0;
    for (int k = threadIdx.x; k < count; k += blockDim.x) m = fmax(m, partial[k]);
    m = wave_max(m);
    if (lane_id() == 0) sh[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int k = 0; k < 256 / WAVE; ++k) r = fmax(r, sh[k]);
        *out = r * scale;
    }
}

// IntegrationTrans (src/simulation.cpp:959-962) + ApplyBoundaryCondArr
// (610-620), and the accepted orientations into the state.
__global__ __launch_bounds__(256) void k_translate(int64_t n, const Rep* __restrict__ reps,
                                                   const int32_t* __restrict__ rep_of,
                                                   const uint64_t* __restrict__ rep_step, double h,
                                                   const double* __restrict__ pos_cur, const double* __restrict__ force,
                                                   const double* __restrict__ zeta_trans, double* __restrict__ pos,
                                                   const double* __restrict__ new_mm, const double* __restrict__ new_ea,
                                                   double* __restrict__ mm, double* __restrict__ ea) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t ri = rep_of[i];
    if (rep_step[ri] >= reps[ri].stop_at) return;  // a replica that has run its steps keeps its state
    const double L = reps[ri].L;
    const double zt = zeta_trans[i];
    for (int d = 0; d < 3; ++d) {
        double x = pos_cur[d * n + i] + h * (force[d * n + i] / zt);
        if (x > L / 2) x -= L;
        if (x < -L / 2) x += L;
        pos[d * n + i] = x;
        mm[d * n + i] = new_mm[d * n + i];
        ea[d * n + i] = new_ea[d * n + i];
    }
}

// ------------------------------------------------------------------ noise
// Philox4x32-10 keyed by the replica's seed, counter (particle within the
// replica, call, replica's step), Box-Muller normals (philox.h, shared with
// the host copy that the published known-answer vectors check).
__global__ __launch_bounds__(256) void k_noise(int64_t n, const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                                               const uint64_t* __restrict__ rep_step, double dt,
                                               const double* __restrict__ zeta_rot,
                                               const double* __restrict__ vol_mag, const double* __restrict__ zeta_trans,
                                               double kT, double alpha_g, double gyro, double sat_mag, int32_t en_t,
                                               int32_t en_b, int32_t en_f, double* __restrict__ n_t,
                                               double* __restrict__ n_b, double* __restrict__ n_f) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    // counter (index within the replica, draw, replica's step): a replica's
    // stream does not depend on the other replicas of the context
    const int32_t ri = rep_of[i];
    const Rep& R = reps[ri];
    const uint64_t step = rep_step[ri];
    const uint64_t seed = R.seed;
    const int64_t li = i - R.off;
    double z[10];
    philox_normals(seed, li, step, z);
    // DefThermFluct, src/simulation.cpp:565-602: torque, field, force
    const double sT = en_t ? sqrt((2.0 * kT / dt) * zeta_rot[i]) : 0.0;
    const double sB = en_b ? sqrt(2.0 * kT * alpha_g / (gyro * sat_mag * dt) * (1.0 / vol_mag[i])) : 0.0;
    const double sF = en_f ? sqrt((2.0 * kT / dt) * zeta_trans[i]) : 0.0;
    for (int d = 0; d < 3; ++d) {
        n_t[d * n + i] = z[d] * sT;
        n_b[d * n + i] = z[3 + d] * sB;
        n_f[d * n + i] = z[6 + d] * sF;
    }
}

// Advance every replica's step counter and time by one accepted step h.
// Transport observables of EvalTransCoeffs (src/postprocessing.cpp:167-187).
// Positions unwrapped against the previous true positions (Rapaport's
// nearest-image rule, rounding half away from zero as Eigen's round()) and
// the squared displacement from the reference positions per axis, for the
// chunked per-replica sums.
__global__ void k_unwrap_d2(int64_t n, const Rep* __restrict__ reps, const int32_t* __restrict__ rep_of,
                            const double* __restrict__ pos, double* __restrict__ true_pos,
                            const double* __restrict__ pos_ref, double* __restrict__ d2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double L = reps[rep_of[i]].L;
    for (int d = 0; d < 3; ++d) {
        const int64_t k = (int64_t)d * n + i;
        const double x = pos[k];
        const double tp = x + round((true_pos[k] - x) / L) * L;
        true_pos[k] = tp;
        const double e = tp - pos_ref[k];
        d2[k] = e * e;
    }
}

// Per replica (one block each): t before the step, MSD per axis (its chunk
// sums over the particle count), and Pxy, Pxz, Pyz of the step's force pass
// (its work items' slots in order, then the moved-pair slot) over the box
// volume, into obs[r] and, at the running step index, the series.
__global__ __launch_bounds__(256) void k_rep_obs(const Rep* __restrict__ reps, int32_t nrep,
                                                 const double* __restrict__ part, const double* __restrict__ rep_t,
                                                 const int32_t* __restrict__ work, const int32_t* __restrict__ nwork,
                                                 const int32_t* __restrict__ cell_rep,
                                                 const double* __restrict__ press_partial, int64_t moved_press_off,
                                                 double* __restrict__ obs, const int32_t* __restrict__ counter,
                                                 double* __restrict__ series, int32_t series_cap) {
    const int r = blockIdx.x;
    const Rep& R = reps[r];
    __shared__ int32_t range[2];
    if (threadIdx.x < 2) {
        // first work item of replica r (threadIdx 0) or of r + 1 (threadIdx 1):
        // work items are in cell order and cells in replica order
        const int target = r + (int)threadIdx.x;
        int lo = 0, hi = *nwork;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cell_rep[work[mid] >> 8] < target) lo = mid + 1; else hi = mid;
        }
        range[threadIdx.x] = lo;
    }
    __syncthreads();
    const int64_t s0 = (int64_t)range[0] * PAIR_SLICES, s1 = (int64_t)range[1] * PAIR_SLICES;
    double a = 0, b = 0, c = 0;
    for (int64_t k = s0 + threadIdx.x; k < s1; k += blockDim.x) {
        a += press_partial[3 * k];
        b += press_partial[3 * k + 1];
        c += press_partial[3 * k + 2];
    }
    __shared__ double p3[3];
    block_sum3<256>(a, b, c, p3);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long* pw =
            reinterpret_cast<const unsigned long long*>(press_partial + moved_press_off) + 6 * (int64_t)r;
        const double pm[3] = {fx_value(pw[0], pw[1], R.px_s), fx_value(pw[2], pw[3], R.px_s),
                              fx_value(pw[4], pw[5], R.px_s)};
        const double vol = R.L * R.L * R.L;
        double m[3] = {0, 0, 0};
        for (int k = R.red_off; k < R.red_off + R.nred; ++k)
            for (int d = 0; d < 3; ++d) m[d] += part[3 * (int64_t)k + d];
        const double v[7] = {rep_t[r], m[0] / (double)R.n, m[1] / (double)R.n, m[2] / (double)R.n,
                             (p3[0] + pm[0]) / vol, (p3[1] + pm[1]) / vol, (p3[2] + pm[2]) / vol};
        const int32_t k = counter ? *counter : series_cap;
        for (int f = 0; f < 7; ++f) {
            obs[7 * r + f] = v[f];
            if (series && k < series_cap) series[((int64_t)k * nrep + r) * 7 + f] = v[f];
        }
    }
}

__global__ void k_step_count(int32_t nrep, const Rep* __restrict__ reps, uint64_t* __restrict__ rep_step,
                             double* __restrict__ rep_t, double h, int32_t* __restrict__ series_counter) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep && rep_step[r] < reps[r].stop_at) {
        rep_step[r] += 1;
        rep_t[r] += h;
    }
    if (r == 0 && series_counter) *series_counter += 1;
}

__global__ void k_clear_events(int64_t n, int32_t* __restrict__ head, int32_t* __restrict__ tail,
                               int32_t* __restrict__ n_events, int32_t* __restrict__ n_moved,
                               double* __restrict__ max_disp) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        head[i] = -1;
        tail[i] = -1;
    }
    if (i == 0) {
        *n_events = 0;
        *n_moved = 0;
        *max_disp = 0.0;
    }
}

}  // namespace

// ------------------------------------------------------------------ launchers
#define LAUNCH_CHECK() return hipGetLastError()

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t count, int32_t v, const int32_t* skip) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, p, count, v, skip);
    LAUNCH_CHECK();
}
// ------------------------------------------------------------ peak probe
// Vector FMA issue rate: eight independent chains per lane (issue, not
// latency, limits the rate), the sum stored so that nothing is dead code.
// Packed FP32 is disabled for every compile, so the FP32 probe measures the
// unpacked rate.
template <typename T>
__global__ __launch_bounds__(256) void k_peak_fma(int32_t iters, T a, T b, T* __restrict__ out) {
    T x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (T)(threadIdx.x + k);
    for (int32_t it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = fma(x[k], a, b);
    T sum = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += x[k];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

hipError_t launch_peak_fma(hipStream_t s, bool fp64, int32_t blocks, int32_t iters, void* out) {
    if (fp64)
        hipLaunchKernelGGL(k_peak_fma<double>, dim3(blocks), dim3(256), 0, s, iters, 0.999999, 1e-6, (double*)out);
    else
        hipLaunchKernelGGL(k_peak_fma<float>, dim3(blocks), dim3(256), 0, s, iters, 0.999999f, 1e-6f, (float*)out);
    LAUNCH_CHECK();
}

hipError_t launch_fill_f64(hipStream_t s, double* p, int64_t count, double v) {
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, p, count, v);
    LAUNCH_CHECK();
}
hipError_t launch_scale(hipStream_t s, double* p, int64_t count, double f) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, p, count, f);
    LAUNCH_CHECK();
}

hipError_t launch_prepare(hipStream_t s, int64_t n, const double* pos, double* pos0, double* pos_cur, const Rep* reps,
                          const int32_t* rep_of, int32_t* key_main, int32_t* key_fine, const double* pos_build,
                          const int32_t* list_state, unsigned long long* list_maxd2) {
    hipLaunchKernelGGL(k_prepare, dim3(grid_for(n, 256)), dim3(256), 0, s, n, pos, pos0, pos_cur, reps, rep_of,
                       key_main, key_fine, pos_build, list_state, list_maxd2);
    LAUNCH_CHECK();
}

hipError_t launch_list_decide(hipStream_t s, int32_t nrep, const Rep* reps, int32_t* list_state,
                              unsigned long long* list_maxd2, int32_t* list_mode, unsigned long long* list_builds,
                              int32_t* sort_skip) {
    if (nrep <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_list_decide, dim3(1), dim3(256), 0, s, nrep, reps, list_state, list_maxd2, list_mode,
                       list_builds, sort_skip);
    LAUNCH_CHECK();
}

hipError_t launch_list_keys(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of,
                            const int32_t* list_mode, const double* pos0, double* pos_build, int32_t* key_main,
                            const int32_t* skip) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_list_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, reps, rep_of, list_mode, pos0,
                       pos_build, key_main, skip);
    LAUNCH_CHECK();
}

hipError_t launch_list_build(hipStream_t s, const Geometry& g, int32_t nrep, int32_t max_work, const int4* stencil,
                             int32_t* lcount, int32_t* loff, int32_t* list, int32_t* list_mode,
                             int32_t* list_state, long long* list_used, int32_t* err) {
    if (max_work <= 0 || nrep <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_list_pass<false>, dim3(max_work), dim3(WAVE * PAIR_SLICES), 0, s, g, stencil, lcount,
                       (const int32_t*)loff, list);
    hipLaunchKernelGGL(k_list_scan, dim3(nrep), dim3(1024), 0, s, g.reps, list_mode, list_state,
                       (const int32_t*)lcount, loff, list_used, err);
    hipLaunchKernelGGL(k_list_pass<true>, dim3(max_work), dim3(WAVE * PAIR_SLICES), 0, s, g, stencil, lcount,
                       (const int32_t*)loff, list);
    LAUNCH_CHECK();
}

int32_t scan_tiles(int64_t m) { return (int32_t)((m + SCAN_TILE - 1) / SCAN_TILE); }

// tmp: scan_tiles(m) ints
hipError_t launch_scan(hipStream_t s, const int32_t* count, int32_t m, int32_t* start, int32_t* end, int32_t* cursor,
                       int32_t* total, int32_t* tmp, const int32_t* skip = nullptr) {
    const int32_t nt = std::max(1, scan_tiles(m));
    hipLaunchKernelGGL(k_scan_sums, dim3(nt), dim3(256), 0, s, count, m, tmp, skip);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tmp, nt, total, skip);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, count, m, (const int32_t*)tmp, start, end, cursor,
                       skip);
    LAUNCH_CHECK();
}

hipError_t launch_cell_sort(hipStream_t s, int64_t n, const int32_t* key, int32_t ncells, int32_t* count,
                            int32_t* start, int32_t* end, int32_t* cursor, int32_t* sorted_idx, int32_t* scan_tmp,
                            const int32_t* skip) {
    hipError_t e = launch_fill_i32(s, count, ncells, 0, skip);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_count, dim3(grid_for(n, 256)), dim3(256), 0, s, n, key, count, skip);
    if ((e = launch_scan(s, count, ncells, start, end, cursor, nullptr, scan_tmp, skip)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_place, dim3(grid_for(n, 256)), dim3(256), 0, s, n, key, cursor, sorted_idx, skip);
    LAUNCH_CHECK();
}

hipError_t launch_build_records(hipStream_t s, int64_t n, const int32_t* sorted_idx, const double* pos0,
                                double4* rec_pos, float4* rec_pos32, int32_t* inv_perm, const Rep* reps,
                                const int32_t* rep_of, const double* pos_build, const int32_t* list_mode,
                                float4* rec_build32, int32_t* lcount) {
    hipLaunchKernelGGL(k_records, dim3(grid_for(n, 256)), dim3(256), 0, s, n, sorted_idx, pos0, rec_pos, rec_pos32,
                       inv_perm, reps, rep_of, pos_build, list_mode, rec_build32, lcount);
    LAUNCH_CHECK();
}

hipError_t launch_mark_moved(hipStream_t s, const int32_t* n_moved, const int32_t* moved_list,
                             const int32_t* inv_perm, double4* rec_pos) {
    hipLaunchKernelGGL(k_mark_moved, dim3(64), dim3(256), 0, s, n_moved, moved_list, inv_perm, rec_pos);
    LAUNCH_CHECK();
}

hipError_t launch_build_work(hipStream_t s, int32_t ncells, const int32_t* start, const int32_t* end, int32_t* work,
                             int32_t* nwork, int32_t cap, int32_t* err, const int32_t* skip) {
    hipLaunchKernelGGL(k_work, dim3(1), dim3(1024), 0, s, ncells, start, end, work, nwork, cap, err, skip);
    LAUNCH_CHECK();
}

hipError_t launch_contacts(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const double* pos0,
                           const double* r_mag, const int32_t* fstart, const int32_t* fend, const int32_t* fsorted,
                           double4* frec, int32_t* cand_count, int32_t* cand_off, int32_t* cand_j, int32_t cand_cap,
                           int32_t* cand_base, int32_t* cand_tmp, int32_t* err, int32_t* scan_tmp) {
    hipLaunchKernelGGL(k_frec, dim3(grid_for(n, 256)), dim3(256), 0, s, n, fsorted, pos0, r_mag, frec, cand_base);
    // cand_off has n+1 slots: offsets then the total at [n]; cand_j holds
    // 2*cand_cap ints: partner j then owner i.
    int32_t* cand_i = cand_j + cand_cap;
    const unsigned blocks = (unsigned)((n + 7) / 8);  // 8 particles per block
    hipLaunchKernelGGL(k_contacts, dim3(blocks), dim3(256), 0, s, n, reps, rep_of, pos0, r_mag, fstart, fend, fsorted,
                       (const double4*)frec, cand_count, cand_base, cand_tmp, cand_cap, err);
    hipError_t e = launch_scan(s, cand_count, (int32_t)n, cand_off, nullptr, nullptr, cand_off + n, scan_tmp);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cand_compact, dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const int32_t*)cand_count,
                       (const int32_t*)cand_off, (const int32_t*)cand_base, (const int32_t*)cand_tmp, cand_i, cand_j,
                       cand_cap, err);
    LAUNCH_CHECK();
}

hipError_t launch_resolve(hipStream_t s, int32_t nrep, int64_t n, const Rep* reps, const double* r_mag,
                          const int32_t* cand_off, const int32_t* cand_j, double* pos_cur, const double* pos0,
                          int32_t* ev_head, int32_t* ev_tail, int32_t* ev_next, uint64_t* ev_key, double* ev_pos,
                          int32_t ev_cap, int32_t* n_events, double* max_disp, double disp_limit, int32_t* moved_list,
                          int32_t* n_moved, int32_t* err) {
    // candidate offsets per particle (total at cand_off[n]); cand_j holds
    // partners then owners, each of capacity ev_cap / 2
    const int32_t cand_cap = ev_cap / 2;
    hipLaunchKernelGGL(k_resolve, dim3((nrep + 63) / 64), dim3(64), 0, s, nrep, n, reps, r_mag, cand_off,
                       cand_j + cand_cap, cand_j, pos_cur, pos0, ev_head, ev_tail, ev_next, ev_key, ev_pos, ev_cap,
                       n_events, max_disp, disp_limit, moved_list, n_moved, err);
    LAUNCH_CHECK();
}

hipError_t launch_clear_events(hipStream_t s, int64_t n, int32_t* ev_head, int32_t* ev_tail, int32_t* n_events,
                               int32_t* n_moved, double* max_disp) {
    hipLaunchKernelGGL(k_clear_events, dim3(grid_for(n, 256)), dim3(256), 0, s, n, ev_head, ev_tail, n_events,
                       n_moved, max_disp);
    LAUNCH_CHECK();
}

hipError_t launch_moments(hipStream_t s, int64_t n, const double* mm, const double* mag_mom, const int32_t* inv_perm,
                          const Rep* reps, const int32_t* rep_of, double* mom, double4* rec_mom, float4* rec_mom32) {
    hipLaunchKernelGGL(k_moments, dim3(grid_for(n, 256)), dim3(256), 0, s, n, mm, mag_mom, inv_perm, reps, rep_of, mom,
                       rec_mom, rec_mom32);
    LAUNCH_CHECK();
}

hipError_t launch_rep_sum3(hipStream_t s, const Rep* reps, int32_t nrep, const int2* chunks, int32_t nch, int64_t n,
                           const double* v, double* part, double* out) {
    hipLaunchKernelGGL(k_red_part, dim3(nch), dim3(256), 0, s, chunks, reps, n, v, 0, 3, part);
    hipLaunchKernelGGL(k_rep_sum3, dim3((nrep + 63) / 64), dim3(64), 0, s, reps, nrep, (const double*)part, out);
    LAUNCH_CHECK();
}

hipError_t launch_rep_mean_z(hipStream_t s, const Rep* reps, int32_t nrep, const int2* chunks, int32_t nch, int64_t n,
                             const double* v, double* part, double* out, const int32_t* counter, double* series,
                             int32_t series_cap) {
    hipLaunchKernelGGL(k_red_part, dim3(nch), dim3(256), 0, s, chunks, reps, n, v, 2, 1, part);
    hipLaunchKernelGGL(k_rep_mean_z, dim3((nrep + 63) / 64), dim3(64), 0, s, reps, nrep, (const double*)part, out,
                       counter, series, series_cap);
    LAUNCH_CHECK();
}

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[64];  // device address of `scratch`, per device
}  // namespace

hipError_t launch_mean_z_noalloc(hipStream_t s, int64_t n, const double* mm, double* out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            g_scratch_ptr[dev] = (double*)p;
        }
        part = g_scratch_ptr[dev];
    }
    const int64_t nch = (n + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_red_part_z, dim3((uint32_t)nch), dim3(256), 0, s, n, mm, part);
    hipLaunchKernelGGL(k_mean_z_final, dim3(1), dim3(64), 0, s, nch, n, (const double*)part, out);
    LAUNCH_CHECK();
}

hipError_t launch_observables(hipStream_t s, const Rep* reps, int32_t nrep, const int32_t* rep_of,
                              const int2* chunks, int32_t nch, int64_t n, const double* pos, double* true_pos,
                              const double* pos_ref, double* d2, double* part, const double* rep_t,
                              const int32_t* work, const int32_t* nwork, const int32_t* cell_rep,
                              const double* press_partial, int64_t moved_press_off, double* obs,
                              const int32_t* counter, double* series, int32_t series_cap) {
    hipLaunchKernelGGL(k_unwrap_d2, dim3(grid_for(n, 256)), dim3(256), 0, s, n, reps, rep_of, pos, true_pos, pos_ref,
                       d2);
    hipLaunchKernelGGL(k_red_part, dim3(nch), dim3(256), 0, s, chunks, reps, n, (const double*)d2, 0, 3, part);
    hipLaunchKernelGGL(k_rep_obs, dim3(nrep), dim3(256), 0, s, reps, nrep, (const double*)part, rep_t, work, nwork,
                       cell_rep, press_partial, moved_press_off, obs, counter, series, series_cap);
    LAUNCH_CHECK();
}

hipError_t launch_pairs(hipStream_t s, bool force, int32_t precision, const Geometry& g, const double4* rec_mom,
                        const PairParams& pp, const PairOut& out, int32_t max_work, const int4* stencil,
                        int32_t moved_press_off, int64_t* cnt_pairs) {
    const int64_t n = g.n;
    if (max_work <= 0 || n <= 0) return hipSuccess;
    // the pass's words of the fixed-point sums start at zero (field words are
    // 6n..12n, force words 0..6n)
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(6 * n, 256)), dim3(256), 0, s, (double*)out.acc + (force ? 0 : 6 * n),
                       (force ? 12 : 6) * n, 0.0);
    const int blocks = max_work;
    unsigned long long* cp = (unsigned long long*)cnt_pairs;
    const bool mixed = precision == 1;
#define KP(F, M) \
    hipLaunchKernelGGL((k_pairs<F, M>), dim3(blocks), dim3(WAVE * PAIR_SLICES), 0, s, g, rec_mom, pp, out, stencil, cp)
#define KC(M) \
    hipLaunchKernelGGL((k_pairs_coat<M>), dim3(blocks), dim3(WAVE * PAIR_SLICES), 0, s, g, rec_mom, pp, out, stencil, cp)
    if (force && pp.coating) {
        if (mixed) KC(true); else KC(false);
    } else if (force) {
        if (mixed) KP(true, true); else KP(true, false);
    } else {
        if (mixed) KP(false, true); else KP(false, false);
    }
#undef KP
#undef KC
    // pairs with an endpoint moved by the overlap correction (FP64)
    const int mblocks = 512;
    if (force && pp.coating)
        hipLaunchKernelGGL((k_pairs_moved<true, true>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    else if (force)
        hipLaunchKernelGGL((k_pairs_moved<true, false>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    else
        hipLaunchKernelGGL((k_pairs_moved<false, false>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    LAUNCH_CHECK();
}

hipError_t launch_pairs_list(hipStream_t s, bool force, int32_t precision, const Geometry& g, const double4* rec_mom,
                             const PairParams& pp, const PairOut& out, int32_t max_work, const int4* stencil,
                             int32_t moved_press_off, int64_t* cnt_pairs) {
    const int64_t n = g.n;
    if (max_work <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(6 * n, 256)), dim3(256), 0, s, (double*)out.acc + (force ? 0 : 6 * n),
                       (force ? 12 : 6) * n, 0.0);
    unsigned long long* cp = (unsigned long long*)cnt_pairs;
    const bool mixed = precision == 1;
    const bool coat = force && pp.coating;
#define KL(F, M) \
    hipLaunchKernelGGL((k_pairs_list<F, M>), dim3(max_work), dim3(WAVE * PAIR_SLICES), 0, s, g, rec_mom, pp, out, cp)
    if (coat) {
        // the short-range terms keep the searching kernel (valid between
        // builds: the grid covers the list radius)
        if (mixed)
            hipLaunchKernelGGL((k_pairs_coat<true>), dim3(max_work), dim3(WAVE * PAIR_SLICES), 0, s, g, rec_mom, pp,
                               out, stencil, cp);
        else
            hipLaunchKernelGGL((k_pairs_coat<false>), dim3(max_work), dim3(WAVE * PAIR_SLICES), 0, s, g, rec_mom, pp,
                               out, stencil, cp);
    } else if (force) {
        if (mixed) KL(true, true); else KL(true, false);
    } else {
        if (mixed) KL(false, true); else KL(false, false);
    }
#undef KL
    const int mblocks = 512;
    if (coat)
        hipLaunchKernelGGL((k_pairs_moved<true, true>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    else if (force)
        hipLaunchKernelGGL((k_pairs_moved<true, false>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    else
        hipLaunchKernelGGL((k_pairs_moved<false, false>), dim3(mblocks), dim3(64), 0, s, g, rec_mom, pp, out, stencil,
                           moved_press_off, cp);
    LAUNCH_CHECK();
}

hipError_t launch_pairs_out(hipStream_t s, const Geometry& g, const PairOut& out) {
    if (g.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pairs_out, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, g.sorted_idx, g.rep_of, g.reps,
                       out.acc, out.force, out.field);
    LAUNCH_CHECK();
}

hipError_t launch_cell_order(hipStream_t s, int64_t n, int32_t ncells, const int32_t* key, const int32_t* start,
                             const int32_t* end, const int32_t* placed, int32_t* sorted_idx, const int32_t* skip) {
    if (ncells <= 0 || n <= 0) return hipSuccess;
    const int64_t cb = (ncells + ORDER_SPLIT - 1) / ORDER_SPLIT, pb = (n + WAVE - 1) / WAVE;
    const int32_t cell_blocks = (int32_t)(cb < 4096 ? cb : 4096), chunk_blocks = (int32_t)(pb < 8192 ? pb : 8192);
    hipLaunchKernelGGL(k_cell_order, dim3(cell_blocks + chunk_blocks), dim3(WAVE * ORDER_SPLIT), 0, s, n, ncells,
                       cell_blocks, key, start, end, placed, sorted_idx, skip);
    LAUNCH_CHECK();
}

hipError_t launch_sincos_tables(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const double* pos0,
                                double* tab, float* tab32) {
    hipLaunchKernelGGL(k_sincos_tab, dim3(grid_for(n, 256)), dim3(256), 0, s, n, reps, rep_of, pos0, tab, tab32);
    LAUNCH_CHECK();
}

// Grid of a build kernel whose blocks stride over `items`: all of them, or
// at most `cap` blocks (a multiple of eight, so item and block keep their XCD).
static int build_grid(int32_t items, int32_t cap) {
    if (cap <= 0 || items <= cap) return items;
    return cap < 8 ? 8 : cap & ~7;
}

hipError_t launch_kbuild(hipStream_t s, int64_t n, const Rep* reps, int32_t nrep, int32_t kb_items, const KSeg* segs,
                         const double* tab, const double* mom, double* partial, int32_t max_blocks) {
    if (kb_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kbuild<KSEG>, dim3(build_grid(kb_items, max_blocks)), dim3(WAVE), 0, s, n, reps, nrep,
                       kb_items, segs, tab, mom, partial);
    LAUNCH_CHECK();
}

hipError_t launch_kbuild_mfma(hipStream_t s, int64_t n, const Rep* reps, int32_t nrep, int32_t items,
                              int32_t nc_max, const KRow* rows, const float* tab32, const double* mom,
                              double* partial, int32_t max_blocks) {
    if (items <= 0) return hipSuccess;
    if (nc_max > KBM_NCMAX) return hipErrorInvalidValue;
    const size_t lds = (size_t)KBM_SUB * (2 * KBM_NX + 4 + 2 * tw4_of(nc_max)) * sizeof(float);
    hipLaunchKernelGGL(k_kbuild_mfma, dim3(build_grid(items, max_blocks)), dim3(256), lds, s, n, reps, nrep, items,
                       rows, tab32, mom, partial);
    LAUNCH_CHECK();
}

hipError_t launch_kscatter_mfma(hipStream_t s, bool force, int64_t n, const Rep* reps, int32_t nrep, int32_t items,
                                int32_t nc_max, const KRow* rows, const float4* pack32, const float* tab32,
                                const double* mom, const double* msum, const double* add_force, double* force_out,
                                double* field) {
    if (items <= 0) return hipSuccess;
    if (nc_max > KSM_NCMAX) return hipErrorInvalidValue;
    const int xw = (2 * (nc_max + 1) + 2 + 3) & ~3;
    const size_t lds = (size_t)WAVE * xw * sizeof(float) + (size_t)WAVE * 2 * tw4_of(nc_max) * sizeof(float);
    if (force)
        hipLaunchKernelGGL(k_kscatter_mfma<true>, dim3(items), dim3(256), lds, s, n, reps, nrep, items, rows, pack32,
                           tab32, mom, msum, add_force, force_out, field);
    else
        hipLaunchKernelGGL(k_kscatter_mfma<false>, dim3(items), dim3(256), lds, s, n, reps, nrep, items, rows, pack32,
                           tab32, mom, msum, add_force, force_out, field);
    LAUNCH_CHECK();
}

hipError_t launch_kpack(hipStream_t s, int32_t npk_all, const Rep* reps, int32_t nrep, const KPack* desc,
                        const double* partial, double4* pack, float4* pack32) {
    if (npk_all <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kpack, dim3((npk_all + WAVE - 1) / WAVE), dim3(256), 0, s, npk_all, reps, nrep, desc,
                       partial, pack, pack32);
    LAUNCH_CHECK();
}

hipError_t launch_kscatter(hipStream_t s, bool force, int64_t n, const Rep* reps, const int32_t* rep_of,
                           const KRow* rows, const double4* pack, const double* tab, const double* mom,
                           const double* msum, const double* add_force, double* force_out, double* field) {
    // one block per 64 particles (four wavefronts over quarters of the rows)
    const int blocks = (int)((n + WAVE - 1) / WAVE);
    if (force)
        hipLaunchKernelGGL(k_kscatter<true>, dim3(blocks), dim3(256), 0, s, n, reps, rep_of, rows, pack, tab, mom, msum,
                           add_force, force_out, field);
    else
        hipLaunchKernelGGL(k_kscatter<false>, dim3(blocks), dim3(256), 0, s, n, reps, rep_of, rows, pack, tab, mom,
                           msum, add_force, force_out, field);
    LAUNCH_CHECK();
}

hipError_t launch_stage(hipStream_t s, const StageParams& sp, int32_t stage, int32_t stages, const double* mm0,
                        const double* ea0, const double* ymm, const double* yea, const double* field,
                        const double* n_t, const double* n_b, double* kmm, double* kea, const double* a_next,
                        const double* b_acc, const double* e_err, double* ymm_out, double* yea_out, double* mom,
                        double4* rec_mom, double* err_partial, double* new_mm, double* new_ea, int32_t nblk) {
    Coef a{}, b{}, e{};
    for (int k = 0; k < 8; ++k) {
        a.v[k] = a_next ? a_next[k] : 0.0;
        b.v[k] = b_acc ? b_acc[k] : 0.0;
        e.v[k] = e_err ? e_err[k] : 0.0;
    }
    hipLaunchKernelGGL(k_stage, dim3(nblk), dim3(256), 0, s, sp, stage, stages, mm0, ea0, ymm, yea, field, n_t, n_b,
                       kmm, kea, a, b, e, ymm_out, yea_out, mom, rec_mom, err_partial, new_mm, new_ea);
    LAUNCH_CHECK();
}

hipError_t launch_reduce_max(hipStream_t s, const double* partial, int32_t count, double scale, double* out) {
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, s, partial, count, scale, out);
    LAUNCH_CHECK();
}

hipError_t launch_translate(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const uint64_t* rep_step,
                            double h, const double* pos_cur, const double* force, const double* zeta_trans,
                            double* pos, const double* new_mm, const double* new_ea, double* mm, double* ea) {
    hipLaunchKernelGGL(k_translate, dim3(grid_for(n, 256)), dim3(256), 0, s, n, reps, rep_of, rep_step, h, pos_cur,
                       force, zeta_trans, pos, new_mm, new_ea, mm, ea);
    LAUNCH_CHECK();
}

hipError_t launch_noise(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const uint64_t* rep_step,
                        double dt, const double* zeta_rot, const double* vol_mag, const double* zeta_trans,
                        double kT, double alpha_g, double gyro, double sat_mag, int32_t en_t, int32_t en_b,
                        int32_t en_f, double* n_t, double* n_b, double* n_f) {
    hipLaunchKernelGGL(k_noise, dim3(grid_for(n, 256)), dim3(256), 0, s, n, reps, rep_of, rep_step, dt,
                       zeta_rot, vol_mag, zeta_trans, kT, alpha_g, gyro, sat_mag, en_t, en_b, en_f, n_t, n_b, n_f);
    LAUNCH_CHECK();
}

hipError_t launch_step_count(hipStream_t s, int32_t nrep, const Rep* reps, uint64_t* rep_step, double* rep_t,
                             double h, int32_t* series_counter) {
    hipLaunchKernelGGL(k_step_count, dim3((nrep + 255) / 256), dim3(256), 0, s, nrep, reps, rep_step, rep_t, h,
                       series_counter);
    LAUNCH_CHECK();
}

}  // namespace mds
