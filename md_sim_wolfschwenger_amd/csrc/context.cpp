// context.cpp - device context, step driver and C ABI (host code).
//
// Each entry point maps to one reference function (see include/mdsim_capi.h).
// The step follows RunSimulation's loop body (src/simulation.cpp:466-520);
// the adaptive controller is a restatement of AdjustTimeStep (912-951).
//
// A context holds one or more replicas: independent periodic systems whose
// particles are concatenated in every device array (replica r owns
// [off_r, off_r + n_r)). Every kernel launch covers all replicas; the kernels
// find a particle's, cell's or k-vector's replica in a small table (Rep).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.h"
#include "mdsim_capi.h"
#include "philox.h"

using namespace mds;

namespace {

// Device memory and graph executables that a live context gives up are
// released later, when a context of the device is destroyed: freeing device
// memory waits for every stream of the device, so a context that re-plans or
// re-captures while other contexts on the device run (the serving loop's
// lanes) would otherwise stall until their work drains. Buffers grow by half
// again at least, so the queued memory stays below twice the live size.
struct Retired {
    int dev;
    void* mem;
    hipGraphExec_t exec;
};
std::mutex g_retired_mu;
std::vector<Retired> g_retired;

int current_device() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return dev;
}

void retire_device_memory(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_retired_mu);
    g_retired.push_back({current_device(), p, nullptr});
}

void free_retired(int dev);

// One executable per re-capture: after 1024 of them (a long serving session)
// everything queued on the device is released, one wait in 1024 captures.
void retire_graph_exec(hipGraphExec_t e) {
    if (!e) return;
    const int dev = current_device();
    size_t queued = 0;
    {
        std::lock_guard<std::mutex> lk(g_retired_mu);
        g_retired.push_back({dev, nullptr, e});
        queued = g_retired.size();
    }
    if (queued > 1024) free_retired(dev);
}

void free_retired(int dev) {
    std::vector<Retired> out;
    {
        std::lock_guard<std::mutex> lk(g_retired_mu);
        auto keep = std::stable_partition(g_retired.begin(), g_retired.end(),
                                          [dev](const Retired& r) { return r.dev != dev; });
        out.assign(keep, g_retired.end());
        g_retired.erase(keep, g_retired.end());
    }
    if (out.empty()) return;
    // a queued executable may still be running (the 1024th re-capture
    // releases its own context's earlier ones): drain the device first
    (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    for (const Retired& r : out) {
        if (r.exec) (void)hipGraphExecDestroy(r.exec);
        if (r.mem) (void)hipFree(r.mem);
    }
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // give the memory up without waiting for the device (released later)
    void retire() {
        retire_device_memory(p);
        p = nullptr;
        n = 0;
    }
    // capacity >= count; contents are not kept. A buffer that grows takes
    // half again of its old capacity at least, so layout changes (admitted
    // and retired requests) rarely reallocate.
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) count = std::max(count, n + n / 2);
        retire();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

// A run of particles [src, src + len) of the old layout that goes to
// [dst, dst + len) of the new one.
struct Seg {
    int64_t src, dst, len;
};

// Rebuild a column-major array of `cols` columns (old column stride
// old_n, new stride new_n) from segments of the old one into `spare`, then
// swap the two; rows not covered by a segment are left for the caller to
// fill. The old buffer is kept as the next spare, so a layout change that
// fits the capacity allocates nothing. Stream-ordered.
template <typename T>
hipError_t repack(DevBuf<T>& b, DevBuf<T>& spare, int cols, int64_t old_n, int64_t new_n, int64_t min_n,
                  const std::vector<Seg>& segs, hipStream_t s) {
    const size_t need = std::max<size_t>(1, (size_t)cols * (size_t)std::max(new_n, min_n));
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (spare.n < need) {
        spare.retire();
        if ((e = hipMalloc((void**)&spare.p, need * sizeof(T))) != hipSuccess) return e;
        spare.n = need;
    }
    if (b.p && old_n > 0)
        for (int d = 0; d < cols; ++d)
            for (const Seg& g : segs)
                if (g.len > 0) {
                    e = hipMemcpyAsync(spare.p + (size_t)d * new_n + g.dst, b.p + (size_t)d * old_n + g.src,
                                       (size_t)g.len * sizeof(T), hipMemcpyDeviceToDevice, s);
                    if (e != hipSuccess) return e;
                }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    std::swap(b.p, spare.p);
    std::swap(b.n, spare.n);
    return hipSuccess;
}

// Kernel classes timed by HIP events.
enum KClass {
    KC_SETUP = 0,
    KC_CONTACTS,
    KC_SINCOS,
    KC_MOMENTS,
    KC_PAIRS_FORCE,
    KC_PAIRS_FIELD,
    KC_KBUILD,
    KC_KREDUCE,
    KC_KSCATTER_FORCE,
    KC_KSCATTER_FIELD,
    KC_STAGE,
    KC_TRANSLATE,
    KC_NOISE,
    KC_LIST,
    KC_COUNT
};
const char* kKClassNames[KC_COUNT] = {"setup",   "contacts", "sincos",          "moments",          "pairs_force",
                                      "pairs_field", "kbuild", "kreduce",        "kscatter_force",   "kscatter_field",
                                      "stage",   "translate", "noise",           "list_build"};

// Butcher tableaux of IntegrationRot (src/simulation.cpp:721-862): stage
// coefficients a, accepted combination b (BS: its 4th stage point), error
// combination e of k_Mm, controller exponent.
struct Tableau {
    int stages = 2;
    double a[8][8] = {};
    double b[8] = {};
    double e[8] = {};
    double power = 0.5;
};
Tableau make_tableau(int solver) {
    Tableau t;
    if (solver == MDS_BOGACKI_SHAMPINE) {
        t.stages = 4;
        t.a[1][0] = 1.0 / 2.0;
        t.a[2][1] = 3.0 / 4.0;
        t.a[3][0] = 2.0 / 9.0;
        t.a[3][1] = 1.0 / 3.0;
        t.a[3][2] = 4.0 / 9.0;
        t.b[0] = 2.0 / 9.0;
        t.b[1] = 1.0 / 3.0;
        t.b[2] = 4.0 / 9.0;
        t.e[0] = -5.0 / 72.0;
        t.e[1] = 1.0 / 12.0;
        t.e[2] = 1.0 / 9.0;
        t.e[3] = -1.0 / 8.0;
        t.power = 1.0 / 3.0;
    } else if (solver == MDS_DORMAND_PRINCE) {
        t.stages = 7;
        const double a[7][6] = {{0},
                                {1.0 / 5.0},
                                {3.0 / 40.0, 9.0 / 40.0},
                                {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
                                {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
                                {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
                                {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
        for (int s = 0; s < 7; ++s)
            for (int r = 0; r < 6; ++r) t.a[s][r] = a[s][r];
        for (int r = 0; r < 6; ++r) t.b[r] = a[6][r];
        const double e[7] = {71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0, -17253.0 / 339200.0, 22.0 / 525.0,
                             -1.0 / 40.0};
        for (int r = 0; r < 7; ++r) t.e[r] = e[r];
        t.power = 1.0 / 5.0;
    } else {
        t.stages = 2;
        t.a[1][0] = 1.0;
        t.b[0] = t.b[1] = 0.5;
        t.e[0] = 0.5;
        t.e[1] = -0.5;
        t.power = 1.0 / 2.0;
    }
    return t;
}

int blocks_for(int64_t n, int bs = 256) { return (int)std::max<int64_t>(1, (n + bs - 1) / bs); }

// splitmix64 finaliser (Steele, Lea, Flood, OOPSLA 2014): replica seeds
uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Host-side description of a replica (what the device Rep table is planned from).
struct RepMeta {
    int64_t n = 0, off = 0;
    double L = 0, rc_sr = 0, alpha = 0, rc = 0, field_z = 0, rmax = 0;
    double mref = 1;  // largest magnetic moment (unit of the FP32 pair path)
    int32_t nc = 0;
    bool ewald_set = false;
    uint64_t seed = 0;
    double margin_scale = 0.1;
    double list_scale = 1.0;  // pair-list region over its estimate (grows when a list does not fit)
    uint64_t stop_at = ~(uint64_t)0;  // step count at which the replica stops (mds_replica_stop_at)
};

constexpr int32_t kSeriesCap = 1024;  // <m_z> rows buffered on the device per run chunk

}  // namespace

struct mds_ctx {
    mds_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for the k-space build, which runs beside the pair
    // kernels (same CU mask as `stream`); fork/join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // third stream for the contact search and overlap correction, which run
    // beside the cell sort and the pair-list kernels (same CU mask). The two
    // helper streams exist only while the overlap is on: a process gets a
    // few hardware queues (four by default), and contexts side by side
    // (serving lanes) keep theirs to one stream each.
    hipStream_t aux = nullptr;
    std::vector<uint32_t> cu_mask;  // the context's CU mask (empty: all CUs)
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    bool overlap_k = true;
    std::string err;

    // replicas
    std::vector<RepMeta> meta;
    std::vector<Rep> h_rep;
    int64_t n = 0;  // all particles
    int32_t ncu = 256;          // compute units of the device
    int32_t kspace_waves = -1;  // waves per SIMD the forked S(k) builds may hold (0: no limit, -1: per kernel)
    int32_t nrep = 0;
    bool state_set = false;
    bool planned = false;
    bool geometry_ready = false;
    bool philox = false;
    int32_t reach = 3;
    // host copies of per-particle constants (planning, error model)
    std::vector<double> h_r_mag, h_mag_mom, h_r_hydr;

    // per-particle constants and state (kept across replica changes)
    DevBuf<double> r_mag, mag_mom, zeta_rot, zeta_trans, vol_mag, c1, c2;
    DevBuf<double> pos, mm, ea, n_t, n_b, n_f;
    DevBuf<double> spare[11];  // re-layout targets of the ten arrays above and pos_build
    // pair list (list_skin > 0, the skin as a fraction of r_c): build
    // positions (original order, kept across layouts), per-replica state,
    // this step's mode, largest squared displacement since the build, stream
    // lengths and offsets per (sorted slot, wavefront), the streams, words
    // per replica and builds per replica
    double list_skin = 0.005;
    DevBuf<double> pos_build;
    DevBuf<int32_t> list_state, list_mode, lcount, loff, list;
    DevBuf<int32_t> sort_skip;  // set on the device when every replica keeps its list
    DevBuf<uint64_t> list_maxd2, list_builds;
    DevBuf<int64_t> list_used;
    DevBuf<float4> rec_build32;
    int64_t list_words = 0;
    bool list_on() const { return list_skin > 0; }
    int64_t reserve_n = 0;     // capacity floor in particles (mds_reserve)
    // per-step scratch
    DevBuf<double> pos0, pos_cur, force, field, mom;
    DevBuf<double4> rec_pos, rec_mom;
    DevBuf<double4> frec;    // positions and radii in fine-cell order (contact search)
    DevBuf<uint64_t> pair_acc;  // [12n] cell-sorted fixed-point pair sums
    DevBuf<int32_t> order_tmp;  // [n] main-grid cell lists in placement (atomic) order
    // EvalTransCoeffs observables (mds_observe): unwrapped and reference
    // positions, squared displacements, per-step series and last values
    DevBuf<double> true_pos, pos_ref, obs_d2, obs_series, obs_last;
    bool observe = false;
    DevBuf<float4> rec_pos32, rec_mom32;
    DevBuf<int32_t> rep_of;
    // replica table and per-replica device values
    DevBuf<Rep> reps;
    DevBuf<double> rep_t, mz_out, msum;
    DevBuf<uint64_t> rep_step;
    // integrator
    DevBuf<double> kmm, kea, pt[2], new_mm, new_ea;
    // cell lists
    int32_t work_cap = 0, cand_cap = 0, cells_main = 0, cells_fine = 0;
    std::vector<int4> h_sten;
    DevBuf<int4> stencil;
    DevBuf<int32_t> cell_rep;
    DevBuf<int32_t> key_main, key_fine, cnt_main, start_main, end_main, cur_main, sorted_main, inv_perm, work, nwork;
    DevBuf<int32_t> cnt_fine, start_fine, end_fine, cur_fine, sorted_fine;
    DevBuf<int32_t> cand_count, cand_off, cand_j, ev_head, ev_tail, ev_next, n_events, err_flag, moved_list, n_moved;
    DevBuf<int32_t> cand_base, cand_tmp;  // contact runs: [n + 1] bases and the append counter, [cand_cap] runs
    DevBuf<uint64_t> ev_key;
    DevBuf<double> ev_pos, max_disp;
    // k space
    std::vector<KRow> h_rows;
    std::vector<KRow> h_rows_m;  // the same rows in the matrix-core kernels' order
    std::vector<KSeg> h_segs;
    std::vector<KPack> h_pack;
    int32_t nk_all = 0, kb_items = 0, kbm_items = 0, ksm_items = 0, nc_max = 0;
    // k space per replica (Rep::kmode): replicas on the matrix cores and their
    // largest n_c (sizes the kernels' LDS), and whether any is on the vector kernels
    int32_t nc_mfma_build = 0, nc_mfma_scatter = 0;
    bool any_mfma_build = false, any_mfma_scatter = false, any_vec_build = false, any_vec_scatter = false;
    DevBuf<KRow> rows, rows_m;
    DevBuf<KSeg> segs;
    DevBuf<KPack> pack_desc;
    DevBuf<double4> pack;
    DevBuf<double> kpartial, tab;
    DevBuf<float> tab32;     // FP32 sin/cos tables of the matrix-core replicas
    DevBuf<float4> pack32;   // FP32 tile-major scatter entries of the matrix-core replicas
    // reductions
    DevBuf<double> press_partial, err_partial, scal;  // scal[0] = error norm
    DevBuf<int32_t> scan_tmp, scan_tmp2;  // scan tiles of the main chain and of the contact chain
    std::vector<int2> h_chunks;  // reduction chunks (replica, first particle)
    DevBuf<int2> red_chunks;
    DevBuf<double> red_part;
    DevBuf<uint64_t> pair_counter;
    DevBuf<int32_t> series_counter;
    DevBuf<double> series;
    // HIP graph of one fixed-dt step
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    double graph_h = 0;

    // timing
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    int64_t launches[KC_COUNT] = {};
    double total_ms[KC_COUNT] = {};
    bool count_work = false;
    int64_t last_events = 0;
    // graph-replayed runs: steps per checked chunk, the chunk-start state
    // (pos, mm, ea, unwrapped positions; replica clocks and step counters) and
    // the number of chunks redone after a raised flag
    int64_t check_every = 128;
    DevBuf<double> snap;
    DevBuf<double> snap_t;
    DevBuf<uint64_t> snap_step;
    DevBuf<int32_t> snap_list;
    int64_t chunks_redone = 0;

    hipEvent_t get_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    struct Span {
        mds_ctx* c;
        int cls;
        hipEvent_t a = nullptr;
        Span(mds_ctx* cc, int k) : c(cc), cls(k) {
            if (c->timing) {
                a = c->get_event();
                (void)hipEventRecord(a, c->stream);
            }
        }
        ~Span() {
            if (c->timing) {
                hipEvent_t b = c->get_event();
                (void)hipEventRecord(b, c->stream);
                c->pending.push_back({cls, {a, b}});
            }
        }
    };
    void collect() {
        if (pending.empty()) return;
        (void)hipStreamSynchronize(stream);
        for (auto& p : pending) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
            launches[p.first] += 1;
            total_ms[p.first] += ms;
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        pending.clear();
    }
    // the executable is released with the context's other retired objects
    // (its destruction can wait for the whole device)
    void drop_graph() {
        retire_graph_exec(gexec);
        gexec = nullptr;
    }
    ~mds_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (auto& p : pending) {
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux) {
            (void)hipStreamSynchronize(aux);
            (void)hipStreamDestroy(aux);
        }
        if (ev_fork2) (void)hipEventDestroy(ev_fork2);
        if (ev_join2) (void)hipEventDestroy(ev_join2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

mds_status fail(mds_ctx* c, mds_status st, const std::string& msg) {
    if (c) c->err = msg;
    return st;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(c, e_ == hipErrorOutOfMemory ? MDS_ENOMEM : MDS_EDEVICE,            \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

mds_status upload(mds_ctx* c, double* dst, const double* src, size_t count) {
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return MDS_OK;
}
mds_status download(mds_ctx* c, double* dst, const double* src, size_t count) {
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}
// column-major N x 3 slice [off, off + m) of an array with stride n
mds_status upload_cols(mds_ctx* c, double* dst, int64_t n, int64_t off, int64_t m, const double* src) {
    for (int d = 0; d < 3; ++d)
        HIPCHK(hipMemcpyAsync(dst + (size_t)d * n + off, src + (size_t)d * m, (size_t)m * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    return MDS_OK;
}
mds_status download_cols(mds_ctx* c, double* dst, const double* src, int64_t n, int64_t off, int64_t m) {
    for (int d = 0; d < 3; ++d)
        HIPCHK(hipMemcpyAsync(dst + (size_t)d * m, src + (size_t)d * n + off, (size_t)m * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}

// ---------------------------------------------------------------- planning
// Cell grids of replica r: the main grid's stencil of +-reach cells must
// cover the candidate radius r_c + margin (the margin covers positions moved
// by the overlap correction); the fine grid covers the contact search radius.
// Main-grid cells per dimension for a stencil half-width `reach`, around the
// pair search radius rl (r_c, or the pair list's r_c + skin).
int32_t main_cells(const RepMeta& m, double rl, double margin, int reach) {
    const double cmin = (rl * (1 + 1e-9) + margin) / reach;
    int32_t nc = (int32_t)std::floor(m.L / cmin);
    if (nc < 2 * reach + 1) nc = 1;
    return std::min(nc, 96);
}

// Cell reach 0 picks one per replica: the largest of 4..1 whose cells hold
// at least 32 particles on average (one i-chunk of the pair kernel is up to 64
// particles of one cell, so emptier cells leave lanes idle), else 1.
// Pair search radius: r_c, with a pair list r_c + skin. A list built at
// positions b holds every pair with |b_i - b_j| < r_l; while every particle
// is within skin / 2 of its b, a pair within r_c is within r_l at b.
double list_radius(const mds_ctx* c, const RepMeta& m) { return m.rc * (1.0 + c->list_skin); }

int replica_reach(const mds_ctx* c, const RepMeta& m, double margin) {
    if (c->reach > 0) return c->reach;
    for (int r = 4; r >= 1; --r) {
        const double nc = main_cells(m, list_radius(c, m), margin, r);
        if ((double)m.n / (nc * nc * nc) >= 32.0) return r;
    }
    return 1;
}

void plan_grid(mds_ctx* c, const RepMeta& m, Rep& R) {
    const double dmax = 2 * m.rmax + 1e-10;
    R.margin = m.margin_scale * dmax;
    const int reach = replica_reach(c, m, R.margin);
    const double rl = list_radius(c, m);
    const int32_t nc = main_cells(m, rl, R.margin, reach);
    R.ncell = nc;
    R.cell = m.L / nc;
    // contact-search grid: cells of at least the search diameter, about two
    // per particle at most
    const double fmin = dmax + R.margin;
    int32_t fn = (int32_t)std::floor(m.L / fmin);
    fn = std::min<int32_t>(fn, std::max<int32_t>(3, (int32_t)std::ceil(std::cbrt(2.0 * (double)m.n))));
    if (fn < 3) fn = 1;
    fn = std::min(fn, 128);
    R.fncell = fn;
    R.fcell = m.L / fn;
    // stencil of neighbour-cell offsets whose nearest points lie within the
    // candidate radius (corner cells of the cube of +-reach are dropped)
    // The cell itself comes first, then the offsets positive in (z, y, x)
    // order (the pair kernel's half stencil: each pair of cells once), then
    // their negatives in the same order.
    const int H = nc == 1 ? 0 : reach;
    const double lim = rl * (1 + 1e-9) + R.margin;
    R.sten_off = (int32_t)c->h_sten.size();
    std::vector<int4> pos_half;
    for (int dz = -H; dz <= H; ++dz)
        for (int dy = -H; dy <= H; ++dy)
            for (int dx = -H; dx <= H; ++dx) {
                auto gap = [&](int o) { return std::max(std::abs(o) - 1, 0) * R.cell; };
                const double d2 = gap(dx) * gap(dx) + gap(dy) * gap(dy) + gap(dz) * gap(dz);
                const bool positive = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
                if (d2 < lim * lim && positive) pos_half.push_back(make_int4(dx, dy, dz, 0));
            }
    c->h_sten.push_back(make_int4(0, 0, 0, 0));
    for (const int4& o : pos_half) c->h_sten.push_back(o);
    for (const int4& o : pos_half) c->h_sten.push_back(make_int4(-o.x, -o.y, -o.z, 0));
    R.nhalf = 1 + (int32_t)pos_half.size();
    R.nsten = (int32_t)c->h_sten.size() - R.sten_off;
}

// k-space rows of replica r over the half space n_z > 0, or n_z = 0 and
// (n_y > 0, or n_y = 0 and n_x > 0), |n|^2 <= n_c^2; with the reference's
// full n_z = 0 plane at weight 1/2 (src/interaction.cpp:438-457) this is the
// same sum.
void plan_kspace(mds_ctx* c, const RepMeta& m, Rep& R) {
    const int nc = m.nc;
    const int64_t nc2 = (int64_t)nc * nc;
    R.row_off = (int32_t)c->h_rows.size();
    R.seg_off = (int32_t)c->h_segs.size();
    R.pk_off = (int32_t)c->h_pack.size();
    const double w = std::pow(c->cfg.my_pi / (m.L * m.alpha), 2.0);
    // e(k) = 2 exp(-w n^2) / n^2 (src/interaction.cpp:450-452)
    auto ek = [&](int64_t len) { return len > 0 ? 2.0 * std::exp(-w * (double)len) / (double)len : 0.0; };
    int32_t koff = 0;
    for (int nz = 0; nz <= nc; ++nz)
        for (int ny = (nz == 0 ? 0 : -nc); ny <= nc; ++ny) {
            const int64_t r2 = (int64_t)ny * ny + (int64_t)nz * nz;
            if (r2 > nc2) continue;
            int nxmax = 0;
            while ((int64_t)(nxmax + 1) * (nxmax + 1) + r2 <= nc2) ++nxmax;
            const bool half = (ny == 0 && nz == 0);
            if (half && nxmax == 0) continue;
            c->h_rows.push_back({ny, nz, nxmax, (int32_t)c->h_pack.size() - R.pk_off, koff});
            for (int a = half ? 1 : 0; a <= nxmax; a += KSEG)
                c->h_segs.push_back({ny, nz, a, std::min(nxmax, a + KSEG - 1), koff, nxmax});
            // scatter table: |nx| = a pairs (+a, ny, nz) with (-a, ny, nz)
            for (int a = 0; a <= nxmax; ++a) {
                KPack d;
                if (half) {
                    d.kp = a >= 1 ? koff + a - 1 : -1;
                    d.km = -1;
                } else {
                    d.kp = koff + a + nxmax;
                    d.km = a >= 1 ? koff - a + nxmax : -1;
                }
                d.e = d.kp >= 0 ? ek((int64_t)a * a + r2) : 0.0;
                d.t32 = -1;
                d.pad_ = 0;
                c->h_pack.push_back(d);
            }
            // pad to an even entry count (the scatter takes two at a time)
            if ((nxmax + 1) % 2) c->h_pack.push_back(KPack{-1, -1, 0.0, -1, 0});
            koff += half ? nxmax : 2 * nxmax + 1;
        }
    R.nrows = (int32_t)c->h_rows.size() - R.row_off;
    R.nseg = (int32_t)c->h_segs.size() - R.seg_off;
    R.npk = (int32_t)c->h_pack.size() - R.pk_off;
    // the matrix-core kernels take the rows by decreasing |nx| range (ties in
    // table order), so the rows of one 16-row tile have similar lengths and
    // the tiles' products stop near their longest row; each packed entry
    // gets its slot in the FP32 tile-major copy, ((row / 16) (n_c + 2) + a)
    // 16 + row % 16 with the row's place in that order
    {
        std::vector<int32_t> ord(R.nrows);
        for (int32_t r = 0; r < R.nrows; ++r) ord[r] = r;
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
            return c->h_rows[R.row_off + x].nxmax > c->h_rows[R.row_off + y].nxmax;
        });
        for (int32_t q = 0; q < R.nrows; ++q) {
            const KRow& row = c->h_rows[R.row_off + ord[q]];
            c->h_rows_m.push_back(row);
            const int32_t ne = ((row.nxmax + 2) / 2) * 2;  // entries with the padding one
            for (int32_t a = 0; a < ne; ++a)
                c->h_pack[R.pk_off + row.poff + a].t32 = ((q / 16) * (nc + 2) + a) * 16 + q % 16;
        }
    }
    // quarters of the table by entries for the four wavefronts of a scatter block
    {
        int32_t q = 1, acc = 0;
        R.ksplit[0] = 0;
        for (int r = 0; r < R.nrows && q < 4; ++r) {
            const KRow& row = c->h_rows[R.row_off + r];
            acc += ((row.nxmax + 2) / 2) * 2;
            while (q < 4 && (int64_t)acc * 4 >= (int64_t)q * R.npk) R.ksplit[q++] = r + 1;
        }
        while (q < 4) R.ksplit[q++] = R.nrows;
    }
    R.nk = koff;
    R.nc = nc;
    const double pi = c->cfg.my_pi, f = c->cfg.fac_ewald, L = m.L;
    // InitEwaldConst, src/interaction.cpp:366-369
    R.c_flux = 4.0 * pi * f / std::pow(L, 3.0);
    R.c_force = 8.0 * pi * pi * f / std::pow(L, 4.0);
    R.c_surf = f * 4.0 * pi / (3 * std::pow(L, 3.0));
    R.c_self = f * 4.0 * std::pow(m.alpha, 3.0) / (3.0 * std::sqrt(pi));
    R.two_pi_over_L = 2 * pi / L;
    R.nsegw = (R.nseg + 63) / 64;
    // particle chunks so that the build has ~8 k wavefronts (8 per SIMD)
    int64_t chunk = (m.n * std::max<int64_t>(1, R.nsegw) + 8191) / 8192;
    chunk = std::max<int64_t>(32, std::min<int64_t>(1024, chunk));
    R.kchunk = (int32_t)chunk;
    R.nchunks = (int32_t)((m.n + chunk - 1) / chunk);
    // matrix-core build: blocks of 4 x 16 rows, passes of KBM_NX values of |nx|
    R.nrowg = (R.nrows + KBM_ROWS - 1) / KBM_ROWS;
    R.kbm_pass = (nc + KBM_NX) / KBM_NX;
}

// Ewald constants and FP32 units of replica r.
void plan_ewald(mds_ctx* c, const RepMeta& m, Rep& R) {
    R.alpha = m.alpha;
    R.alpha2 = m.alpha * m.alpha;
    R.alpha4 = R.alpha2 * R.alpha2;
    R.rc = m.rc;
    R.rc_sr = m.rc_sr;
    const double f = c->cfg.fac_ewald;
    if (c->cfg.precision == MDS_PREC_MIXED) {
        R.ell = m.rc;
        R.inv_ell = 1.0 / m.rc;
        R.inv_mref = 1.0 / m.mref;
        R.bscale = f * m.mref / std::pow(R.ell, 3.0);
        R.fscale = f * m.mref * m.mref / std::pow(R.ell, 4.0);
        R.pscale = R.ell * R.fscale;
    } else {
        R.ell = R.inv_ell = R.inv_mref = 1.0;
        R.bscale = R.fscale = R.pscale = f;
    }
    R.al_s = (float)(m.alpha * R.ell);
    R.al2_s = R.al_s * R.al_s;
    R.al4_s = R.al2_s * R.al2_s;
    // Fixed-point units of the pair sums: s1 = 2^(e - 17) below the field
    // (M / D^3) and force (M^2 / D^4) scales of the largest moment M at
    // distance D = r_max, in the sums' units; resolution s1 2^-40, range
    // 2^50 s1 (about 1e10 times the scale).
    const double M = m.mref * R.inv_mref, D = m.rmax * R.inv_ell;
    const double scale[2] = {M / (D * D * D), M * M / (D * D * D * D)};
    for (int q = 0; q < 2; ++q) {
        const int e = (std::isfinite(scale[q]) && scale[q] > 0) ? std::ilogb(scale[q]) - 17 : 0;
        R.fx_s[q] = std::ldexp(1.0, e);
        R.fx_is[q] = std::ldexp(1.0, -e);
    }
    // pressure of the pairs with a moved endpoint (r u F summed by atomics in
    // any order): fixed point too, in units of the force unit times the box
    // edge (SI), so that the sum does not depend on the order of the adds
    const double ps = R.fx_s[1] * R.fscale * m.L;
    const int ep = (std::isfinite(ps) && ps > 0) ? std::ilogb(ps) + 1 : 0;
    R.px_s = std::ldexp(1.0, ep);
    R.px_is = std::ldexp(1.0, -ep);
}

// Words of a replica's pair-list region: the pairs within r_l of a uniform
// suspension with a quarter extra for the quad padding of each candidate's
// hits and two header words per window, times the replica's growth factor.
int64_t list_capacity(const RepMeta& m, double rl) {
    const double frac = std::min(1.0, 4.0 / 3.0 * 3.14159265358979 * rl * rl * rl / (m.L * m.L * m.L));
    const double pairs = 0.5 * (double)m.n * (double)(m.n - 1) * frac;
    const double words = (1.25 * pairs + 16.0 * (double)m.n + 4096.0) * m.list_scale;
    return (int64_t)std::min(words, 4.0e18);
}

// Replica list states VALID -> STALE (rebuilt from their build positions at
// the next geometry pass); INVALID stays.
mds_status stale_lists(mds_ctx* c) {
    if (!c->list_on() || c->nrep <= 0 || !c->list_state.p) return MDS_OK;
    std::vector<int32_t> st(c->nrep);
    HIPCHK(hipMemcpyAsync(st.data(), c->list_state.p, c->nrep * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t& v : st)
        if (v != LIST_INVALID) v = LIST_STALE;
    HIPCHK(hipMemcpyAsync(c->list_state.p, st.data(), c->nrep * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}

// Plan grids, stencils and k space of every replica, size the scratch
// buffers and upload the tables.
mds_status plan_all(mds_ctx* c) {
    if (c->nrep <= 0 || c->n <= 0) return fail(c, MDS_EINVAL, "no particles");
    for (const RepMeta& m : c->meta)
        if (!m.ewald_set) return fail(c, MDS_EINVAL, "Ewald parameters not set");
    c->drop_graph();
    c->h_rep.assign(c->nrep, Rep{});
    c->h_sten.clear();
    c->h_rows.clear();
    c->h_rows_m.clear();
    c->h_segs.clear();
    c->h_pack.clear();
    std::vector<int32_t> cell_rep, rep_of((size_t)c->n);
    c->h_chunks.clear();
    int64_t cells = 0, fcells = 0, work_cap = 0, tab = 0, part = 0, kb = 0, kbm = 0, ksm = 0, kall = 0;
    int64_t tab32 = 0, pk32 = 0, list_words = 0;
    int32_t nc_max = 0;
    int32_t nc_mfma_build = 0, nc_mfma_scatter = 0;
    bool any_mfma_build = false, any_mfma_scatter = false, any_vec_build = false, any_vec_scatter = false;
    for (int r = 0; r < c->nrep; ++r) {
        const RepMeta& m = c->meta[r];
        Rep& R = c->h_rep[r];
        R.off = m.off;
        R.n = m.n;
        R.L = m.L;
        R.field_z = m.field_z;
        R.seed = m.seed;
        R.stop_at = m.stop_at;
        plan_grid(c, m, R);
        plan_ewald(c, m, R);
        plan_kspace(c, m, R);
        R.rlist = list_radius(c, m);
        R.skin_half2 = 0.25 * (R.rlist - m.rc) * (R.rlist - m.rc);
        R.list_base = list_words;
        R.list_cap = c->list_on() ? list_capacity(m, R.rlist) : 0;
        if (R.list_cap >= ((int64_t)1 << 31))
            return fail(c, MDS_ECAPACITY, "the pair list of a replica exceeds 2^31 words (use a smaller r_c)");
        list_words += R.list_cap;
        const int64_t ncm = (int64_t)R.ncell * R.ncell * R.ncell, ncf = (int64_t)R.fncell * R.fncell * R.fncell;
        R.cell_off = (int32_t)cells;
        R.fcell_off = (int32_t)fcells;
        cells += ncm;
        fcells += ncf;
        cell_rep.insert(cell_rep.end(), (size_t)ncm, r);
        work_cap += (m.n + 63) / 64 + ncm;
        R.tab_off = tab;
        tab += m.n * 3 * (int64_t)(R.nc + 1) * 2;
        R.k_off = (int32_t)kall;
        kall += R.nk;
        R.part_off = part;
        part += 2 * (int64_t)R.nk * R.nchunks;
        R.kb_off = (int32_t)kb;
        kb += (R.nk > 0) ? (int64_t)R.nsegw * R.nchunks : 0;
        R.ksm_off = (int32_t)ksm;
        ksm += (m.n + 63) / 64;
        nc_max = std::max(nc_max, R.nc);
        // the mixed path takes each replica's k space to the matrix cores when
        // its n_c fits their LDS tiles; the choice is the replica's own, so
        // its results do not depend on the other replicas of the context
        const bool mixed = c->cfg.precision == MDS_PREC_MIXED;
        R.kmode = (mixed && R.nc <= KBM_NCMAX ? 1 : 0) | (mixed && R.nc <= KSM_NCMAX ? 2 : 0);
        if (R.nk > 0) {
            if (R.kmode & 1) nc_mfma_build = std::max(nc_mfma_build, R.nc), any_mfma_build = true;
            else any_vec_build = true;
        }
        if (R.kmode & 2) nc_mfma_scatter = std::max(nc_mfma_scatter, R.nc), any_mfma_scatter = true;
        else any_vec_scatter = true;
        R.tab32_off = tab32;
        if (R.kmode) tab32 += m.n * 3 * (int64_t)tw4_of(R.nc);
        R.pk32_off = pk32;
        if (R.kmode & 2) pk32 += (int64_t)((R.nrows + 15) / 16) * (R.nc + 2) * 16;
        R.kbm_off = (int32_t)kbm;
        kbm += (R.nk > 0) ? (int64_t)((R.nrowg + 3) / 4) * R.kbm_pass * 8 * ((R.nchunks + 7) / 8) : 0;
        std::fill(rep_of.begin() + m.off, rep_of.begin() + m.off + m.n, r);
        R.red_off = (int32_t)c->h_chunks.size();
        for (int64_t b = m.off; b < m.off + m.n; b += RED_CHUNK_PARTICLES) c->h_chunks.push_back(make_int2(r, (int)b));
        R.nred = (int32_t)c->h_chunks.size() - R.red_off;
    }
    // work items pack (cell << 8) | chunk into an int32; pair-queue entries
    // (j << 6) | lane, with bit 31 free
    if (cells >= (1 << 23) || fcells >= (1 << 30) || work_cap >= (1 << 30) || kb >= (1 << 30) || kbm >= (1 << 30))
        return fail(c, MDS_ECAPACITY, "too many cells or k-space work items in one context");
    if (c->n >= (int64_t)1 << 25) return fail(c, MDS_ECAPACITY, "more than 2^25 particles in one context");
    c->cells_main = (int32_t)cells;
    c->cells_fine = (int32_t)fcells;
    c->work_cap = (int32_t)work_cap;
    c->cand_cap = (int32_t)std::min<int64_t>(2 * c->n + 4096, (int64_t)1 << 29);
    c->nk_all = (int32_t)kall;
    c->kb_items = (int32_t)kb;
    c->kbm_items = (int32_t)kbm;
    c->ksm_items = (int32_t)ksm;
    c->nc_max = nc_max;
    c->nc_mfma_build = nc_mfma_build;
    c->nc_mfma_scatter = nc_mfma_scatter;
    c->any_mfma_build = any_mfma_build;
    c->any_mfma_scatter = any_mfma_scatter;
    c->any_vec_build = any_vec_build;
    c->any_vec_scatter = any_vec_scatter;
    HIPCHK(c->cnt_main.ensure(cells));
    HIPCHK(c->start_main.ensure(cells));
    HIPCHK(c->end_main.ensure(cells));
    HIPCHK(c->cur_main.ensure(cells));
    HIPCHK(c->cell_rep.ensure(cells));
    HIPCHK(c->work.ensure(work_cap));
    HIPCHK(c->cnt_fine.ensure(fcells));
    HIPCHK(c->start_fine.ensure(fcells));
    HIPCHK(c->end_fine.ensure(fcells));
    HIPCHK(c->cur_fine.ensure(fcells));
    HIPCHK(c->cand_j.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->cand_tmp.ensure((size_t)c->cand_cap));
    HIPCHK(c->ev_next.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->ev_key.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->ev_pos.ensure(6 * (size_t)c->cand_cap));
    HIPCHK(c->moved_list.ensure(2 * (size_t)c->cand_cap));
    // per-wavefront slots of the pair kernels, then 6 fixed-point words per
    // replica for the moved pairs
    HIPCHK(c->press_partial.ensure(3 * (size_t)work_cap * PAIR_SLICES + 6 * (size_t)c->nrep));
    HIPCHK(c->stencil.ensure(std::max<size_t>(1, c->h_sten.size())));
    HIPCHK(c->rows.ensure(std::max<size_t>(1, c->h_rows.size())));
    HIPCHK(c->rows_m.ensure(std::max<size_t>(1, c->h_rows_m.size())));
    HIPCHK(c->segs.ensure(std::max<size_t>(1, c->h_segs.size())));
    HIPCHK(c->pack_desc.ensure(std::max<size_t>(1, c->h_pack.size())));
    HIPCHK(c->pack.ensure(std::max<size_t>(1, c->h_pack.size())));
    HIPCHK(c->kpartial.ensure(std::max<int64_t>(1, part)));
    HIPCHK(c->tab.ensure(std::max<int64_t>(1, tab)));
    HIPCHK(c->tab32.ensure(std::max<int64_t>(1, tab32)));
    HIPCHK(c->pack32.ensure(std::max<int64_t>(1, pk32)));
    HIPCHK(c->reps.ensure(c->nrep));
    HIPCHK(c->rep_of.ensure(c->n));
    HIPCHK(c->msum.ensure(3 * (size_t)c->nrep));
    HIPCHK(c->mz_out.ensure(c->nrep));
    HIPCHK(c->series.ensure((size_t)kSeriesCap * c->nrep));
    HIPCHK(c->scan_tmp.ensure(scan_tiles(std::max<int64_t>({cells, fcells, c->n})) + 1));
    HIPCHK(c->scan_tmp2.ensure(scan_tiles(std::max<int64_t>(fcells, c->n)) + 1));
    HIPCHK(c->red_chunks.ensure(c->h_chunks.size()));
    HIPCHK(c->red_part.ensure(3 * c->h_chunks.size()));
    c->list_words = list_words;
    if (c->list_on()) {
        HIPCHK(c->list.ensure(std::max<int64_t>(1, list_words)));
        HIPCHK(c->list_mode.ensure(c->nrep));
        HIPCHK(c->sort_skip.ensure(1));
        HIPCHK(c->list_maxd2.ensure(c->nrep));
        HIPCHK(c->list_used.ensure(c->nrep));
    }
    hipStream_t s = c->stream;
    HIPCHK(hipMemcpyAsync(c->reps.p, c->h_rep.data(), c->nrep * sizeof(Rep), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->rep_of.p, rep_of.data(), rep_of.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->cell_rep.p, cell_rep.data(), cell_rep.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                          s));
    HIPCHK(hipMemcpyAsync(c->red_chunks.p, c->h_chunks.data(), c->h_chunks.size() * sizeof(int2),
                          hipMemcpyHostToDevice, s));
    if (!c->h_sten.empty())
        HIPCHK(hipMemcpyAsync(c->stencil.p, c->h_sten.data(), c->h_sten.size() * sizeof(int4), hipMemcpyHostToDevice,
                              s));
    if (!c->h_rows.empty()) {
        HIPCHK(hipMemcpyAsync(c->rows.p, c->h_rows.data(), c->h_rows.size() * sizeof(KRow), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->rows_m.p, c->h_rows_m.data(), c->h_rows_m.size() * sizeof(KRow),
                              hipMemcpyHostToDevice, s));
    }
    if (!c->h_segs.empty())
        HIPCHK(hipMemcpyAsync(c->segs.p, c->h_segs.data(), c->h_segs.size() * sizeof(KSeg), hipMemcpyHostToDevice, s));
    if (!c->h_pack.empty())
        HIPCHK(hipMemcpyAsync(c->pack_desc.p, c->h_pack.data(), c->h_pack.size() * sizeof(KPack),
                              hipMemcpyHostToDevice, s));
    // the FP32 scatter tiles' slots that no entry writes (rows past the
    // replica's last, |nx| past the row's end) stay zero
    if (pk32 > 0) HIPCHK(launch_fill_f64(s, (double*)c->pack32.p, 2 * pk32, 0.0));
    if (c->list_on()) {
        HIPCHK(launch_fill_f64(s, (double*)c->list_maxd2.p, c->nrep, 0.0));
        HIPCHK(launch_fill_f64(s, (double*)c->list_used.p, c->nrep, 0.0));
        // a list is a function of its build positions: the replicas that have
        // them rebuild it from them (new regions, grids or offsets)
        mds_status st = stale_lists(c);
        if (st != MDS_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(s));
    c->planned = true;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status size_scratch_buffers(mds_ctx* c);

// Per-particle scratch of n particles (contents not kept). Some of it
// carries state from step to step (the main grid's sort and work items, the
// pair lists' streams, which a kept list reuses): if any of that moves, the
// plan is redone, which rebuilds the lists from their build positions.
mds_status size_scratch(mds_ctx* c) {
    const void* kept[] = {c->key_main.p, c->sorted_main.p, c->inv_perm.p, c->lcount.p, c->loff.p, c->rec_build32.p};
    mds_status st = size_scratch_buffers(c);
    const void* now[] = {c->key_main.p, c->sorted_main.p, c->inv_perm.p, c->lcount.p, c->loff.p, c->rec_build32.p};
    if (st == MDS_OK && c->planned && !std::equal(std::begin(kept), std::end(kept), std::begin(now))) {
        c->planned = false;
        c->geometry_ready = false;
        c->drop_graph();
    }
    return st;
}

mds_status size_scratch_buffers(mds_ctx* c) {
    const int64_t n = std::max(c->n, c->reserve_n);
    const size_t n3 = 3 * (size_t)n;
    for (auto* b : {&c->pos0, &c->pos_cur, &c->force, &c->field, &c->mom, &c->new_mm, &c->new_ea}) HIPCHK(b->ensure(n3));
    HIPCHK(c->pt[0].ensure(2 * n3));
    HIPCHK(c->pt[1].ensure(2 * n3));
    const int stages = make_tableau(c->cfg.fixed_dt ? MDS_HEUN_EULER : c->cfg.solver).stages;
    HIPCHK(c->kmm.ensure((size_t)stages * n3));
    HIPCHK(c->kea.ensure((size_t)stages * n3));
    HIPCHK(c->rec_pos.ensure(n));
    HIPCHK(c->rec_pos32.ensure(n));
    HIPCHK(c->frec.ensure(n));
    HIPCHK(c->rec_mom.ensure(n));
    HIPCHK(c->rec_mom32.ensure(n));
    HIPCHK(c->pair_acc.ensure(12 * (size_t)n));
    HIPCHK(c->order_tmp.ensure(n));
    for (auto* b : {&c->key_main, &c->key_fine, &c->sorted_main, &c->sorted_fine, &c->inv_perm, &c->cand_count,
                    &c->ev_head, &c->ev_tail})
        HIPCHK(b->ensure(n));
    HIPCHK(c->cand_off.ensure(n + 1));
    HIPCHK(c->cand_base.ensure(n + 1));
    HIPCHK(c->err_partial.ensure(blocks_for(n)));
    if (c->list_on()) {
        HIPCHK(c->rec_build32.ensure(n));
        HIPCHK(c->lcount.ensure((size_t)PAIR_SLICES * n));
        HIPCHK(c->loff.ensure((size_t)PAIR_SLICES * n));
    }
    return MDS_OK;
}

mds_status ensure_planned(mds_ctx* c) {
    if (c->planned) return MDS_OK;
    mds_status st = size_scratch(c);
    if (st != MDS_OK) return st;
    return plan_all(c);
}

Geometry geometry(mds_ctx* c) {
    Geometry g{};
    g.n = c->n;
    g.L = c->meta.empty() ? 0 : c->meta[0].L;
    g.reps = c->reps.p;
    g.rep_of = c->rep_of.p;
    g.cell_rep = c->cell_rep.p;
    g.ncell = c->h_rep.empty() ? 1 : c->h_rep[0].ncell;
    g.reach = c->reach;
    g.cell = c->h_rep.empty() ? 0 : c->h_rep[0].cell;
    g.cell_start = c->start_main.p;
    g.cell_end = c->end_main.p;
    g.sorted_idx = c->sorted_main.p;
    g.rec_pos = c->rec_pos.p;
    g.work = c->work.p;
    g.nwork = c->nwork.p;
    g.ev_head = c->ev_head.p;
    g.ev_next = c->ev_next.p;
    g.ev_key = c->ev_key.p;
    g.ev_pos = c->ev_pos.p;
    g.n_events = c->n_events.p;
    g.max_disp = c->max_disp.p;
    g.moved_list = c->moved_list.p;
    g.n_moved = c->n_moved.p;
    g.inv_perm = c->inv_perm.p;
    g.rec_pos32 = c->rec_pos32.p;
    g.rec_mom32 = c->rec_mom32.p;
    if (c->list_on()) {
        g.rec_build32 = c->rec_build32.p;
        g.list_mode = c->list_mode.p;
        g.lcount = c->lcount.p;
        g.loff = c->loff.p;
        g.list = c->list.p;
    }
    return g;
}

// Constants shared by all replicas (the per-replica ones come from Rep).
PairParams pair_params(mds_ctx* c) {
    PairParams pp{};
    pp.ir_pi = c->cfg.ir_pi;
    pp.fac = c->cfg.fac_ewald;
    pp.coating = c->cfg.enable_coating_pot;
    pp.eff_len = c->cfg.eff_len_surf_mol;
    pp.ster_fac = c->cfg.ster_rep_fac;
    pp.el_fac = c->cfg.electrostat_rep_fac;
    pp.kappa = c->cfg.kappa;
    pp.hamaker = c->cfg.hamaker;
    pp.r_mag = c->r_mag.p;
    return pp;
}

// Step-start positions, cell keys and the replicas' largest displacements
// since their list builds.
mds_status prepare_keys(mds_ctx* c) {
    const bool list = c->list_on();
    mds_ctx::Span sp(c, KC_SETUP);
    // with a pair list the main grid sorts the build positions: a kept
    // replica's cells, order and work items come out as built
    HIPCHK(launch_prepare(c->stream, c->n, c->pos.p, c->pos0.p, c->pos_cur.p, c->reps.p, c->rep_of.p, c->key_main.p,
                          c->key_fine.p, list ? c->pos_build.p : nullptr, c->list_state.p,
                          (unsigned long long*)c->list_maxd2.p));
    return MDS_OK;
}

// The sin/cos tables of EvalSinCos: with the keys, everything the k-space
// chain reads of the geometry.
mds_status prepare_sincos(mds_ctx* c, hipStream_t s = nullptr) {
    mds_ctx::Span sp(c, KC_SINCOS);
    HIPCHK(launch_sincos_tables(s ? s : c->stream, c->n, c->reps.p, c->rep_of.p, c->pos0.p, c->tab.p, c->tab32.p));
    return MDS_OK;
}

mds_status prepare_positions(mds_ctx* c) {
    mds_status st = prepare_keys(c);
    if (st != MDS_OK) return st;
    return prepare_sincos(c);
}

// Whether the contact search runs on the third stream (not when kernels are
// timed, which keeps them in order on one stream).
bool contacts_split(const mds_ctx* c) { return c->overlap_k && !c->timing; }

// Contact candidates and the sequential overlap correction of
// CompInteractVecs. They read the step-start positions and fine-grid keys
// only, and write nothing that the sin/cos tables, the k-space chain, the
// cell sort or the pair lists read: with `split` they run on the third
// stream from the point of this call (after prepare_keys), and k_mark_moved,
// which needs them, joins them.
mds_status contact_chain(mds_ctx* c, bool split) {
    const int64_t n = c->n;
    hipStream_t cs = split ? c->aux : c->stream;
    if (split) {
        HIPCHK(hipEventRecord(c->ev_fork2, c->stream));
        HIPCHK(hipStreamWaitEvent(c->aux, c->ev_fork2, 0));
    }
    {
        mds_ctx::Span sp(c, KC_CONTACTS);
        HIPCHK(launch_cell_sort(cs, n, c->key_fine.p, c->cells_fine, c->cnt_fine.p, c->start_fine.p, c->end_fine.p,
                                c->cur_fine.p, c->sorted_fine.p, c->scan_tmp2.p));
        HIPCHK(launch_clear_events(cs, n, c->ev_head.p, c->ev_tail.p, c->n_events.p, c->n_moved.p, c->max_disp.p));
        HIPCHK(launch_contacts(cs, n, c->reps.p, c->rep_of.p, c->pos0.p, c->r_mag.p, c->start_fine.p, c->end_fine.p,
                               c->sorted_fine.p, c->frec.p, c->cand_count.p, c->cand_off.p, c->cand_j.p,
                               c->cand_cap, c->cand_base.p, c->cand_tmp.p, c->err_flag.p, c->scan_tmp2.p));
        // displacement bound: half the smallest contact-search margin
        double dl = 1e300;
        for (const Rep& R : c->h_rep) dl = std::min(dl, 0.5 * R.margin);
        HIPCHK(launch_resolve(cs, c->nrep, n, c->reps.p, c->r_mag.p, c->cand_off.p, c->cand_j.p, c->pos_cur.p,
                              c->pos0.p, c->ev_head.p, c->ev_tail.p, c->ev_next.p, c->ev_key.p, c->ev_pos.p,
                              2 * c->cand_cap, c->n_events.p, c->max_disp.p, dl, c->moved_list.p, c->n_moved.p,
                              c->err_flag.p));
    }
    if (split) HIPCHK(hipEventRecord(c->ev_join2, c->aux));
    return MDS_OK;
}

// The pair geometry of CompInteractVecs after prepare_positions: cell lists
// (or the pair lists), contact candidates, sequential overlap correction.
// `forked`: contact_chain was already started on the third stream.
mds_status prepare_cells(mds_ctx* c, bool forked = false) {
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    const bool list = c->list_on();
    const bool split = contacts_split(c);
    mds_status st;
    if (split && !forked && (st = contact_chain(c, true)) != MDS_OK) return st;
    {
        mds_ctx::Span sp(c, KC_SETUP);
        // when every replica keeps its list (most steps) the main grid's
        // keys, sort, order and work items are those of the last step: the
        // kernels that make them return at once; the records always follow
        // the current positions
        const int32_t* skip = nullptr;
        if (list) {
            HIPCHK(launch_list_decide(s, c->nrep, c->reps.p, c->list_state.p, (unsigned long long*)c->list_maxd2.p,
                                      c->list_mode.p, (unsigned long long*)c->list_builds.p, c->sort_skip.p));
            skip = c->sort_skip.p;
            HIPCHK(launch_list_keys(s, n, c->reps.p, c->rep_of.p, c->list_mode.p, c->pos0.p, c->pos_build.p,
                                    c->key_main.p, skip));
        }
        HIPCHK(launch_cell_sort(s, n, c->key_main.p, c->cells_main, c->cnt_main.p, c->start_main.p, c->end_main.p,
                                c->cur_main.p, c->order_tmp.p, c->scan_tmp.p, skip));
        HIPCHK(launch_cell_order(s, n, c->cells_main, c->key_main.p, c->start_main.p, c->end_main.p, c->order_tmp.p,
                                 c->sorted_main.p, skip));
        HIPCHK(launch_build_records(s, n, c->sorted_main.p, c->pos0.p, c->rec_pos.p, c->rec_pos32.p, c->inv_perm.p,
                                    c->reps.p, c->rep_of.p, list ? c->pos_build.p : nullptr, c->list_mode.p,
                                    c->rec_build32.p, c->lcount.p));
        HIPCHK(launch_build_work(s, c->cells_main, c->start_main.p, c->end_main.p, c->work.p, c->nwork.p,
                                 c->work_cap, c->err_flag.p, skip));
    }
    if (list) {
        mds_ctx::Span sp(c, KC_LIST);
        HIPCHK(launch_list_build(s, geometry(c), c->nrep, c->work_cap, c->stencil.p, c->lcount.p, c->loff.p,
                                 c->list.p, c->list_mode.p, c->list_state.p, (long long*)c->list_used.p,
                                 c->err_flag.p));
    }
    if (split) {
        HIPCHK(hipStreamWaitEvent(s, c->ev_join2, 0));
    } else if ((st = contact_chain(c, false)) != MDS_OK) {
        return st;
    }
    {
        mds_ctx::Span sp(c, KC_CONTACTS);
        HIPCHK(launch_mark_moved(s, c->n_moved.p, c->moved_list.p, c->inv_perm.p, c->rec_pos.p));
    }
    c->geometry_ready = true;
    return MDS_OK;
}

// EvalSinCos + the pair geometry of CompInteractVecs for the current
// positions.
mds_status prepare_geometry_once(mds_ctx* c) {
    mds_status st = prepare_positions(c);
    if (st != MDS_OK) return st;
    return prepare_cells(c);
}

// With `check`, the error flags are cleared first and the overlap
// correction's displacement bound is verified before use: a move larger than
// half the contact-search margin could have hidden a contact, so the search
// is redone with a 4x wider margin. Without it, flags accumulate until read.
mds_status prepare_geometry(mds_ctx* c, bool check) {
    for (int attempt = 0;; ++attempt) {
        if (check) HIPCHK(launch_fill_i32(c->stream, c->err_flag.p, 1, 0));
        mds_status st = prepare_geometry_once(c);
        if (st != MDS_OK || !check) return st;
        int32_t f = 0;
        HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if ((f & 128) && attempt <= 4) {
            // a pair list outgrew its region: regions sized from the words
            // the build counted, with a quarter to spare
            std::vector<int64_t> used(c->nrep);
            HIPCHK(hipMemcpy(used.data(), c->list_used.p, c->nrep * sizeof(int64_t), hipMemcpyDeviceToHost));
            for (int r = 0; r < c->nrep; ++r)
                if (used[r] > c->h_rep[r].list_cap)
                    c->meta[r].list_scale *= 1.25 * (double)used[r] / (double)std::max<int64_t>(1, c->h_rep[r].list_cap);
            if ((st = plan_all(c)) != MDS_OK) return st;
            continue;
        }
        double ms = 0;
        for (const RepMeta& m : c->meta) ms = std::max(ms, m.margin_scale);
        if (!(f & 32) || ms >= 2.0 || attempt > 4) return MDS_OK;
        for (RepMeta& m : c->meta) m.margin_scale = std::min(2.0, m.margin_scale * 4);
        if ((st = plan_all(c)) != MDS_OK) return st;
    }
}

// k space of an evaluation (CompDipInteractEwaldF / CompDipFieldEwaldF) for
// the moments in c->mom (their per-replica sums in c->msum): S(k), its chunk
// reduction, then the scatter that writes the field (and force) with the
// self, surface and thermal-force terms; the pair sums are added to it last.
// The chain reads only the sin/cos tables and the moments, so it is forked
// onto the side stream (`forked`) and runs beside the cell and pair kernels
// (in order on one stream when kernels are timed).
mds_status moments_orig(mds_ctx* c, const double* mm_dev, hipStream_t s = nullptr);

// tables_mm (optional): the chain first computes its own inputs on its
// stream, the sin/cos tables and the moments of tables_mm.
mds_status kspace_chain(mds_ctx* c, bool force_pass, bool& forked, const double* tables_mm = nullptr) {
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    const bool forked_here = c->overlap_k && !c->timing;
    auto kspace = [&](hipStream_t ks) -> mds_status {
        if (tables_mm) {
            mds_status st = prepare_sincos(c, ks);
            if (st != MDS_OK) return st;
            if ((st = moments_orig(c, tables_mm, ks)) != MDS_OK) return st;
        }
        if (c->nk_all > 0) {
            {
                mds_ctx::Span sp(c, KC_KBUILD);
                // beside the pair kernels the builds take at most
                // kspace_waves waves per SIMD (a strided grid); in order, all.
                // By default (-1) the matrix-core build holds one wave per
                // SIMD and the vector build is not capped: C3 mixed step
                // 4.32 -> 4.10 ms (256 blocks; 64-192 and 384-512 slower),
                // FP64 step 8.30 -> 8.56 ms if its build were capped too.
                const int ncu = c->ncu;
                const int wm = c->kspace_waves < 0 ? 1 : c->kspace_waves;
                const int wv = c->kspace_waves < 0 ? 0 : c->kspace_waves;
                if (c->any_mfma_build)
                    HIPCHK(launch_kbuild_mfma(ks, n, c->reps.p, c->nrep, c->kbm_items, c->nc_mfma_build, c->rows_m.p,
                                              c->tab32.p, c->mom.p, c->kpartial.p,
                                              forked_here && wm > 0 ? wm * ncu : 0));
                if (c->any_vec_build)
                    HIPCHK(launch_kbuild(ks, n, c->reps.p, c->nrep, c->kb_items, c->segs.p, c->tab.p, c->mom.p,
                                         c->kpartial.p, forked_here && wv > 0 ? 4 * wv * ncu : 0));
            }
            {
                mds_ctx::Span sp(c, KC_KREDUCE);
                HIPCHK(launch_kpack(ks, (int32_t)c->h_pack.size(), c->reps.p, c->nrep, c->pack_desc.p, c->kpartial.p,
                                    c->pack.p, c->pack32.p));
            }
        }
        mds_ctx::Span sp(c, force_pass ? KC_KSCATTER_FORCE : KC_KSCATTER_FIELD);
        if (c->any_mfma_scatter)
            HIPCHK(launch_kscatter_mfma(ks, force_pass, n, c->reps.p, c->nrep, c->ksm_items, c->nc_mfma_scatter,
                                        c->rows_m.p, c->pack32.p, c->tab32.p, c->mom.p, c->msum.p,
                                        force_pass ? c->n_f.p : nullptr, c->force.p, c->field.p));
        if (c->any_vec_scatter)
            HIPCHK(launch_kscatter(ks, force_pass, n, c->reps.p, c->rep_of.p, c->rows.p, c->pack.p, c->tab.p,
                                   c->mom.p, c->msum.p, force_pass ? c->n_f.p : nullptr, c->force.p, c->field.p));
        return MDS_OK;
    };
    forked = c->overlap_k && !c->timing;
    if (forked) {
        HIPCHK(hipEventRecord(c->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        mds_status st = kspace(c->side);
        if (st != MDS_OK) return st;
        HIPCHK(hipEventRecord(c->ev_join, c->side));
        return MDS_OK;
    }
    return kspace(s);
}

// Real space of an evaluation (CompDipInteractEwaldR / CompDipFieldEwaldR,
// with pressure and the short-range terms in the force pass) for the moments
// in c->rec_mom, added to what the k-space chain wrote (joined first when it
// was forked).
mds_status pair_pass(mds_ctx* c, bool force_pass, bool forked) {
    hipStream_t s = c->stream;
    Geometry g = geometry(c);
    PairParams pp = pair_params(c);
    PairOut po{force_pass ? c->force.p : nullptr, c->field.p, c->press_partial.p,
               (unsigned long long*)c->pair_acc.p, c->err_flag.p};
    const size_t moved_press = 3 * (size_t)c->work_cap * PAIR_SLICES;
    mds_ctx::Span sp(c, force_pass ? KC_PAIRS_FORCE : KC_PAIRS_FIELD);
    int64_t* cnt = c->count_work ? (int64_t*)c->pair_counter.p : nullptr;
    if (force_pass) HIPCHK(launch_fill_f64(s, c->press_partial.p + moved_press, 6 * c->nrep, 0.0));
    if (c->list_on())
        HIPCHK(launch_pairs_list(s, force_pass, c->cfg.precision, g, c->rec_mom.p, pp, po, c->work_cap, c->stencil.p,
                                 c->work_cap * PAIR_SLICES, cnt));
    else
        HIPCHK(launch_pairs(s, force_pass, c->cfg.precision, g, c->rec_mom.p, pp, po, c->work_cap, c->stencil.p,
                            c->work_cap * PAIR_SLICES, cnt));
    if (forked) HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
    HIPCHK(launch_pairs_out(s, g, po));
    return MDS_OK;
}

// Real-space + k-space Ewald (+ short range): force_pass
// CompDipInteractEwaldR/F with forces, pressure and the thermal force added,
// else the field-only CompDipFieldEwaldR/F of CompDemagField.
mds_status evaluate(mds_ctx* c, bool force_pass) {
    bool forked = false;
    mds_status st = kspace_chain(c, force_pass, forked);
    if (st != MDS_OK) return st;
    return pair_pass(c, force_pass, forked);
}

// Moments in original order and their per-replica sums (what the k-space
// chain reads), without the cell-sorted copies.
mds_status moments_orig(mds_ctx* c, const double* mm_dev, hipStream_t s) {
    mds_ctx::Span sp(c, KC_MOMENTS);
    if (!s) s = c->stream;
    HIPCHK(launch_moments(s, c->n, mm_dev, c->mag_mom.p, nullptr, c->reps.p, c->rep_of.p, c->mom.p, nullptr,
                          nullptr));
    HIPCHK(launch_rep_sum3(s, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), c->n, c->mom.p,
                           c->red_part.p, c->msum.p));
    return MDS_OK;
}

// The cell-sorted copies (FP64, FP32 in replica units) the pair kernels read.
mds_status moments_sorted(mds_ctx* c, const double* mm_dev) {
    mds_ctx::Span sp(c, KC_MOMENTS);
    HIPCHK(launch_moments(c->stream, c->n, mm_dev, c->mag_mom.p, c->inv_perm.p, c->reps.p, c->rep_of.p, nullptr,
                          c->rec_mom.p, c->rec_mom32.p));
    return MDS_OK;
}

mds_status moments_of(mds_ctx* c, const double* mm_dev) {
    mds_ctx::Span sp(c, KC_MOMENTS);
    HIPCHK(launch_moments(c->stream, c->n, mm_dev, c->mag_mom.p, c->inv_perm.p, c->reps.p, c->rep_of.p, c->mom.p,
                          c->rec_mom.p, c->rec_mom32.p));
    HIPCHK(launch_rep_sum3(c->stream, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), c->n,
                           c->mom.p, c->red_part.p, c->msum.p));
    return MDS_OK;
}

StageParams stage_params(mds_ctx* c, double h) {
    StageParams sp{};
    sp.n = c->n;
    sp.h = h;
    sp.anis_const = c->cfg.anis_const;
    sp.vel_mm_const = c->cfg.vel_mm_const;
    sp.mag_damp = c->cfg.mag_damp;
    sp.mobil = c->cfg.enable_mobilization;
    sp.c1 = c->c1.p;
    sp.c2 = c->c2.p;
    sp.mag_mom = c->mag_mom.p;
    sp.inv_perm = c->inv_perm.p;
    sp.rep_t = c->rep_t.p;
    sp.t_mag = c->cfg.t_mag;
    sp.reps = c->reps.p;
    sp.rep_of = c->rep_of.p;
    sp.rec_mom32 = c->rec_mom32.p;
    return sp;
}

mds_status check_ready(mds_ctx* c) {
    if (!c) return MDS_EINVAL;
    if (c->n <= 0) return fail(c, MDS_EINVAL, "particles not set");
    for (const RepMeta& m : c->meta)
        if (!m.ewald_set) return fail(c, MDS_EINVAL, "Ewald parameters not set");
    if (!c->state_set) return fail(c, MDS_EINVAL, "state not set");
    HIPCHK(hipSetDevice(c->device));
    return ensure_planned(c);
}

mds_status read_flags(mds_ctx* c) {
    int32_t f = 0, nev = 0;
    if (!c->err_flag.p) return MDS_OK;
    HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&nev, c->n_events.p, sizeof(nev), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last_events = nev;
    if (f & 32) return fail(c, MDS_EOVERLAP, "overlap correction moved a particle beyond the contact search margin");
    if (f & 64) return fail(c, MDS_ENONFINITE, "a pair sum is not finite or beyond the fixed-point range");
    if (f & 128) return fail(c, MDS_ECAPACITY, "a pair list exceeded its region");
    if (f) return fail(c, MDS_ECAPACITY, "device capacity exceeded (flag " + std::to_string(f) + ")");
    return MDS_OK;
}

// The step's EvalTransCoeffs observables (src/postprocessing.cpp:167-187),
// after translation and <m_z>, before the step counter advances.
mds_status observables(mds_ctx* c) {
    if (!c->observe) return MDS_OK;
    HIPCHK(launch_observables(c->stream, c->reps.p, c->nrep, c->rep_of.p, c->red_chunks.p, (int32_t)c->h_chunks.size(),
                              c->n, c->pos.p, c->true_pos.p, c->pos_ref.p, c->obs_d2.p, c->red_part.p, c->rep_t.p,
                              c->work.p, c->nwork.p, c->cell_rep.p, c->press_partial.p,
                              3 * (int64_t)c->work_cap * PAIR_SLICES, c->obs_last.p, c->series_counter.p,
                              c->obs_series.p, kSeriesCap));
    return MDS_OK;
}

// The device work of one step of size h after the force evaluation's inputs
// are in place: noise, geometry, force pass, RK stages (one tableau pass, no
// controller), translation, <m_z>, counters. Used for fixed-dt steps and for
// the graph capture.
mds_status fixed_step_body(mds_ctx* c, double h, bool check) {
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t s = c->stream;
    if (c->philox) {
        mds_ctx::Span sp(c, KC_NOISE);
        HIPCHK(launch_noise(s, n, c->reps.p, c->rep_of.p, c->rep_step.p, h, c->zeta_rot.p, c->vol_mag.p,
                            c->zeta_trans.p, c->cfg.kB * c->cfg.temp, c->cfg.mag_damp, c->cfg.gyro_mr,
                            c->cfg.sat_mag, c->cfg.enable_therm_torque, c->cfg.enable_therm_field,
                            c->cfg.enable_therm_force, c->n_t.p, c->n_b.p, c->n_f.p));
    }
    mds_status st;
    if (!check && c->overlap_k && !c->timing) {
        // After the keys, three streams: the contact search (it needs only
        // the keys, and ends before the cell sort and the pair lists), the
        // force pass's k-space chain with its own sin/cos tables and
        // moments (the step-start positions and moments), and the cell sort,
        // lists and pair kernels. The inputs are the same, so the results
        // are those of the in-order step.
        if ((st = prepare_keys(c)) != MDS_OK) return st;
        if ((st = contact_chain(c, true)) != MDS_OK) return st;
        bool forked = false;
        if ((st = kspace_chain(c, true, forked, c->mm.p)) != MDS_OK) return st;
        if ((st = prepare_cells(c, true)) != MDS_OK) return st;
        if ((st = moments_sorted(c, c->mm.p)) != MDS_OK) return st;
        if ((st = pair_pass(c, true, forked)) != MDS_OK) return st;
    } else {
        if ((st = prepare_geometry(c, check)) != MDS_OK) return st;
        if ((st = moments_of(c, c->mm.p)) != MDS_OK) return st;
        if ((st = evaluate(c, true)) != MDS_OK) return st;
    }
    const Tableau tb = make_tableau(MDS_HEUN_EULER);
    const int nb = blocks_for(n);
    const StageParams sp = stage_params(c, h);
    for (int sgi = 0; sgi < tb.stages; ++sgi) {
        const double* ym = sgi == 0 ? c->mm.p : c->pt[(sgi - 1) & 1].p;
        const double* ye = sgi == 0 ? c->ea.p : c->pt[(sgi - 1) & 1].p + n3;
        double* yo = c->pt[sgi & 1].p;
        if (sgi > 0) {
            HIPCHK(launch_rep_sum3(s, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), n, c->mom.p,
                                   c->red_part.p, c->msum.p));
            if ((st = evaluate(c, false)) != MDS_OK) return st;
        }
        mds_ctx::Span spn(c, KC_STAGE);
        const bool last = sgi + 1 == tb.stages;
        HIPCHK(launch_stage(s, sp, sgi, tb.stages, c->mm.p, c->ea.p, ym, ye, c->field.p, c->n_t.p, c->n_b.p, c->kmm.p,
                            c->kea.p, last ? nullptr : tb.a[sgi + 1], tb.b, tb.e, yo, yo + n3, c->mom.p, c->rec_mom.p,
                            c->err_partial.p, c->new_mm.p, c->new_ea.p, nb));
    }
    {
        mds_ctx::Span sp2(c, KC_TRANSLATE);
        HIPCHK(launch_translate(s, n, c->reps.p, c->rep_of.p, c->rep_step.p, h, c->pos_cur.p, c->force.p, c->zeta_trans.p, c->pos.p,
                                c->new_mm.p, c->new_ea.p, c->mm.p, c->ea.p));
        HIPCHK(launch_rep_mean_z(s, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), n, c->mm.p,
                                 c->red_part.p, c->mz_out.p, c->series_counter.p, c->series.p, kSeriesCap));
        if ((st = observables(c)) != MDS_OK) return st;
        HIPCHK(launch_step_count(s, c->nrep, c->reps.p, c->rep_step.p, c->rep_t.p, h, c->series_counter.p));
    }
    return MDS_OK;
}

// One RunSimulation iteration with the adaptive controller (single
// replica). Stages run back to back on the stream; the controller reads the
// error norm (one sync per attempt).
mds_status adaptive_step(mds_ctx* c, double* t, double* dt, double* dt_used) {
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t s = c->stream;
    if (c->philox) {
        mds_ctx::Span sp(c, KC_NOISE);
        HIPCHK(launch_noise(s, n, c->reps.p, c->rep_of.p, c->rep_step.p, *dt, c->zeta_rot.p, c->vol_mag.p,
                            c->zeta_trans.p, c->cfg.kB * c->cfg.temp, c->cfg.mag_damp, c->cfg.gyro_mr,
                            c->cfg.sat_mag, c->cfg.enable_therm_torque, c->cfg.enable_therm_field,
                            c->cfg.enable_therm_force, c->n_t.p, c->n_b.p, c->n_f.p));
    }
    mds_status st = prepare_geometry(c, true);
    if (st != MDS_OK) return st;
    if ((st = moments_of(c, c->mm.p)) != MDS_OK) return st;
    if ((st = evaluate(c, true)) != MDS_OK) return st;

    const Tableau tb = make_tableau(c->cfg.solver);
    const int nb = blocks_for(n);
    bool no_failed = true;
    double h = *dt;
    int tries = 0;
    for (;;) {
        h = *dt;
        const StageParams sp = stage_params(c, h);
        for (int sgi = 0; sgi < tb.stages; ++sgi) {
            const double* ym = sgi == 0 ? c->mm.p : c->pt[(sgi - 1) & 1].p;
            const double* ye = sgi == 0 ? c->ea.p : c->pt[(sgi - 1) & 1].p + n3;
            double* yo = c->pt[sgi & 1].p;
            if (sgi > 0) {
                HIPCHK(launch_rep_sum3(s, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), n, c->mom.p,
                                   c->red_part.p, c->msum.p));
                if ((st = evaluate(c, false)) != MDS_OK) return st;
            }
            mds_ctx::Span spn(c, KC_STAGE);
            const bool last = sgi + 1 == tb.stages;
            HIPCHK(launch_stage(s, sp, sgi, tb.stages, c->mm.p, c->ea.p, ym, ye, c->field.p, c->n_t.p, c->n_b.p,
                                c->kmm.p, c->kea.p, last ? nullptr : tb.a[sgi + 1], tb.b, tb.e, yo, yo + n3, c->mom.p,
                                c->rec_mom.p, c->err_partial.p, c->new_mm.p, c->new_ea.p, nb));
        }
        // AdjustTimeStep (src/simulation.cpp:912-951)
        double err = 0;
        HIPCHK(launch_reduce_max(s, c->err_partial.p, nb, h, c->scal.p));
        HIPCHK(hipMemcpyAsync(&err, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double hn;
        const double tol = c->cfg.abs_tol, dmin = c->cfg.delta_t_min;
        if (err > tol) {
            if (no_failed) {
                no_failed = false;
                hn = std::max(dmin, h * std::max(0.1, 0.9 * std::pow(tol / err, tb.power)));
            } else {
                hn = std::max(dmin, h * 0.5);
            }
            const double f = std::sqrt(hn / h);
            HIPCHK(launch_scale(s, c->n_b.p, n3, f));
            HIPCHK(launch_scale(s, c->n_t.p, n3, f));
            HIPCHK(launch_scale(s, c->n_f.p, n3, f));
            *dt = std::max(hn, dmin);
            if (++tries > 200 || (h <= dmin && hn <= dmin))
                return fail(c, MDS_ENOCONV, "step rejected at deltaTmin (the reference loops forever here)");
            continue;
        }
        hn = no_failed ? h * std::min(std::max(0.9 * std::pow(tol / err, tb.power), 0.2), 5.0) : h;
        *dt = std::max(hn, dmin);  // the min with deltaTmax is overwritten (949-950)
        break;
    }
    {
        mds_ctx::Span sp(c, KC_TRANSLATE);
        // accepted, renormalised state -> mm/ea; IntegrationTrans + wrap; <m_z>
        HIPCHK(launch_translate(s, n, c->reps.p, c->rep_of.p, c->rep_step.p, h, c->pos_cur.p, c->force.p, c->zeta_trans.p, c->pos.p,
                                c->new_mm.p, c->new_ea.p, c->mm.p, c->ea.p));
        HIPCHK(launch_rep_mean_z(s, c->reps.p, c->nrep, c->red_chunks.p, (int32_t)c->h_chunks.size(), n, c->mm.p,
                                 c->red_part.p, c->mz_out.p, c->series_counter.p, c->series.p, kSeriesCap));
        if ((st = observables(c)) != MDS_OK) return st;
        HIPCHK(launch_step_count(s, c->nrep, c->reps.p, c->rep_step.p, c->rep_t.p, h, c->series_counter.p));
    }
    *t += h;
    if (dt_used) *dt_used = h;
    return MDS_OK;
}

// Capture one fixed-dt step (no host synchronisation inside) into a graph.
mds_status capture_step(mds_ctx* c, double h) {
    c->drop_graph();
    const bool timing = c->timing;
    c->timing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        c->timing = timing;
        HIPCHK(e);
    }
    mds_status st = fixed_step_body(c, h, false);
    e = hipStreamEndCapture(c->stream, &graph);
    c->timing = timing;
    if (st != MDS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    HIPCHK(e);
    e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->gexec = nullptr;
        HIPCHK(e);
    }
    c->graph_h = h;
    return MDS_OK;
}

mds_status set_rep_time(mds_ctx* c, int r, double t) {
    HIPCHK(hipMemcpyAsync(c->rep_t.p + r, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}

// Per-particle constants of m particles (InitParticleProps,
// src/simulation.cpp:118-214; velEaConst1/2 at 209-210).
struct Props {
    std::vector<double> rh, vm, mu, zr, zt, c1, c2;
};
mds_status particle_props(mds_ctx* c, int64_t m, const double* r_mag, const double* r_hydr, Props& P) {
    P.rh.resize(m);
    P.vm.resize(m);
    P.mu.resize(m);
    P.zr.resize(m);
    P.zt.resize(m);
    P.c1.resize(m);
    P.c2.resize(m);
    mds_status st = mds_particle_props(&c->cfg, m, r_mag, P.rh.data(), P.vm.data(), P.mu.data(), P.zr.data(),
                                       P.zt.data());
    if (st != MDS_OK) return fail(c, st, "particle properties");
    if (r_hydr) {
        // hydrodynamic radii given: friction coefficients follow them
        for (int64_t i = 0; i < m; ++i) {
            P.rh[i] = r_hydr[i];
            P.zr[i] = 8.0 * c->cfg.my_pi * c->cfg.vis * std::pow(P.rh[i], 3.0);
            P.zt[i] = 6.0 * c->cfg.my_pi * c->cfg.vis * P.rh[i];
        }
    }
    for (int64_t i = 0; i < m; ++i) {
        P.c1[i] = 2.0 * c->cfg.anis_en * (P.vm[i] / P.zr[i]);
        P.c2[i] = 1.0 / P.zr[i];
    }
    return MDS_OK;
}

// Rebuild the kept per-particle arrays for a new particle layout.
mds_status relayout(mds_ctx* c, int64_t old_n, int64_t new_n, const std::vector<Seg>& segs) {
    hipStream_t s = c->stream;
    c->observe = false;  // reference positions no longer match the layout
    const int64_t mn = c->reserve_n;
    DevBuf<double>* one[7] = {&c->r_mag, &c->mag_mom, &c->zeta_rot, &c->zeta_trans, &c->vol_mag, &c->c1, &c->c2};
    for (int k = 0; k < 7; ++k) HIPCHK(repack(*one[k], c->spare[k], 1, old_n, new_n, mn, segs, s));
    DevBuf<double>* three[4] = {&c->pos, &c->mm, &c->ea, &c->pos_build};
    for (int k = 0; k < 4; ++k) HIPCHK(repack(*three[k], c->spare[7 + k], 3, old_n, new_n, mn, segs, s));
    // injected noise does not survive a change of layout
    for (auto* b : {&c->n_t, &c->n_b, &c->n_f}) {
        HIPCHK(b->ensure(3 * (size_t)std::max(new_n, mn)));
        HIPCHK(launch_fill_f64(s, b->p, 3 * new_n, 0.0));
    }
    return MDS_OK;
}

// Per-replica device values (time, step counter) for a new replica list
// (c->nrep is still the old count).
mds_status relayout_rep_values(mds_ctx* c, const std::vector<int>& keep, int added, uint64_t counter) {
    const int old_nrep = c->rep_t.p ? c->nrep : 0;
    std::vector<double> t_old(old_nrep);
    std::vector<uint64_t> s_old(old_nrep), b_old(old_nrep);
    std::vector<int32_t> l_old(old_nrep);
    if (old_nrep > 0) {
        HIPCHK(hipMemcpyAsync(t_old.data(), c->rep_t.p, old_nrep * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(s_old.data(), c->rep_step.p, old_nrep * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(l_old.data(), c->list_state.p, old_nrep * sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(b_old.data(), c->list_builds.p, old_nrep * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    std::vector<double> t_new;
    std::vector<uint64_t> s_new, b_new;
    std::vector<int32_t> l_new;
    for (int k : keep) {
        t_new.push_back(t_old[k]);
        s_new.push_back(s_old[k]);
        l_new.push_back(l_old[k]);
        b_new.push_back(b_old[k]);
    }
    for (int a = 0; a < added; ++a) {
        t_new.push_back(0.0);
        s_new.push_back(counter);
        l_new.push_back(LIST_INVALID);  // no build positions yet
        b_new.push_back(0);
    }
    const size_t m = t_new.size();
    if (c->rep_t.n < m) {
        // grow with headroom; the contents are rewritten below
        HIPCHK(c->rep_t.ensure(2 * m));
        HIPCHK(c->rep_step.ensure(2 * m));
        HIPCHK(c->list_state.ensure(2 * m));
        HIPCHK(c->list_builds.ensure(2 * m));
    }
    if (m > 0) {
        HIPCHK(hipMemcpyAsync(c->rep_t.p, t_new.data(), m * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->rep_step.p, s_new.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->list_state.p, l_new.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->list_builds.p, b_new.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return MDS_OK;
}

void refresh_offsets(mds_ctx* c) {
    int64_t off = 0;
    for (RepMeta& m : c->meta) {
        m.off = off;
        off += m.n;
    }
    c->n = off;
    c->nrep = (int32_t)c->meta.size();
    c->planned = false;
    c->geometry_ready = false;
    c->drop_graph();
}

// Append a replica of m particles (state optional).
// Append `count` replicas of ns[k] particles. Per-particle inputs cover all
// added particles in order (column-major over them for the N x 3 state);
// per-replica inputs have `count` entries (NULL: box from volFrac, the
// configured field, seed 0).
mds_status add_replicas(mds_ctx* c, int32_t count, const int64_t* ns, const double* r_mag, const double* r_hydr,
                        const double* l_box, const double* field_z, const uint64_t* seeds, const double* pos,
                        const double* pos_mm, const double* pos_ea, int32_t* first) {
    if (count <= 0 || !ns || !r_mag) return fail(c, MDS_EINVAL, "bad particle arguments");
    int64_t m = 0;
    for (int32_t k = 0; k < count; ++k) {
        if (ns[k] <= 0) return fail(c, MDS_EINVAL, "empty replica");
        m += ns[k];
    }
    if (c->n + m > ((int64_t)1 << 25)) return fail(c, MDS_EINVAL, "at most 2^25 particles per context");
    if (c->nrep + count > (1 << 20)) return fail(c, MDS_EINVAL, "too many replicas");
    for (int64_t i = 0; i < m; ++i)
        if (!(r_mag[i] > 0) || !std::isfinite(r_mag[i])) return fail(c, MDS_EINVAL, "radii must be positive");
    if (pos || pos_mm || pos_ea) {
        if (!(pos && pos_mm && pos_ea)) return fail(c, MDS_EINVAL, "state needs pos, pos_mm and pos_ea");
        for (int64_t k = 0; k < 3 * m; ++k)
            if (!std::isfinite(pos[k]) || !std::isfinite(pos_mm[k]) || !std::isfinite(pos_ea[k]))
                return fail(c, MDS_ENONFINITE, "non-finite state");
    }
    HIPCHK(hipSetDevice(c->device));
    Props P;
    mds_status st = particle_props(c, m, r_mag, r_hydr, P);
    if (st != MDS_OK) return st;
    std::vector<RepMeta> added;
    int64_t o = 0;
    for (int32_t k = 0; k < count; ++k) {
        const int64_t nk = ns[k];
        RepMeta rm;
        rm.n = nk;
        rm.L = (l_box && l_box[k] > 0) ? l_box[k] : mds_box_length(&c->cfg, nk, P.rh.data() + o);
        rm.field_z = field_z ? field_z[k] : c->cfg.mag_flux_dens;
        rm.seed = seeds ? seeds[k] : 0;
        rm.rmax = *std::max_element(r_mag + o, r_mag + o + nk);
        rm.mref = *std::max_element(P.mu.begin() + o, P.mu.begin() + o + nk);
        if (!(rm.mref > 0)) rm.mref = 1.0;
        st = mds_cutoff_sr(&c->cfg, nk, r_mag + o, P.rh.data() + o, P.mu.data() + o, &rm.rc_sr);
        if (st != MDS_OK) return fail(c, st, "short-range cut-off (InitCutOffSR)");
        added.push_back(rm);
        o += nk;
    }
    const int64_t old_n = c->n, new_n = old_n + m;
    std::vector<Seg> segs{{0, 0, old_n}};
    if ((st = relayout(c, old_n, new_n, segs)) != MDS_OK) return st;
    std::vector<int> keep;
    for (int r = 0; r < c->nrep; ++r) keep.push_back(r);
    if ((st = relayout_rep_values(c, keep, count, 0)) != MDS_OK) return st;
    const double* cols[7] = {r_mag, P.mu.data(), P.zr.data(), P.zt.data(), P.vm.data(), P.c1.data(), P.c2.data()};
    DevBuf<double>* dst[7] = {&c->r_mag, &c->mag_mom, &c->zeta_rot, &c->zeta_trans, &c->vol_mag, &c->c1, &c->c2};
    for (int k = 0; k < 7; ++k)
        HIPCHK(hipMemcpyAsync(dst[k]->p + old_n, cols[k], m * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (pos) {
        if ((st = upload_cols(c, c->pos.p, new_n, old_n, m, pos)) != MDS_OK) return st;
        if ((st = upload_cols(c, c->mm.p, new_n, old_n, m, pos_mm)) != MDS_OK) return st;
        if ((st = upload_cols(c, c->ea.p, new_n, old_n, m, pos_ea)) != MDS_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->h_r_mag.insert(c->h_r_mag.end(), r_mag, r_mag + m);
    c->h_mag_mom.insert(c->h_mag_mom.end(), P.mu.begin(), P.mu.end());
    c->h_r_hydr.insert(c->h_r_hydr.end(), P.rh.begin(), P.rh.end());
    if (first) *first = c->nrep;
    c->meta.insert(c->meta.end(), added.begin(), added.end());
    refresh_offsets(c);
    return MDS_OK;
}

// r_c / L of replica m for the autotuner's candidate f: f itself for one
// replica; with several, f is the value for a replica of the mean size and
// each replica's is scaled by (N_mean / N)^(1/6), the optimum of a cost
// ~ N rho r_c^3 (real space) + N (L / r_c)^3 (k space) at fixed accuracy,
// kept in [0.1, 0.5].
double rc_frac_of(const mds_ctx* c, double f, const RepMeta& m) {
    if (c->nrep <= 1) return f;
    const double mean = (double)c->n / (double)c->nrep;
    return std::min(0.5, std::max(0.1, f * std::pow(mean / (double)m.n, 1.0 / 6.0)));
}

mds_status set_rep_ewald(mds_ctx* c, int r, double alpha, double r_cut, int32_t nc) {
    RepMeta& m = c->meta[r];
    if (!(alpha > 0) || !(r_cut > 0) || nc < 0 || nc > 64) return fail(c, MDS_EINVAL, "bad Ewald parameters");
    if (r_cut > 0.5 * m.L * (1 + 1e-12)) return fail(c, MDS_EINVAL, "r_c above L/2 breaks the minimum image");
    m.alpha = alpha;
    m.rc = r_cut;
    m.nc = nc;
    m.ewald_set = true;
    c->planned = false;
    c->geometry_ready = false;
    c->drop_graph();
    return MDS_OK;
}

}  // namespace

extern "C" {

int mds_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

mds_status mds_create(const mds_config* cfg, int device, mds_ctx** out) {
    if (!cfg || !out) return MDS_EINVAL;
    *out = nullptr;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || device < 0 || device >= cnt) return MDS_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) return MDS_EDEVICE;
    if (cfg->solver < 0 || cfg->solver > 2) return MDS_ECONFIG;
    if (cfg->precision != MDS_PREC_FP64 && cfg->precision != MDS_PREC_MIXED) return MDS_EINVAL;
    mds_ctx* c = new mds_ctx();
    c->cfg = *cfg;
    c->device = device;
    c->ncu = std::max(1, mds_device_cus(device));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return MDS_EDEVICE;
    }
    if (c->pair_counter.ensure(1) != hipSuccess || c->series_counter.ensure(1) != hipSuccess ||
        c->scal.ensure(8) != hipSuccess || c->n_events.ensure(1) != hipSuccess || c->err_flag.ensure(1) != hipSuccess ||
        c->max_disp.ensure(1) != hipSuccess || c->nwork.ensure(1) != hipSuccess || c->n_moved.ensure(1) != hipSuccess) {
        delete c;
        return MDS_ENOMEM;
    }
    (void)hipMemset(c->pair_counter.p, 0, sizeof(uint64_t));
    (void)hipMemset(c->series_counter.p, 0, sizeof(int32_t));
    (void)hipMemset(c->err_flag.p, 0, sizeof(int32_t));
    (void)hipMemset(c->n_events.p, 0, sizeof(int32_t));
    (void)hipMemset(c->max_disp.p, 0, sizeof(double));
    (void)hipMemset(c->n_moved.p, 0, sizeof(int32_t));
    *out = c;
    return MDS_OK;
}

// Drop every replica (the buffers stay allocated for reuse).
static void clear_replicas(mds_ctx* c) {
    c->observe = false;
    c->meta.clear();
    c->h_r_mag.clear();
    c->h_mag_mom.clear();
    c->h_r_hydr.clear();
    c->rep_t.retire();
    c->rep_step.retire();
    c->n = 0;
    c->nrep = 0;
}

void mds_destroy(mds_ctx* c) {
    if (!c) return;
    const int dev = c->device;
    (void)hipSetDevice(dev);
    delete c;
    free_retired(dev);
}

const char* mds_last_error(const mds_ctx* c) { return c ? c->err.c_str() : "null context"; }

mds_status mds_set_particles(mds_ctx* c, int64_t n, const double* r_mag, const double* r_hydr, double l_box) {
    if (!c || n <= 0 || !r_mag) return fail(c, MDS_EINVAL, "bad particle arguments");
    if (n > (int64_t)1 << 25) return fail(c, MDS_EINVAL, "at most 2^25 particles per context");
    HIPCHK(hipSetDevice(c->device));
    // a single replica: drop any previous ones
    const uint64_t seed = c->meta.empty() ? 0 : c->meta[0].seed;
    clear_replicas(c);
    const double lb = l_box;
    mds_status st = add_replicas(c, 1, &n, r_mag, r_hydr, &lb, nullptr, &seed, nullptr, nullptr, nullptr, nullptr);
    if (st != MDS_OK) return st;
    c->state_set = false;
    return MDS_OK;
}

mds_status mds_set_ewald(mds_ctx* c, double alpha, double r_cut, int32_t nc) {
    if (!c || c->n <= 0) return fail(c, MDS_EINVAL, "set particles first");
    for (int r = 0; r < c->nrep; ++r) {
        mds_status st = set_rep_ewald(c, r, alpha, r_cut, nc);
        if (st != MDS_OK) return st;
    }
    HIPCHK(hipSetDevice(c->device));
    return ensure_planned(c);
}

mds_status mds_set_tiling(mds_ctx* c, int32_t cell_reach) {
    if (!c || cell_reach < 0 || cell_reach > 8) return fail(c, MDS_EINVAL, "cell reach must be 0..8");
    c->reach = cell_reach;
    c->planned = false;
    c->geometry_ready = false;
    c->drop_graph();
    return MDS_OK;
}

mds_status mds_set_pair_list(mds_ctx* c, double skin, double region_scale) {
    if (!c || !(skin >= 0) || skin > 0.25) return fail(c, MDS_EINVAL, "pair-list skin must be in [0, 0.25] of r_c");
    if (!(region_scale > 0) || region_scale > 64) return fail(c, MDS_EINVAL, "pair-list region scale must be in (0, 64]");
    HIPCHK(hipSetDevice(c->device));
    c->list_skin = skin;
    for (RepMeta& m : c->meta) m.list_scale = region_scale;
    c->planned = false;
    c->geometry_ready = false;
    c->drop_graph();
    // every list is built afresh from the current positions
    if (c->nrep > 0) {
        HIPCHK(launch_fill_i32(c->stream, c->list_state.p, c->nrep, LIST_INVALID));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return MDS_OK;
}

int64_t mds_pair_list_words(int64_t n, double l_box, double r_cut, double skin) {
    if (n <= 0 || !(l_box > 0) || !(r_cut > 0) || !(skin > 0)) return 0;
    RepMeta m;
    m.n = n;
    m.L = l_box;
    m.rc = r_cut;
    return list_capacity(m, r_cut * (1.0 + skin));
}

mds_status mds_pair_list_stats(mds_ctx* c, int32_t index, int64_t* builds, int64_t* words, int64_t* capacity) {
    if (!c || index < 0 || index >= c->nrep) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    uint64_t b = 0;
    int64_t w = 0;
    if (c->list_on() && c->planned) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(&b, c->list_builds.p + index, sizeof(b), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&w, c->list_used.p + index, sizeof(w), hipMemcpyDeviceToHost));
    } else if (c->list_on()) {
        HIPCHK(hipMemcpy(&b, c->list_builds.p + index, sizeof(b), hipMemcpyDeviceToHost));
    }
    if (builds) *builds = (int64_t)b;
    if (words) *words = w;
    if (capacity) *capacity = (c->list_on() && c->planned) ? c->h_rep[index].list_cap : 0;
    return MDS_OK;
}

mds_status mds_get_tiling(const mds_ctx* c, int32_t* cell_reach, int32_t* ncell, int32_t* nstencil) {
    if (!c) return MDS_EINVAL;
    if (cell_reach) *cell_reach = c->reach;
    if (!c->planned) {
        // planned lazily: report the last plan if any
        if (ncell) *ncell = c->h_rep.empty() ? 0 : c->h_rep[0].ncell;
        if (nstencil) *nstencil = c->h_rep.empty() ? 0 : c->h_rep[0].nsten;
        return MDS_OK;
    }
    if (ncell) *ncell = c->h_rep.empty() ? 0 : c->h_rep[0].ncell;
    if (nstencil) *nstencil = c->h_rep.empty() ? 0 : c->h_rep[0].nsten;
    return MDS_OK;
}

mds_status mds_autotune(mds_ctx* c, const double* rc_fracs, int32_t n_rc, const int32_t* reaches, int32_t n_reach,
                        int32_t reps, double* best_rc_frac, int32_t* best_reach, double* best_ms) {
    // CompEwaldParams (src/interaction.cpp:164-324) picks r_c in [0.2, 0.5] L
    // by timing the field sums on the CPU; here every candidate (r_c, cell
    // reach) is timed on the device over the step's geometry (cell lists,
    // contact search, tables: their cost depends on the tiling) and one force
    // and one field evaluation, with alpha and k_c of every replica from the
    // same error model, so each candidate has the accuracy the configuration
    // asks for.
    if (!c || c->n <= 0 || !c->state_set || !rc_fracs || n_rc <= 0 || !reaches || n_reach <= 0)
        return fail(c, MDS_EINVAL, "autotune needs particles, state and candidates");
    HIPCHK(hipSetDevice(c->device));
    if (reps < 1) reps = 1;
    double best = 1e300, brc = -1;
    int32_t breach = c->reach;
    std::vector<double> b_alpha(c->nrep), cur_alpha(c->nrep);
    std::vector<int32_t> b_nc(c->nrep), cur_nc(c->nrep);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool timing = c->timing;
    c->timing = false;
    mds_status st = MDS_OK;
    for (int a = 0; a < n_rc && st == MDS_OK; ++a) {
        bool ok = true;
        for (int r = 0; r < c->nrep && ok; ++r) {
            const RepMeta& m = c->meta[r];
            const double rc = rc_frac_of(c, rc_fracs[a], m) * m.L;
            double alpha = 0, kc = 0;
            if (mds_ewald_params_for_rc(&c->cfg, m.n, c->h_mag_mom.data() + m.off, c->h_r_hydr.data() + m.off, m.L,
                                        rc, &alpha, &kc) != MDS_OK) {
                ok = false;
                break;
            }
            const int32_t nc = (int32_t)std::floor(kc + 0.5);  // src/interaction.cpp:304
            if (nc > 64 || rc > 0.5 * m.L) ok = false;
            cur_alpha[r] = alpha;
            cur_nc[r] = nc;
        }
        if (!ok) continue;
        for (int r = 0; r < c->nrep && st == MDS_OK; ++r)
            st = set_rep_ewald(c, r, cur_alpha[r], rc_frac_of(c, rc_fracs[a], c->meta[r]) * c->meta[r].L, cur_nc[r]);
        for (int b = 0; b < n_reach && st == MDS_OK; ++b) {
            c->reach = reaches[b];
            c->planned = false;
            if ((st = ensure_planned(c)) != MDS_OK) break;
            if ((st = prepare_geometry(c, true)) != MDS_OK) break;
            float best_ms_here = 1e30f;
            for (int r = 0; r < reps + 1 && st == MDS_OK; ++r) {
                HIPCHK(hipEventRecord(e0, c->stream));
                if ((st = prepare_geometry(c, false)) != MDS_OK) break;
                if ((st = moments_of(c, c->mm.p)) != MDS_OK) break;
                if ((st = evaluate(c, true)) != MDS_OK) break;
                if ((st = evaluate(c, false)) != MDS_OK) break;
                HIPCHK(hipEventRecord(e1, c->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) best_ms_here = std::min(best_ms_here, ms);  // the first run warms up
            }
            if (best_ms_here < best) {
                best = best_ms_here;
                brc = rc_fracs[a];
                breach = reaches[b];
                b_alpha = cur_alpha;
                b_nc = cur_nc;
            }
        }
    }
    c->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != MDS_OK) return st;
    if (brc < 0) return fail(c, MDS_ENOCONV, "no admissible Ewald candidate");
    c->reach = breach;
    for (int r = 0; r < c->nrep; ++r)
        if ((st = set_rep_ewald(c, r, b_alpha[r], rc_frac_of(c, brc, c->meta[r]) * c->meta[r].L, b_nc[r])) != MDS_OK)
            return st;
    if ((st = ensure_planned(c)) != MDS_OK) return st;
    if (best_rc_frac) *best_rc_frac = brc;
    if (best_reach) *best_reach = breach;
    if (best_ms) *best_ms = best;
    return MDS_OK;
}

mds_status mds_get_params(const mds_ctx* c, double* l_box, double* r_cut_sr, double* alpha, double* r_cut,
                          int32_t* nc) {
    if (!c || c->meta.empty()) return MDS_EINVAL;
    const RepMeta& m = c->meta[0];
    if (l_box) *l_box = m.L;
    if (r_cut_sr) *r_cut_sr = m.rc_sr;
    if (alpha) *alpha = m.alpha;
    if (r_cut) *r_cut = m.rc;
    if (nc) *nc = m.nc;
    return MDS_OK;
}

mds_status mds_set_state(mds_ctx* c, const double* pos, const double* pos_mm, const double* pos_ea) {
    if (!c || c->n <= 0 || !pos || !pos_mm || !pos_ea) return fail(c, MDS_EINVAL, "bad state");
    const size_t n3 = 3 * (size_t)c->n;
    for (size_t k = 0; k < n3; ++k)
        if (!std::isfinite(pos[k]) || !std::isfinite(pos_mm[k]) || !std::isfinite(pos_ea[k]))
            return fail(c, MDS_ENONFINITE, "non-finite state");
    HIPCHK(hipSetDevice(c->device));
    mds_status st;
    if ((st = upload(c, c->pos.p, pos, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->mm.p, pos_mm, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->ea.p, pos_ea, n3)) != MDS_OK) return st;
    // new positions: every list is built afresh
    if (c->nrep > 0) HIPCHK(launch_fill_i32(c->stream, c->list_state.p, c->nrep, LIST_INVALID));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->state_set = true;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status mds_get_state(mds_ctx* c, double* pos, double* pos_mm, double* pos_ea) {
    if (!c || c->n <= 0) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const size_t n3 = 3 * (size_t)c->n;
    mds_status st;
    if (pos && (st = download(c, pos, c->pos.p, n3)) != MDS_OK) return st;
    if (pos_mm && (st = download(c, pos_mm, c->mm.p, n3)) != MDS_OK) return st;
    if (pos_ea && (st = download(c, pos_ea, c->ea.p, n3)) != MDS_OK) return st;
    return MDS_OK;
}

mds_status mds_set_noise(mds_ctx* c, const double* tt, const double* tb, const double* tf) {
    if (!c || c->n <= 0) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const size_t n3 = 3 * (size_t)c->n;
    mds_status st;
    double* dst[3] = {c->n_t.p, c->n_b.p, c->n_f.p};
    const double* src[3] = {tt, tb, tf};
    for (int k = 0; k < 3; ++k) {
        if (src[k]) {
            if ((st = upload(c, dst[k], src[k], n3)) != MDS_OK) return st;
        } else {
            HIPCHK(launch_fill_f64(c->stream, dst[k], n3, 0.0));
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->philox = false;
    c->drop_graph();
    return MDS_OK;
}

mds_status mds_use_philox(mds_ctx* c, uint64_t seed, uint64_t counter) {
    if (!c || c->nrep <= 0) return fail(c, MDS_EINVAL, "set particles first");
    HIPCHK(hipSetDevice(c->device));
    c->philox = true;
    if (c->nrep == 1) {
        c->meta[0].seed = seed;
        c->planned = false;
    }
    std::vector<uint64_t> v(c->nrep, counter);
    HIPCHK(hipMemcpyAsync(c->rep_step.p, v.data(), v.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->drop_graph();
    return MDS_OK;
}

mds_status mds_mean_z_noalloc(const double* pos_mm, int64_t n, double* mz, void* stream) {
    if (!pos_mm || !mz || n <= 0) return MDS_EINVAL;
    if (n > NOALLOC_MAX_PARTICLES) return MDS_ECAPACITY;
    return launch_mean_z_noalloc((hipStream_t)stream, n, pos_mm, mz) == hipSuccess ? MDS_OK : MDS_EDEVICE;
}

mds_status mds_get_noise(mds_ctx* c, double* tt, double* tb, double* tf) {
    if (!c || c->n <= 0) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n3 = 3 * (size_t)c->n;
    mds_status st;
    if (tt && (st = download(c, tt, c->n_t.p, n3)) != MDS_OK) return st;
    if (tb && (st = download(c, tb, c->n_b.p, n3)) != MDS_OK) return st;
    if (tf && (st = download(c, tf, c->n_f.p, n3)) != MDS_OK) return st;
    return MDS_OK;
}

void mds_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = c[k];
}

void mds_philox_normals(uint64_t seed, int64_t index, uint64_t step, double* z9) {
    double z[10];
    philox_normals(seed, index, step, z);
    for (int k = 0; k < 9; ++k) z9[k] = z[k];
}

mds_status mds_comp_interactions(mds_ctx* c, double* force, double* field, double* press, double* pos_after) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if ((st = prepare_geometry(c, true)) != MDS_OK) return st;
    if ((st = moments_of(c, c->mm.p)) != MDS_OK) return st;
    if ((st = evaluate(c, true)) != MDS_OK) return st;
    if ((st = read_flags(c)) != MDS_OK) return st;
    const size_t n3 = 3 * (size_t)c->n;
    if (force && (st = download(c, force, c->force.p, n3)) != MDS_OK) return st;
    if (field && (st = download(c, field, c->field.p, n3)) != MDS_OK) return st;
    if (pos_after && (st = download(c, pos_after, c->pos_cur.p, n3)) != MDS_OK) return st;
    if (press) {
        // pressure sums over all replicas (one replica: the reference's)
        const size_t slots = (size_t)c->work_cap * PAIR_SLICES;
        std::vector<double> pp(3 * slots);
        int32_t nw = 0;
        HIPCHK(hipMemcpy(&nw, c->nwork.p, sizeof(nw), hipMemcpyDeviceToHost));
        const size_t used = (size_t)nw * PAIR_SLICES;
        HIPCHK(hipMemcpy(pp.data(), c->press_partial.p, 3 * used * sizeof(double), hipMemcpyDeviceToHost));
        press[0] = press[1] = press[2] = 0;
        for (size_t w = 0; w < used; ++w)
            for (int d = 0; d < 3; ++d) press[d] += pp[3 * w + d];
        // moved pairs: two fixed-point words per component and replica
        std::vector<long long> pm(6 * (size_t)c->nrep);
        HIPCHK(hipMemcpy(pm.data(), c->press_partial.p + 3 * slots, pm.size() * sizeof(long long),
                         hipMemcpyDeviceToHost));
        for (int32_t r = 0; r < c->nrep; ++r) {
            const double u = c->h_rep[r].px_s;
            for (int d = 0; d < 3; ++d)
                press[d] += std::fma((double)pm[6 * (size_t)r + 2 * d], u,
                                     (double)pm[6 * (size_t)r + 2 * d + 1] * (u * (1.0 / 1099511627776.0)));
        }
    }
    return MDS_OK;
}

mds_status mds_comp_demag_field(mds_ctx* c, const double* pos_mm, double* field) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if (!c->geometry_ready) return fail(c, MDS_EINVAL, "call mds_comp_interactions first");
    if (!pos_mm || !field) return MDS_EINVAL;
    const size_t n3 = 3 * (size_t)c->n;
    if ((st = upload(c, c->pt[0].p, pos_mm, n3)) != MDS_OK) return st;
    if ((st = moments_of(c, c->pt[0].p)) != MDS_OK) return st;
    if ((st = evaluate(c, false)) != MDS_OK) return st;
    return download(c, field, c->field.p, n3);
}

mds_status mds_rk_coeffs(mds_ctx* c, double t, const double* pos_mm, const double* pos_ea, const double* field,
                         double* k_mm, double* k_ea) {
    if (!c || c->n <= 0 || !pos_mm || !pos_ea || !field) return MDS_EINVAL;
    if (c->nrep != 1) return fail(c, MDS_EINVAL, "mds_rk_coeffs takes a single replica");
    HIPCHK(hipSetDevice(c->device));
    mds_status st = size_scratch(c);
    if (st != MDS_OK) return st;
    if (!c->planned) {
        // the stage kernel needs the replica table (field, time), not the plan
        Rep R{};
        R.off = 0;
        R.n = c->n;
        R.L = c->meta[0].L;
        R.field_z = c->meta[0].field_z;
        HIPCHK(c->reps.ensure(1));
        HIPCHK(c->rep_of.ensure(c->n));
        HIPCHK(hipMemcpy(c->reps.p, &R, sizeof(Rep), hipMemcpyHostToDevice));
        HIPCHK(launch_fill_i32(c->stream, c->rep_of.p, c->n, 0));
    }
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    if ((st = set_rep_time(c, 0, t)) != MDS_OK) return st;
    if ((st = upload(c, c->pt[0].p, pos_mm, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->pt[0].p + n3, pos_ea, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->field.p, field, n3)) != MDS_OK) return st;
    const StageParams sp = stage_params(c, 0.0);
    const double zero[8] = {0};
    // a stage of a one-stage tableau: k only (the accepted-state outputs go to scratch)
    HIPCHK(launch_stage(c->stream, sp, 0, 1, c->pt[0].p, c->pt[0].p + n3, c->pt[0].p, c->pt[0].p + n3, c->field.p,
                        c->n_t.p, c->n_b.p, c->kmm.p, c->kea.p, nullptr, zero, zero, c->pt[1].p, c->pt[1].p + n3,
                        c->mom.p, c->rec_mom.p, c->err_partial.p, c->new_mm.p, c->new_ea.p, blocks_for(n)));
    if (k_mm && (st = download(c, k_mm, c->kmm.p, n3)) != MDS_OK) return st;
    if (k_ea && (st = download(c, k_ea, c->kea.p, n3)) != MDS_OK) return st;
    return MDS_OK;
}

mds_status mds_step(mds_ctx* c, double* t, double* dt, double* dt_used, double* mz) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if (!t || !dt) return MDS_EINVAL;
    if (c->nrep == 1 && (st = set_rep_time(c, 0, *t)) != MDS_OK) return st;
    if (c->cfg.fixed_dt) {
        HIPCHK(launch_fill_i32(c->stream, c->series_counter.p, 1, 0));
        if ((st = fixed_step_body(c, *dt, true)) != MDS_OK) return st;
        *t += *dt;
        if (dt_used) *dt_used = *dt;
    } else {
        if (c->nrep != 1) return fail(c, MDS_EINVAL, "adaptive time stepping takes a single replica");
        if ((st = adaptive_step(c, t, dt, dt_used)) != MDS_OK) return st;
    }
    if ((st = read_flags(c)) != MDS_OK) return st;
    if (mz) HIPCHK(hipMemcpy(mz, c->mz_out.p, sizeof(double), hipMemcpyDeviceToHost));
    return MDS_OK;
}

namespace {
// Save (restore = false) or restore the state a graph-replayed chunk starts
// from: positions, moments, easy axes, the unwrapped positions of the
// observables, and each replica's clock and step counter.
mds_status snapshot(mds_ctx* c, bool restore) {
    const size_t n3 = 3 * (size_t)c->n, nr = (size_t)c->nrep;
    const size_t need = 5 * n3;
    if (!restore) {
        HIPCHK(c->snap.ensure(need));
        HIPCHK(c->snap_t.ensure(nr));
        HIPCHK(c->snap_step.ensure(nr));
        HIPCHK(c->snap_list.ensure(nr));
    }
    // the pair lists' build positions and states: a restored chunk's lists
    // are rebuilt from the build positions it started with
    double* live[5] = {c->pos.p, c->mm.p, c->ea.p, c->observe ? c->true_pos.p : nullptr,
                       c->list_on() ? c->pos_build.p : nullptr};
    for (int k = 0; k < 5; ++k) {
        if (!live[k]) continue;
        double* sp = c->snap.p + k * n3;
        HIPCHK(hipMemcpyAsync(restore ? live[k] : sp, restore ? sp : live[k], n3 * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipMemcpyAsync(restore ? c->rep_t.p : c->snap_t.p, restore ? c->snap_t.p : c->rep_t.p, nr * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(restore ? c->rep_step.p : c->snap_step.p, restore ? c->snap_step.p : c->rep_step.p,
                          nr * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    if (c->list_on()) {
        HIPCHK(hipMemcpyAsync(restore ? c->list_state.p : c->snap_list.p, restore ? c->snap_list.p : c->list_state.p,
                              nr * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
        if (restore) return stale_lists(c);
    }
    return MDS_OK;
}

mds_status run_impl(mds_ctx* c, int64_t steps, double* t, double* dt, double* mz_series, double* obs) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if (!t || !dt || steps < 0) return MDS_EINVAL;
    if (steps == 0) return MDS_OK;
    if (obs && !c->observe) return fail(c, MDS_EINVAL, "call mds_observe first");
    if (c->nrep == 1 && (st = set_rep_time(c, 0, *t)) != MDS_OK) return st;
    const int32_t R = c->nrep;
    if (!c->cfg.fixed_dt) {
        if (R != 1) return fail(c, MDS_EINVAL, "adaptive time stepping takes a single replica");
        for (int64_t k = 0; k < steps; ++k) {
            if ((st = adaptive_step(c, t, dt, nullptr)) != MDS_OK) return st;
            if (mz_series)
                HIPCHK(hipMemcpyAsync(mz_series + k, c->mz_out.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            if (obs)
                HIPCHK(hipMemcpyAsync(obs + 7 * k, c->obs_last.p, 7 * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        return read_flags(c);
    }
    const double h = *dt;
    // Steps in chunks of at most check_every (and of the device <m_z>
    // buffer). A chunk's first step checks the overlap margin (and clears the
    // flags); later ones replay a graph when enabled, unchecked, so the flags
    // are read after the chunk. A chunk that raised the overlap-margin or a
    // capacity flag is rolled back to its start and redone eagerly, every
    // step checked (a wider margin where a move needs it); its Philox draws
    // are functions of the restored step counters, so the redone chunk is the
    // run that was checked throughout. A step that still raises a flag ends
    // the run with the state it started from.
    int64_t done = 0;
    while (done < steps) {
        const int64_t m = std::min<int64_t>(std::min<int64_t>(kSeriesCap, c->check_every), steps - done);
        const bool guard = c->use_graph && m > 1;
        if (guard && (st = snapshot(c, false)) != MDS_OK) return st;
        HIPCHK(launch_fill_i32(c->stream, c->series_counter.p, 1, 0));
        for (int64_t k = 0; k < m; ++k) {
            if (k == 0 || !c->use_graph) {
                if ((st = fixed_step_body(c, h, k == 0)) != MDS_OK) return st;
            } else {
                if (!c->gexec || c->graph_h != h) {
                    if ((st = capture_step(c, h)) != MDS_OK) return st;
                }
                HIPCHK(hipGraphLaunch(c->gexec, c->stream));
            }
        }
        if (guard) {
            int32_t f = 0;
            HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (f & 64) return read_flags(c);
            if (f) {
                ++c->chunks_redone;
                if ((st = snapshot(c, true)) != MDS_OK) return st;
                c->drop_graph();
                HIPCHK(launch_fill_i32(c->stream, c->series_counter.p, 1, 0));
                for (int64_t k = 0; k < m; ++k) {
                    if ((st = snapshot(c, false)) != MDS_OK) return st;
                    if ((st = fixed_step_body(c, h, true)) != MDS_OK) return st;
                    HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
                    HIPCHK(hipStreamSynchronize(c->stream));
                    if (f) {
                        st = read_flags(c);
                        mds_status rs = snapshot(c, true);
                        HIPCHK(hipStreamSynchronize(c->stream));
                        for (int64_t q = 0; q < done + k; ++q) *t += h;
                        if (rs != MDS_OK) return rs;
                        return fail(c, st, std::string(mds_last_error(c)) + " at step " + std::to_string(done + k) +
                                               " of the run (the state is the one that step started from)");
                    }
                }
            }
        }
        if (mz_series)
            HIPCHK(hipMemcpyAsync(mz_series + done * R, c->series.p, (size_t)m * R * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
        if (obs)
            HIPCHK(hipMemcpyAsync(obs + done * R * 7, c->obs_series.p, (size_t)m * R * 7 * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
        if (mz_series || obs) HIPCHK(hipStreamSynchronize(c->stream));
        done += m;
    }
    for (int64_t k = 0; k < steps; ++k) *t += h;  // as tInt accumulates (src/simulation.cpp:518)
    return read_flags(c);
}
}  // namespace

mds_status mds_replica_stop_at(mds_ctx* c, int32_t index, uint64_t step) {
    if (!c || index < 0 || index >= c->nrep) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    c->meta[index].stop_at = step;
    if (c->planned && (size_t)index < c->h_rep.size()) {
        c->h_rep[index].stop_at = step;
        HIPCHK(hipMemcpyAsync(c->reps.p + index, &c->h_rep[index], sizeof(Rep), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return MDS_OK;
}

mds_status mds_set_check_interval(mds_ctx* c, int64_t steps) {
    if (!c || steps < 1) return MDS_EINVAL;
    c->check_every = steps;
    return MDS_OK;
}

int64_t mds_chunks_redone(const mds_ctx* c) { return c ? c->chunks_redone : 0; }

mds_status mds_run(mds_ctx* c, int64_t steps, double* t, double* dt, double* mz_series) {
    return run_impl(c, steps, t, dt, mz_series, nullptr);
}

mds_status mds_run_observed(mds_ctx* c, int64_t steps, double* t, double* dt, double* mz_series, double* obs) {
    if (!obs) return MDS_EINVAL;
    return run_impl(c, steps, t, dt, mz_series, obs);
}

mds_status mds_observe(mds_ctx* c, int32_t on) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    c->drop_graph();
    c->observe = false;
    if (!on) return MDS_OK;
    if ((st = ensure_planned(c)) != MDS_OK) return st;
    const size_t n3 = 3 * (size_t)c->n;
    for (auto* b : {&c->true_pos, &c->pos_ref, &c->obs_d2}) HIPCHK(b->ensure(n3));
    HIPCHK(c->obs_series.ensure((size_t)kSeriesCap * c->nrep * 7));
    HIPCHK(c->obs_last.ensure((size_t)c->nrep * 7));
    // positions at this call: the reference and the first true positions
    // (src/simulation.cpp:108-109)
    HIPCHK(hipMemcpyAsync(c->pos_ref.p, c->pos.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->true_pos.p, c->pos.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(launch_fill_f64(c->stream, c->obs_last.p, (int64_t)c->nrep * 7, 0.0));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->observe = true;
    return MDS_OK;
}

mds_status mds_last_observables(mds_ctx* c, double* obs) {
    if (!c || !obs) return MDS_EINVAL;
    if (!c->observe) return fail(c, MDS_EINVAL, "call mds_observe first");
    HIPCHK(hipMemcpyAsync(obs, c->obs_last.p, (size_t)c->nrep * 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}

mds_status mds_synchronize(mds_ctx* c) {
    if (!c) return MDS_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    return read_flags(c);
}

mds_status mds_enable_timing(mds_ctx* c, int32_t on) {
    if (!c) return MDS_EINVAL;
    c->collect();
    c->timing = on != 0;
    c->count_work = on != 0;
    for (int k = 0; k < KC_COUNT; ++k) {
        c->launches[k] = 0;
        c->total_ms[k] = 0;
    }
    if (c->pair_counter.p) HIPCHK(hipMemset(c->pair_counter.p, 0, sizeof(uint64_t)));
    c->drop_graph();
    return MDS_OK;
}

int32_t mds_kernel_class_count(void) { return KC_COUNT; }
const char* mds_kernel_class_name(int32_t id) { return (id >= 0 && id < KC_COUNT) ? kKClassNames[id] : ""; }

mds_status mds_kernel_stats(mds_ctx* c, int64_t* launches, double* total_ms) {
    if (!c) return MDS_EINVAL;
    c->collect();
    for (int k = 0; k < KC_COUNT; ++k) {
        if (launches) launches[k] = c->launches[k];
        if (total_ms) total_ms[k] = c->total_ms[k];
    }
    return MDS_OK;
}

mds_status mds_work_counts(mds_ctx* c, int64_t* pairs, int64_t* candidates, int64_t* kvecs, int64_t* overlap_events) {
    if (!c) return MDS_EINVAL;
    uint64_t p = 0;
    if (c->pair_counter.p) HIPCHK(hipMemcpy(&p, c->pair_counter.p, sizeof(p), hipMemcpyDeviceToHost));
    int32_t nev = 0;
    if (c->n_events.p) HIPCHK(hipMemcpy(&nev, c->n_events.p, sizeof(nev), hipMemcpyDeviceToHost));
    if (pairs) *pairs = (int64_t)p;
    if (candidates) *candidates = 0;
    if (kvecs) *kvecs = c->nk_all;
    if (overlap_events) *overlap_events = nev / 2;
    return MDS_OK;
}

// ------------------------------------------------------------------ replicas
mds_status mds_add_replica(mds_ctx* c, int64_t n, const double* r_mag, const double* r_hydr, double l_box,
                           double field_z, uint64_t seed, const double* pos, const double* pos_mm,
                           const double* pos_ea, int32_t* index) {
    if (!c) return MDS_EINVAL;
    const bool had_state = c->state_set || c->nrep == 0;
    mds_status st = add_replicas(c, 1, &n, r_mag, r_hydr, &l_box, &field_z, &seed, pos, pos_mm, pos_ea, index);
    if (st != MDS_OK) return st;
    c->state_set = had_state && pos != nullptr;
    return MDS_OK;
}

mds_status mds_add_replicas(mds_ctx* c, int32_t count, const int64_t* n, const double* r_mag, const double* r_hydr,
                            const double* l_box, const double* field_z, const uint64_t* seeds, const double* pos,
                            const double* pos_mm, const double* pos_ea, int32_t* first_index) {
    if (!c) return MDS_EINVAL;
    const bool had_state = c->state_set || c->nrep == 0;
    mds_status st = add_replicas(c, count, n, r_mag, r_hydr, l_box, field_z, seeds, pos, pos_mm, pos_ea, first_index);
    if (st != MDS_OK) return st;
    c->state_set = had_state && pos != nullptr;
    return MDS_OK;
}

mds_status mds_remove_replica(mds_ctx* c, int32_t index) {
    if (!c || index < 0 || index >= c->nrep) return fail(c, MDS_EINVAL, "no such replica");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const RepMeta gone = c->meta[index];
    const int64_t old_n = c->n, new_n = old_n - gone.n;
    std::vector<Seg> segs{{0, 0, gone.off}, {gone.off + gone.n, gone.off, old_n - gone.off - gone.n}};
    mds_status st = relayout(c, old_n, new_n, segs);
    if (st != MDS_OK) return st;
    std::vector<int> keep;
    for (int r = 0; r < c->nrep; ++r)
        if (r != index) keep.push_back(r);
    if ((st = relayout_rep_values(c, keep, 0, 0)) != MDS_OK) return st;
    auto cut = [&](std::vector<double>& v) { v.erase(v.begin() + gone.off, v.begin() + gone.off + gone.n); };
    cut(c->h_r_mag);
    cut(c->h_mag_mom);
    cut(c->h_r_hydr);
    c->meta.erase(c->meta.begin() + index);
    refresh_offsets(c);
    if (c->nrep == 0) c->state_set = false;
    return MDS_OK;
}

int32_t mds_replica_count(const mds_ctx* c) { return c ? c->nrep : 0; }

mds_status mds_replica_set_ewald(mds_ctx* c, int32_t index, double alpha, double r_cut, int32_t nc) {
    if (!c || index < 0 || index >= c->nrep) return fail(c, MDS_EINVAL, "no such replica");
    return set_rep_ewald(c, index, alpha, r_cut, nc);
}

mds_status mds_replica_info(mds_ctx* c, int32_t index, int64_t* n, int64_t* offset, double* l_box, double* alpha,
                            double* r_cut, int32_t* nc, double* t, uint64_t* steps) {
    if (!c || index < 0 || index >= c->nrep) return fail(c, MDS_EINVAL, "no such replica");
    const RepMeta& m = c->meta[index];
    if (n) *n = m.n;
    if (offset) *offset = m.off;
    if (l_box) *l_box = m.L;
    if (alpha) *alpha = m.alpha;
    if (r_cut) *r_cut = m.rc;
    if (nc) *nc = m.nc;
    if (t || steps) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (t) HIPCHK(hipMemcpyAsync(t, c->rep_t.p + index, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (steps)
            HIPCHK(hipMemcpyAsync(steps, c->rep_step.p + index, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return MDS_OK;
}

mds_status mds_replica_get_state(mds_ctx* c, int32_t index, double* pos, double* pos_mm, double* pos_ea) {
    if (!c || index < 0 || index >= c->nrep) return fail(c, MDS_EINVAL, "no such replica");
    HIPCHK(hipSetDevice(c->device));
    const RepMeta& m = c->meta[index];
    mds_status st;
    if (pos && (st = download_cols(c, pos, c->pos.p, c->n, m.off, m.n)) != MDS_OK) return st;
    if (pos_mm && (st = download_cols(c, pos_mm, c->mm.p, c->n, m.off, m.n)) != MDS_OK) return st;
    if (pos_ea && (st = download_cols(c, pos_ea, c->ea.p, c->n, m.off, m.n)) != MDS_OK) return st;
    return MDS_OK;
}

mds_status mds_replica_mz(mds_ctx* c, double* mz) {
    if (!c || !mz || c->nrep <= 0) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    if (!c->mz_out.p) return fail(c, MDS_EINVAL, "no step taken");
    HIPCHK(hipMemcpyAsync(mz, c->mz_out.p, c->nrep * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MDS_OK;
}

mds_status mds_reserve(mds_ctx* c, int64_t n_particles) {
    if (!c || n_particles < 0 || n_particles > ((int64_t)1 << 25)) return fail(c, MDS_EINVAL, "bad reservation");
    HIPCHK(hipSetDevice(c->device));
    c->reserve_n = n_particles;
    if (c->n > 0) {
        // grow the kept arrays now (identity layout)
        std::vector<Seg> segs{{0, 0, c->n}};
        mds_status st = relayout(c, c->n, c->n, segs);
        if (st != MDS_OK) return st;
    }
    mds_status st = size_scratch(c);
    if (st != MDS_OK) return st;
    for (int k = 0; k < 11; ++k) {
        const size_t need = (size_t)(k < 7 ? 1 : 3) * (size_t)std::max<int64_t>(1, n_particles);
        if (c->spare[k].n < need) HIPCHK(c->spare[k].ensure(need));
    }
    return MDS_OK;
}

}  // extern "C"

namespace {
// Streams of a context, on its CU mask: the main stream always, the two
// helper streams (k-space chain, contact search) while `helpers` is set.
// Existing streams are drained and replaced.
mds_status make_streams(mds_ctx* c, bool helpers) {
    HIPCHK(hipSetDevice(c->device));
    c->collect();
    c->drop_graph();
    for (hipStream_t q : {c->stream, c->side, c->aux})
        if (q) HIPCHK(hipStreamSynchronize(q));
    const int nq = helpers ? 3 : 1;
    hipStream_t s[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < nq; ++k) {
        const hipError_t e = c->cu_mask.empty()
                                 ? hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking)
                                 : hipExtStreamCreateWithCUMask(&s[k], (uint32_t)c->cu_mask.size(), c->cu_mask.data());
        if (e != hipSuccess) {
            for (int q = 0; q < k; ++q) (void)hipStreamDestroy(s[q]);
            return fail(c, MDS_EDEVICE, std::string("stream creation: ") + hipGetErrorString(e));
        }
    }
    for (hipStream_t q : {c->stream, c->side, c->aux})
        if (q) (void)hipStreamDestroy(q);
    c->stream = s[0];
    c->side = s[1];
    c->aux = s[2];
    return MDS_OK;
}
}  // namespace

extern "C" {

mds_status mds_set_cu_mask(mds_ctx* c, const uint32_t* mask, int32_t words) {
    if (!c || words < 0 || (words > 0 && !mask)) return fail(c, MDS_EINVAL, "bad CU mask");
    c->cu_mask.assign(mask, mask + words);
    return make_streams(c, c->side != nullptr);
}

mds_status mds_peak_probe(int device, double* fp64_tflops, double* fp32_tflops, double* copy_gbs) {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || device < 0 || device >= cnt) return MDS_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) return MDS_EDEVICE;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return MDS_EDEVICE;
    const int32_t blocks = 8 * cus, iters = 4096;  // 8 waves per SIMD
    const size_t copy_bytes = (size_t)1 << 30;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *out = nullptr, *src = nullptr, *dst = nullptr;
    mds_status st = MDS_OK;
    auto timed = [&](auto&& launch, int reps, float* ms) -> bool {
        if (launch() != hipSuccess) return false;  // warm-up
        if (hipEventRecord(e0, s) != hipSuccess) return false;
        for (int r = 0; r < reps; ++r)
            if (launch() != hipSuccess) return false;
        return hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
               hipEventElapsedTime(ms, e0, e1) == hipSuccess;
    };
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)) != hipSuccess ||
        hipMalloc(&src, copy_bytes) != hipSuccess || hipMalloc(&dst, copy_bytes) != hipSuccess) {
        st = MDS_ENOMEM;
    } else {
        const double flop = 2.0 * 8.0 * iters * 256.0 * blocks;
        float ms = 0;
        for (int fp64 = 1; fp64 >= 0 && st == MDS_OK; --fp64) {
            if (!timed([&] { return launch_peak_fma(s, fp64 != 0, blocks, iters, out); }, 5, &ms)) st = MDS_EDEVICE;
            double* dst_v = fp64 ? fp64_tflops : fp32_tflops;
            if (st == MDS_OK && dst_v) *dst_v = 5.0 * flop / (ms * 1e-3) / 1e12;
        }
        if (st == MDS_OK && hipMemsetAsync(src, 0, copy_bytes, s) != hipSuccess) st = MDS_EDEVICE;
        if (st == MDS_OK &&
            !timed([&] { return hipMemcpyAsync(dst, src, copy_bytes, hipMemcpyDeviceToDevice, s); }, 5, &ms))
            st = MDS_EDEVICE;
        // bytes read plus bytes written
        if (st == MDS_OK && copy_gbs) *copy_gbs = 5.0 * 2.0 * (double)copy_bytes / (ms * 1e-3) / 1e9;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (out) (void)hipFree(out);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return st;
}

int32_t mds_device_cus(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return v;
}

mds_status mds_set_kspace_waves(mds_ctx* c, int32_t waves) {
    if (!c || waves < -1 || waves > 16) return fail(c, MDS_EINVAL, "k-space waves per SIMD must be in [-1, 16]");
    c->kspace_waves = waves;
    c->drop_graph();
    return MDS_OK;
}

mds_status mds_set_kspace_overlap(mds_ctx* c, int32_t on) {
    if (!c) return MDS_EINVAL;
    c->overlap_k = on != 0;
    c->drop_graph();
    // the helper streams exist only while they are used
    if (c->overlap_k != (c->side != nullptr)) return make_streams(c, c->overlap_k);
    return MDS_OK;
}

mds_status mds_set_graph(mds_ctx* c, int32_t on) {
    if (!c) return MDS_EINVAL;
    c->use_graph = on != 0;
    c->drop_graph();
    return MDS_OK;
}

// ---------------------------------------------------------------- checkpoint
// A binary checkpoint holds what a later run needs to continue as if it had
// never stopped: per replica its particles, box, field, seed, Ewald
// parameters, contact margin, clock and step counter (the Philox counter),
// the state, and with observables on the reference and unwrapped positions;
// the cell tiling, the noise source, and the caller's t and dt. Doubles are
// stored as the context holds them, so a run resumed with Philox noise
// matches the uninterrupted run bit for bit.
//
// Layout (host byte order, little-endian on every ROCm host):
//   "MDSCKPT1"  u32 version  i32 precision solver fixed_dt nrep reach philox
//   observe  f64 t dt  i64 n_total
//   nrep x { i64 n  f64 L field_z  u64 seed  i32 ewald_set nc  f64 alpha r_cut
//            margin_scale t  u64 step }
//   f64 r_mag[n_total] r_hydr[n_total] pos[3 n_total] pos_mm[..] pos_ea[..]
//   (observe) f64 pos_ref[3 n_total] true_pos[3 n_total]      (column-major)
//   u64 FNV-1a 64 of all preceding bytes
}  // extern "C"

namespace {

constexpr char kCkptMagic[8] = {'M', 'D', 'S', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 2;

uint64_t fnv1a64(const unsigned char* p, size_t n) {
    uint64_t h = 14695981039346656037ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct CkptOut {
    std::vector<unsigned char> b;
    void raw(const void* p, size_t n) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        b.insert(b.end(), q, q + n);
    }
    template <typename T>
    void put(T v) { raw(&v, sizeof(T)); }
};

struct CkptIn {
    const unsigned char* p;
    size_t n, at = 0;
    bool ok = true;
    void raw(void* dst, size_t k) {
        if (!ok || at + k > n) {
            ok = false;
            return;
        }
        std::memcpy(dst, p + at, k);
        at += k;
    }
    template <typename T>
    T get() {
        T v{};
        raw(&v, sizeof(T));
        return v;
    }
};

struct CkptRep {
    int64_t n;
    double L, field_z;
    uint64_t seed;
    int32_t ewald_set, nc;
    double alpha, rc, margin, t;
    uint64_t step;
};

}  // namespace

extern "C" {

mds_status mds_save_checkpoint(mds_ctx* c, const char* path, double t, double dt) {
    if (!c || !path) return MDS_EINVAL;
    if (c->n <= 0 || !c->state_set) return fail(c, MDS_EINVAL, "nothing to save: particles and state not set");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    std::vector<double> rt(c->nrep);
    std::vector<uint64_t> rs(c->nrep);
    HIPCHK(hipMemcpy(rt.data(), c->rep_t.p, c->nrep * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rs.data(), c->rep_step.p, c->nrep * sizeof(uint64_t), hipMemcpyDeviceToHost));
    CkptOut o;
    o.raw(kCkptMagic, 8);
    o.put<uint32_t>(kCkptVersion);
    o.put<int32_t>(c->cfg.precision);
    o.put<int32_t>(c->cfg.solver);
    o.put<int32_t>(c->cfg.fixed_dt);
    o.put<int32_t>(c->nrep);
    o.put<int32_t>(c->reach);
    o.put<int32_t>(c->philox ? 1 : 0);
    o.put<int32_t>(c->observe ? 1 : 0);
    o.put<double>(t);
    o.put<double>(dt);
    o.put<int64_t>(n);
    for (int r = 0; r < c->nrep; ++r) {
        const RepMeta& m = c->meta[r];
        o.put<int64_t>(m.n);
        o.put<double>(m.L);
        o.put<double>(m.field_z);
        o.put<uint64_t>(m.seed);
        o.put<int32_t>(m.ewald_set ? 1 : 0);
        o.put<int32_t>(m.nc);
        o.put<double>(m.alpha);
        o.put<double>(m.rc);
        o.put<double>(m.margin_scale);
        o.put<double>(rt[r]);
        o.put<uint64_t>(rs[r]);
    }
    o.raw(c->h_r_mag.data(), n * sizeof(double));
    o.raw(c->h_r_hydr.data(), n * sizeof(double));
    std::vector<double> tmp(n3);
    std::vector<const DevBuf<double>*> arrays = {&c->pos, &c->mm, &c->ea};
    if (c->observe) {
        arrays.push_back(&c->pos_ref);
        arrays.push_back(&c->true_pos);
    }
    for (const DevBuf<double>* a : arrays) {
        HIPCHK(hipMemcpy(tmp.data(), a->p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        o.raw(tmp.data(), n3 * sizeof(double));
    }
    // pair lists: the skin, each replica's state and the build positions (a
    // resumed run rebuilds the lists from them, so it evaluates the pairs in
    // the order the uninterrupted run does)
    o.put<int32_t>(c->list_on() ? 1 : 0);
    o.put<double>(c->list_skin);
    if (c->list_on()) {
        std::vector<int32_t> ls(c->nrep);
        HIPCHK(hipMemcpy(ls.data(), c->list_state.p, c->nrep * sizeof(int32_t), hipMemcpyDeviceToHost));
        o.raw(ls.data(), c->nrep * sizeof(int32_t));
        HIPCHK(hipMemcpy(tmp.data(), c->pos_build.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        o.raw(tmp.data(), n3 * sizeof(double));
    }
    o.put<uint64_t>(fnv1a64(o.b.data(), o.b.size()));
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, MDS_EINVAL, std::string("cannot open ") + path + " for writing");
    const size_t w = std::fwrite(o.b.data(), 1, o.b.size(), f);
    const int cl = std::fclose(f);
    if (w != o.b.size() || cl != 0) return fail(c, MDS_EINVAL, std::string("short write to ") + path);
    return MDS_OK;
}

mds_status mds_load_checkpoint(mds_ctx* c, const char* path, double* t, double* dt) {
    if (!c || !path) return MDS_EINVAL;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, MDS_EINVAL, std::string("cannot open ") + path);
    std::vector<unsigned char> b;
    unsigned char chunk[1 << 16];
    size_t k;
    while ((k = std::fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + k);
    std::fclose(f);
    if (b.size() < 8 + sizeof(uint64_t) || std::memcmp(b.data(), kCkptMagic, 8) != 0)
        return fail(c, MDS_EINVAL, "not a checkpoint file");
    uint64_t sum;
    std::memcpy(&sum, b.data() + b.size() - sizeof(uint64_t), sizeof(uint64_t));
    if (fnv1a64(b.data(), b.size() - sizeof(uint64_t)) != sum) return fail(c, MDS_EINVAL, "checkpoint checksum mismatch");
    CkptIn in{b.data(), b.size() - sizeof(uint64_t)};
    in.at = 8;
    if (in.get<uint32_t>() != kCkptVersion) return fail(c, MDS_EINVAL, "unsupported checkpoint version");
    const int32_t precision = in.get<int32_t>(), solver = in.get<int32_t>(), fixed = in.get<int32_t>();
    if (precision != c->cfg.precision || solver != c->cfg.solver || fixed != c->cfg.fixed_dt)
        return fail(c, MDS_EINVAL, "checkpoint written with another precision, solver or fixedDt setting");
    const int32_t nrep = in.get<int32_t>(), reach = in.get<int32_t>();
    const int32_t philox = in.get<int32_t>(), observe = in.get<int32_t>();
    const double t0 = in.get<double>(), dt0 = in.get<double>();
    const int64_t n = in.get<int64_t>();
    if (!in.ok || nrep <= 0 || nrep > (1 << 20) || n <= 0 || n > ((int64_t)1 << 25) || reach < 0 || reach > 8)
        return fail(c, MDS_EINVAL, "corrupt checkpoint header");
    std::vector<CkptRep> reps(nrep);
    int64_t total = 0;
    for (CkptRep& r : reps) {
        r.n = in.get<int64_t>();
        r.L = in.get<double>();
        r.field_z = in.get<double>();
        r.seed = in.get<uint64_t>();
        r.ewald_set = in.get<int32_t>();
        r.nc = in.get<int32_t>();
        r.alpha = in.get<double>();
        r.rc = in.get<double>();
        r.margin = in.get<double>();
        r.t = in.get<double>();
        r.step = in.get<uint64_t>();
        if (r.n <= 0 || !(r.L > 0)) in.ok = false;
        total += r.n;
    }
    const size_t n3 = 3 * (size_t)n;
    const size_t need = (2 * (size_t)n + (observe ? 5 : 3) * n3) * sizeof(double);
    if (!in.ok || total != n || in.n - in.at < need + sizeof(int32_t) + sizeof(double))
        return fail(c, MDS_EINVAL, "corrupt checkpoint body");
    std::vector<double> r_mag(n), r_hydr(n), st(3 * n3), obs_pos, pos_build;
    in.raw(r_mag.data(), n * sizeof(double));
    in.raw(r_hydr.data(), n * sizeof(double));
    in.raw(st.data(), 3 * n3 * sizeof(double));
    if (observe) {
        obs_pos.resize(2 * n3);
        in.raw(obs_pos.data(), 2 * n3 * sizeof(double));
    }
    const int32_t list_on = in.get<int32_t>();
    const double list_skin = in.get<double>();
    std::vector<int32_t> list_state;
    if (list_on) {
        if (in.n - in.at != nrep * sizeof(int32_t) + n3 * sizeof(double) || !(list_skin > 0))
            return fail(c, MDS_EINVAL, "corrupt checkpoint pair-list section");
        list_state.resize(nrep);
        pos_build.resize(n3);
        in.raw(list_state.data(), nrep * sizeof(int32_t));
        in.raw(pos_build.data(), n3 * sizeof(double));
    } else if (in.n != in.at) {
        return fail(c, MDS_EINVAL, "corrupt checkpoint body");
    }
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    clear_replicas(c);
    c->list_skin = list_on ? list_skin : 0.0;
    std::vector<int64_t> ns(nrep);
    std::vector<double> Ls(nrep), fz(nrep);
    std::vector<uint64_t> seeds(nrep);
    for (int r = 0; r < nrep; ++r) {
        ns[r] = reps[r].n;
        Ls[r] = reps[r].L;
        fz[r] = reps[r].field_z;
        seeds[r] = reps[r].seed;
    }
    mds_status s = add_replicas(c, nrep, ns.data(), r_mag.data(), r_hydr.data(), Ls.data(), fz.data(), seeds.data(),
                                st.data(), st.data() + n3, st.data() + 2 * n3, nullptr);
    if (s != MDS_OK) return s;
    for (int r = 0; r < nrep; ++r) {
        c->meta[r].margin_scale = reps[r].margin;
        if (reps[r].ewald_set && (s = set_rep_ewald(c, r, reps[r].alpha, reps[r].rc, reps[r].nc)) != MDS_OK) return s;
    }
    std::vector<double> rt(nrep);
    std::vector<uint64_t> rs(nrep);
    for (int r = 0; r < nrep; ++r) {
        rt[r] = reps[r].t;
        rs[r] = reps[r].step;
    }
    HIPCHK(hipMemcpy(c->rep_t.p, rt.data(), nrep * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->rep_step.p, rs.data(), nrep * sizeof(uint64_t), hipMemcpyHostToDevice));
    if (list_on) {
        // rebuilt from the stored build positions at the next geometry pass
        for (int32_t& v : list_state)
            if (v != LIST_INVALID) v = LIST_STALE;
        HIPCHK(hipMemcpy(c->list_state.p, list_state.data(), nrep * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->pos_build.p, pos_build.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    }
    c->reach = reach;
    c->planned = false;
    c->geometry_ready = false;
    c->philox = philox != 0;
    c->state_set = true;
    c->drop_graph();
    if (observe) {
        if ((s = mds_observe(c, 1)) != MDS_OK) return s;
        HIPCHK(hipMemcpy(c->pos_ref.p, obs_pos.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->true_pos.p, obs_pos.data() + n3, n3 * sizeof(double), hipMemcpyHostToDevice));
    }
    if (t) *t = t0;
    if (dt) *dt = dt0;
    return MDS_OK;
}

int64_t mds_device_bytes(const mds_ctx* c) {
    if (!c) return 0;
    int64_t b = 0;
    auto add = [&](size_t cnt, size_t sz) { b += (int64_t)(cnt * sz); };
    for (auto* d : {&c->r_mag, &c->mag_mom, &c->zeta_rot, &c->zeta_trans, &c->vol_mag, &c->c1, &c->c2, &c->pos, &c->mm,
                    &c->ea, &c->n_t, &c->n_b, &c->n_f, &c->pos0, &c->pos_cur, &c->force, &c->field, &c->mom, &c->kmm,
                    &c->kea, &c->pt[0], &c->pt[1], &c->new_mm, &c->new_ea, &c->ev_pos, &c->kpartial,
                    &c->tab, &c->press_partial, &c->series, &c->msum, &c->mz_out, &c->rep_t, &c->err_partial,
                    &c->true_pos, &c->pos_ref, &c->obs_d2, &c->obs_series, &c->obs_last, &c->snap, &c->snap_t})
        add(d->n, sizeof(double));
    add(c->pair_acc.n + c->order_tmp.n / 2, sizeof(uint64_t));
    for (auto* d : {&c->rep_of, &c->cell_rep, &c->key_main, &c->key_fine, &c->cnt_main, &c->start_main, &c->end_main,
                    &c->cur_main, &c->sorted_main, &c->inv_perm, &c->work, &c->cnt_fine, &c->start_fine, &c->end_fine,
                    &c->cur_fine, &c->sorted_fine, &c->cand_count, &c->cand_off, &c->cand_j, &c->ev_head, &c->ev_tail,
                    &c->ev_next, &c->moved_list, &c->cand_base, &c->cand_tmp})
        add(d->n, sizeof(int32_t));
    add(c->rec_pos.n + c->rec_mom.n + c->frec.n, sizeof(double4));
    add(c->rec_pos32.n + c->rec_mom32.n, sizeof(float4));
    add(c->ev_key.n + c->rep_step.n + c->snap_step.n, sizeof(uint64_t));
    add(c->stencil.n, sizeof(int4));
    add(c->rows.n + c->rows_m.n, sizeof(KRow));
    add(c->segs.n, sizeof(KSeg));
    add(c->pack_desc.n, sizeof(KPack));
    add(c->scan_tmp.n + c->scan_tmp2.n, sizeof(int32_t));
    for (const auto& b : c->spare) add(b.n, sizeof(double));
    add(c->red_chunks.n, sizeof(int2));
    add(c->red_part.n, sizeof(double));
    add(c->pack.n, sizeof(double4));
    add(c->tab32.n, sizeof(float));
    add(c->pack32.n, sizeof(float4));
    add(c->reps.n, sizeof(Rep));
    add(c->pos_build.n, sizeof(double));
    add(c->list.n + c->lcount.n + c->loff.n + c->list_state.n + c->list_mode.n + c->snap_list.n, sizeof(int32_t));
    add(c->list_maxd2.n + c->list_builds.n + c->list_used.n, sizeof(uint64_t));
    add(c->rec_build32.n, sizeof(float4));
    return b;
}

}  // extern "C"
