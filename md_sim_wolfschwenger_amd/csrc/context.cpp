// context.cpp - device context, step driver and C ABI (host code).
//
// Each entry point maps to one reference function (see include/mdsim_capi.h).
// The step follows RunSimulation's loop body (src/simulation.cpp:466-520);
// the adaptive controller is a restatement of AdjustTimeStep (912-951).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "kernels.h"
#include "mdsim_capi.h"

using namespace mds;

namespace {

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
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

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
    KC_COUNT
};
const char* kKClassNames[KC_COUNT] = {"setup",   "contacts", "sincos",          "moments",          "pairs_force",
                                      "pairs_field", "kbuild", "kreduce",        "kscatter_force",   "kscatter_field",
                                      "stage",   "translate", "noise"};

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

}  // namespace

struct mds_ctx {
    mds_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int64_t n = 0;
    double L = 0, rc_sr = 0, alpha = 0, rc = 0;
    int32_t nc = 0;
    bool ewald_set = false;
    bool state_set = false;
    bool geometry_ready = false;
    bool philox = false;
    uint64_t seed = 0;

    // per-particle constants
    DevBuf<double> r_mag, mag_mom, zeta_rot, zeta_trans, vol_mag, c1, c2;
    // state
    DevBuf<double> pos, pos0, pos_cur, mm, ea;
    // noise, outputs
    DevBuf<double> n_t, n_b, n_f, force, field, mom;
    DevBuf<double4> rec_pos, rec_mom;
    DevBuf<float4> rec_pos32;
    double mref = 1.0;  // largest magnetic moment (unit of the FP32 pair path)
    std::vector<double> h_mag_mom, h_r_hydr;
    // integrator
    DevBuf<double> kmm, kea, pt[2], new_mm, new_ea;
    // cell lists
    int32_t ncell = 1, reach = 3, fncell = 1;
    double cell = 0, fcell = 0, margin = 0, margin_scale = 0.1;
    int32_t work_cap = 0, cand_cap = 0, nsten = 0;
    DevBuf<int4> stencil;
    DevBuf<int32_t> key_main, key_fine, cnt_main, start_main, end_main, cur_main, sorted_main, inv_perm, work, nwork;
    DevBuf<int32_t> cnt_fine, start_fine, end_fine, cur_fine, sorted_fine;
    DevBuf<int32_t> cand_count, cand_off, cand_j, ev_head, ev_tail, ev_next, n_events, err_flag, moved_list, n_moved;
    DevBuf<uint64_t> ev_key;
    DevBuf<double> ev_pos, max_disp;
    // k space
    std::vector<KRow> h_rows;
    std::vector<KSeg> h_segs;
    int32_t nk = 0, nchunks = 1, kchunk = 256;
    DevBuf<KRow> rows;
    DevBuf<KSeg> segs;
    DevBuf<double> ek, sk, kpartial, tab;
    double c_force = 0, c_flux = 0, c_self = 0, c_surf = 0;
    // reductions
    DevBuf<double> msum_partial, msum, press_partial, err_partial, scal;  // scal: [0]=err [1]=mz [2..4]=press
    DevBuf<uint64_t> step_counter, pair_counter;
    DevBuf<int32_t> series_counter;
    DevBuf<double> series;

    // timing
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    int64_t launches[KC_COUNT] = {};
    double total_ms[KC_COUNT] = {};
    bool count_work = false;
    int64_t last_pairs = 0, last_events = 0;

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
    ~mds_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
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

// Cell grids: the main grid's stencil of +-reach cells must cover the
// candidate radius r_c + margin (the margin covers positions moved by the
// overlap correction); the fine grid covers the contact search radius.
mds_status plan_grids(mds_ctx* c) {
    const int64_t n = c->n;
    std::vector<double> rm(n);
    HIPCHK(hipMemcpy(rm.data(), c->r_mag.p, n * sizeof(double), hipMemcpyDeviceToHost));
    const double dmax = 2 * *std::max_element(rm.begin(), rm.end()) + 1e-10;
    if (c->margin_scale <= 0) c->margin_scale = 0.1;
    c->margin = c->margin_scale * dmax;
    const double cmin = (c->rc * (1 + 1e-9) + c->margin) / c->reach;
    int32_t nc = (int32_t)std::floor(c->L / cmin);
    if (nc < 2 * c->reach + 1) nc = 1;
    nc = std::min(nc, 96);
    c->ncell = nc;
    c->cell = c->L / nc;
    const double fmin = dmax + c->margin;
    int32_t fn = (int32_t)std::floor(c->L / fmin);
    if (fn < 3) fn = 1;
    fn = std::min(fn, 128);
    c->fncell = fn;
    c->fcell = c->L / fn;
    // stencil of neighbour-cell offsets whose nearest points lie within the
    // candidate radius (corner cells of the cube of +-reach are dropped)
    {
        std::vector<int4> st;
        const int R = nc == 1 ? 0 : c->reach;
        const double lim = c->rc * (1 + 1e-9) + c->margin;
        for (int dz = -R; dz <= R; ++dz)
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    auto gap = [&](int o) { return std::max(std::abs(o) - 1, 0) * c->cell; };
                    const double d2 = gap(dx) * gap(dx) + gap(dy) * gap(dy) + gap(dz) * gap(dz);
                    if (d2 < lim * lim) st.push_back(make_int4(dx, dy, dz, 0));
                }
        c->nsten = (int32_t)st.size();
        HIPCHK(c->stencil.ensure(st.size()));
        HIPCHK(hipMemcpy(c->stencil.p, st.data(), st.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    const int64_t ncm = (int64_t)nc * nc * nc, ncf = (int64_t)fn * fn * fn;
    c->work_cap = (int32_t)((n + 63) / 64 + ncm);
    c->cand_cap = (int32_t)(2 * n + 4096);
    HIPCHK(c->cnt_main.ensure(ncm));
    HIPCHK(c->start_main.ensure(ncm));
    HIPCHK(c->end_main.ensure(ncm));
    HIPCHK(c->cur_main.ensure(ncm));
    HIPCHK(c->work.ensure(c->work_cap));
    HIPCHK(c->cnt_fine.ensure(ncf));
    HIPCHK(c->start_fine.ensure(ncf));
    HIPCHK(c->end_fine.ensure(ncf));
    HIPCHK(c->cur_fine.ensure(ncf));
    HIPCHK(c->cand_j.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->ev_next.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->ev_key.ensure(2 * (size_t)c->cand_cap));
    HIPCHK(c->ev_pos.ensure(6 * (size_t)c->cand_cap));
    HIPCHK(c->press_partial.ensure(3 * ((size_t)c->work_cap + 1)));
    HIPCHK(c->moved_list.ensure(2 * (size_t)c->cand_cap));
    return MDS_OK;
}

// k-space rows of the half space n_z > 0, or n_z = 0 and (n_y > 0, or n_y = 0
// and n_x > 0), |n|^2 <= n_c^2; with the reference's full n_z = 0 plane at
// weight 1/2 (src/interaction.cpp:438-457) this is the same sum.
mds_status plan_kspace(mds_ctx* c) {
    const int nc = c->nc;
    const int64_t nc2 = (int64_t)nc * nc;
    c->h_rows.clear();
    c->h_segs.clear();
    std::vector<double> ek;
    const double w = std::pow(c->cfg.my_pi / (c->L * c->alpha), 2.0);
    int32_t koff = 0;
    for (int nz = 0; nz <= nc; ++nz)
        for (int ny = (nz == 0 ? 0 : -nc); ny <= nc; ++ny) {
            const int64_t r2 = (int64_t)ny * ny + (int64_t)nz * nz;
            if (r2 > nc2) continue;
            int nxmax = 0;
            while ((int64_t)(nxmax + 1) * (nxmax + 1) + r2 <= nc2) ++nxmax;
            const bool half = (ny == 0 && nz == 0);
            if (half && nxmax == 0) continue;
            c->h_rows.push_back({ny, nz, nxmax, koff});
            const int a0 = half ? 1 : 0;
            for (int a = a0; a <= nxmax; a += KSEG) c->h_segs.push_back({ny, nz, a, std::min(nxmax, a + KSEG - 1), koff, nxmax});
            const int lo = half ? 1 : -nxmax;
            for (int nx = lo; nx <= nxmax; ++nx) {
                const double len = (double)nx * nx + (double)r2;
                ek.push_back(2.0 * std::exp(-w * len) / len);
            }
            koff += half ? nxmax : 2 * nxmax + 1;
        }
    c->nk = koff;
    const double pi = c->cfg.my_pi, f = c->cfg.fac_ewald, L = c->L;
    // InitEwaldConst, src/interaction.cpp:366-369
    c->c_flux = 4.0 * pi * f / std::pow(L, 3.0);
    c->c_force = 8.0 * pi * pi * f / std::pow(L, 4.0);
    c->c_surf = f * 4.0 * pi / (3 * std::pow(L, 3.0));
    c->c_self = f * 4.0 * std::pow(c->alpha, 3.0) / (3.0 * std::sqrt(pi));
    const int64_t nsegw = ((int64_t)c->h_segs.size() + 63) / 64;
    int64_t chunk = (c->n * std::max<int64_t>(1, nsegw) + 4095) / 4096;
    chunk = std::max<int64_t>(32, std::min<int64_t>(1024, chunk));
    c->kchunk = (int32_t)chunk;
    c->nchunks = (int32_t)((c->n + chunk - 1) / chunk);
    HIPCHK(c->rows.ensure(std::max<size_t>(1, c->h_rows.size())));
    HIPCHK(c->segs.ensure(std::max<size_t>(1, c->h_segs.size())));
    HIPCHK(c->ek.ensure(std::max(1, c->nk)));
    HIPCHK(c->sk.ensure(4 * (size_t)std::max(1, c->nk)));
    HIPCHK(c->kpartial.ensure(2 * (size_t)std::max(1, c->nk) * c->nchunks));
    HIPCHK(c->tab.ensure((size_t)c->n * 3 * (nc + 1) * 2));
    if (!c->h_rows.empty())
        HIPCHK(hipMemcpy(c->rows.p, c->h_rows.data(), c->h_rows.size() * sizeof(KRow), hipMemcpyHostToDevice));
    if (!c->h_segs.empty())
        HIPCHK(hipMemcpy(c->segs.p, c->h_segs.data(), c->h_segs.size() * sizeof(KSeg), hipMemcpyHostToDevice));
    if (!ek.empty()) HIPCHK(hipMemcpy(c->ek.p, ek.data(), ek.size() * sizeof(double), hipMemcpyHostToDevice));
    return MDS_OK;
}

Geometry geometry(mds_ctx* c) {
    Geometry g{};
    g.n = c->n;
    g.L = c->L;
    g.ncell = c->ncell;
    g.reach = c->reach;
    g.cell = c->cell;
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
    return g;
}

PairParams pair_params(mds_ctx* c) {
    PairParams pp{};
    pp.alpha = c->alpha;
    pp.alpha2 = c->alpha * c->alpha;
    pp.alpha4 = pp.alpha2 * pp.alpha2;
    pp.ir_pi = c->cfg.ir_pi;
    pp.fac = c->cfg.fac_ewald;
    pp.rc = c->rc;
    pp.rc_sr = c->rc_sr;
    pp.coating = c->cfg.enable_coating_pot;
    pp.eff_len = c->cfg.eff_len_surf_mol;
    pp.ster_fac = c->cfg.ster_rep_fac;
    pp.el_fac = c->cfg.electrostat_rep_fac;
    pp.kappa = c->cfg.kappa;
    pp.hamaker = c->cfg.hamaker;
    pp.r_mag = c->r_mag.p;
    const double f = c->cfg.fac_ewald;
    if (c->cfg.precision == MDS_PREC_MIXED) {
        pp.ell = c->rc;
        pp.inv_ell = 1.0 / c->rc;
        pp.inv_mref = 1.0 / c->mref;
        pp.bscale = f * c->mref / std::pow(pp.ell, 3.0);
        pp.fscale = f * c->mref * c->mref / std::pow(pp.ell, 4.0);
        pp.pscale = pp.ell * pp.fscale;
    } else {
        pp.ell = pp.inv_ell = pp.inv_mref = 1.0;
        pp.bscale = pp.fscale = pp.pscale = f;
    }
    pp.al_s = (float)(c->alpha * pp.ell);
    pp.al2_s = pp.al_s * pp.al_s;
    pp.al4_s = pp.al2_s * pp.al2_s;
    return pp;
}

// EvalSinCos + the pair geometry of CompInteractVecs for the current
// positions: cell lists, contact candidates, sequential overlap correction.
mds_status prepare_geometry_once(mds_ctx* c) {
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    {
        mds_ctx::Span sp(c, KC_SETUP);
        HIPCHK(launch_prepare(s, n, c->pos.p, c->pos0.p, c->pos_cur.p, c->L, c->ncell, c->cell, c->key_main.p,
                              c->fncell, c->fcell, c->key_fine.p));
        HIPCHK(launch_cell_sort(s, n, c->key_main.p, c->ncell * c->ncell * c->ncell, c->cnt_main.p, c->start_main.p,
                                c->end_main.p, c->cur_main.p, c->sorted_main.p, nullptr));
        HIPCHK(launch_build_records(s, n, c->sorted_main.p, c->pos0.p, c->r_mag.p, c->rec_pos.p, c->rec_pos32.p,
                                    c->inv_perm.p));
        HIPCHK(launch_build_work(s, c->ncell * c->ncell * c->ncell, c->start_main.p, c->end_main.p, c->work.p,
                                 c->nwork.p, c->work_cap, c->err_flag.p));
    }
    {
        mds_ctx::Span sp(c, KC_CONTACTS);
        HIPCHK(launch_cell_sort(s, n, c->key_fine.p, c->fncell * c->fncell * c->fncell, c->cnt_fine.p,
                                c->start_fine.p, c->end_fine.p, c->cur_fine.p, c->sorted_fine.p, nullptr));
        HIPCHK(launch_clear_events(s, n, c->ev_head.p, c->ev_tail.p));
        HIPCHK(launch_contacts(s, n, c->L, c->pos0.p, c->r_mag.p, c->fncell, c->fcell, c->start_fine.p,
                               c->end_fine.p, c->sorted_fine.p, c->margin, c->cand_count.p, c->cand_off.p,
                               c->cand_j.p, c->cand_cap, nullptr, c->err_flag.p));
        HIPCHK(launch_resolve(s, n, c->L, c->rc, c->r_mag.p, c->cand_off.p, c->cand_j.p, c->pos_cur.p, c->pos0.p,
                              c->ev_head.p, c->ev_tail.p, c->ev_next.p, c->ev_key.p, c->ev_pos.p, 2 * c->cand_cap,
                              c->n_events.p, c->max_disp.p, 0.5 * c->margin, c->moved_list.p, c->n_moved.p,
                              c->err_flag.p));
    }
    {
        mds_ctx::Span sp(c, KC_SINCOS);
        HIPCHK(launch_sincos_tables(s, n, c->nc, c->L, 2 * c->cfg.my_pi, c->pos0.p, c->tab.p));
    }
    c->geometry_ready = true;
    return MDS_OK;
}

// With `check`, the overlap correction's displacement bound is verified
// before use: a move larger than half the contact-search margin could have
// hidden a contact, so the search is redone with a 4x wider margin.
mds_status prepare_geometry(mds_ctx* c, bool check) {
    for (int attempt = 0;; ++attempt) {
        HIPCHK(launch_fill_i32(c->stream, c->err_flag.p, 1, 0));
        mds_status st = prepare_geometry_once(c);
        if (st != MDS_OK || !check) return st;
        int32_t f = 0;
        HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (!(f & 32) || c->margin_scale >= 2.0 || attempt > 4) return MDS_OK;
        c->margin_scale = std::min(2.0, c->margin_scale * 4);
        if ((st = plan_grids(c)) != MDS_OK) return st;
    }
}

// Real-space + k-space Ewald (+ short range) for the moments in c->mom /
// c->rec_mom (their sum in c->msum). force_pass: CompDipInteractEwaldR/F
// with forces, pressure and the thermal force added; else the field-only
// CompDipFieldEwaldR/F of CompDemagField.
mds_status evaluate(mds_ctx* c, bool force_pass) {
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    Geometry g = geometry(c);
    PairParams pp = pair_params(c);
    PairOut po{force_pass ? c->force.p : nullptr, c->field.p, c->press_partial.p};
    {
        mds_ctx::Span sp(c, force_pass ? KC_PAIRS_FORCE : KC_PAIRS_FIELD);
        int64_t* cnt = c->count_work ? (int64_t*)c->pair_counter.p : nullptr;

        if (force_pass) HIPCHK(launch_fill_f64(s, c->press_partial.p + 3 * (size_t)c->work_cap, 3, 0.0));
        HIPCHK(launch_pairs(s, force_pass, c->cfg.precision, g, c->rec_mom.p, pp, po, c->work_cap, c->stencil.p,
                            c->nsten, cnt));
        HIPCHK(launch_pairs_moved(s, force_pass, g, c->rec_mom.p, pp, po, c->stencil.p, c->nsten, c->work_cap, cnt));
    }
    if (c->nk > 0) {
        {
            mds_ctx::Span sp(c, KC_KBUILD);
            HIPCHK(launch_kbuild(s, c->cfg.precision, n, c->nc, (int32_t)c->h_segs.size(), c->segs.p, c->tab.p,
                                 c->mom.p, c->kchunk, c->nchunks, c->nk, c->kpartial.p));
        }
        {
            mds_ctx::Span sp(c, KC_KREDUCE);
            HIPCHK(launch_kreduce(s, c->nk, c->nchunks, c->kpartial.p, c->sk.p));
        }
    }
    {
        mds_ctx::Span sp(c, force_pass ? KC_KSCATTER_FORCE : KC_KSCATTER_FIELD);
        HIPCHK(launch_kscatter(s, force_pass, c->cfg.precision, n, c->nc, (int32_t)c->h_rows.size(), c->rows.p,
                               c->ek.p, c->sk.p, c->tab.p, c->mom.p, c->msum.p, c->c_force, c->c_flux, c->c_self,
                               c->c_surf, force_pass ? c->n_f.p : nullptr, c->force.p, c->field.p));
    }
    return MDS_OK;
}

mds_status moments_of(mds_ctx* c, const double* mm_dev) {
    mds_ctx::Span sp(c, KC_MOMENTS);
    const int nb = blocks_for(c->n);
    HIPCHK(launch_moments(c->stream, c->n, mm_dev, c->mag_mom.p, c->inv_perm.p, c->mom.p, c->rec_mom.p,
                          c->msum_partial.p, nb));
    HIPCHK(launch_reduce_sum3(c->stream, c->msum_partial.p, nb, c->msum.p));
    return MDS_OK;
}

StageParams stage_params(mds_ctx* c, double h, double bext) {
    StageParams sp{};
    sp.n = c->n;
    sp.h = h;
    sp.bext_z = bext;
    sp.anis_const = c->cfg.anis_const;
    sp.vel_mm_const = c->cfg.vel_mm_const;
    sp.mag_damp = c->cfg.mag_damp;
    sp.mobil = c->cfg.enable_mobilization;
    sp.c1 = c->c1.p;
    sp.c2 = c->c2.p;
    sp.mag_mom = c->mag_mom.p;
    sp.inv_perm = c->inv_perm.p;
    return sp;
}

mds_status check_ready(mds_ctx* c) {
    if (!c) return MDS_EINVAL;
    if (c->n <= 0) return fail(c, MDS_EINVAL, "particles not set");
    if (!c->ewald_set) return fail(c, MDS_EINVAL, "Ewald parameters not set");
    if (!c->state_set) return fail(c, MDS_EINVAL, "state not set");
    return MDS_OK;
}

mds_status read_flags(mds_ctx* c) {
    int32_t f = 0, nev = 0;
    HIPCHK(hipMemcpyAsync(&f, c->err_flag.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&nev, c->n_events.p, sizeof(nev), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last_events = nev;
    if (f & 32) return fail(c, MDS_EOVERLAP, "overlap correction moved a particle beyond the contact search margin");
    if (f) return fail(c, MDS_ECAPACITY, "device capacity exceeded (flag " + std::to_string(f) + ")");
    return MDS_OK;
}

// One RunSimulation iteration. Stages run back to back on the stream; the
// adaptive controller reads the error norm (one sync per attempt).
mds_status step_impl(mds_ctx* c, double* t, double* dt, double* dt_used, double* mz_out, bool sync) {
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t s = c->stream;
    const double bext = (*t < c->cfg.t_mag) ? c->cfg.mag_flux_dens : 0.0;  // DefExtField
    if (c->philox) {
        mds_ctx::Span sp(c, KC_NOISE);
        HIPCHK(launch_noise(s, n, c->seed, c->step_counter.p, 0, *dt, c->zeta_rot.p, c->vol_mag.p, c->zeta_trans.p,
                            c->cfg.kB * c->cfg.temp, c->cfg.mag_damp, c->cfg.gyro_mr, c->cfg.sat_mag,
                            c->cfg.enable_therm_torque, c->cfg.enable_therm_field, c->cfg.enable_therm_force,
                            c->n_t.p, c->n_b.p, c->n_f.p));
    }
    mds_status st = prepare_geometry(c, sync);
    if (st != MDS_OK) return st;
    if ((st = moments_of(c, c->mm.p)) != MDS_OK) return st;
    if ((st = evaluate(c, true)) != MDS_OK) return st;

    const Tableau tb = make_tableau(c->cfg.fixed_dt ? MDS_HEUN_EULER : c->cfg.solver);
    const int nb = blocks_for(n);
    bool no_failed = true;
    double h = *dt;
    int tries = 0;
    for (;;) {
        h = *dt;
        const StageParams sp = stage_params(c, h, bext);
        for (int sgi = 0; sgi < tb.stages; ++sgi) {
            const double* ym = sgi == 0 ? c->mm.p : c->pt[(sgi - 1) & 1].p;
            const double* ye = sgi == 0 ? c->ea.p : c->pt[(sgi - 1) & 1].p + n3;
            double* yo = c->pt[sgi & 1].p;
            if (sgi > 0) {
                HIPCHK(launch_reduce_sum3(s, c->msum_partial.p, nb, c->msum.p));
                if ((st = evaluate(c, false)) != MDS_OK) return st;
            }
            mds_ctx::Span spn(c, KC_STAGE);
            const bool last = sgi + 1 == tb.stages;
            HIPCHK(launch_stage(s, sp, sgi, tb.stages, c->mm.p, c->ea.p, ym, ye, c->field.p, c->n_t.p, c->n_b.p,
                                c->kmm.p, c->kea.p, last ? nullptr : tb.a[sgi + 1], tb.b, tb.e, yo, yo + n3, c->mom.p,
                                c->rec_mom.p, c->msum_partial.p, c->err_partial.p, c->new_mm.p, c->new_ea.p, nb));
        }
        if (c->cfg.fixed_dt) break;
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
        // accept + renormalised state -> mm/ea; IntegrationTrans + wrap; <m_z>
        HIPCHK(launch_translate(s, n, h, c->L, c->pos_cur.p, c->force.p, c->zeta_trans.p, c->pos.p, c->new_mm.p,
                                c->msum_partial.p, nb));
        HIPCHK(hipMemcpyAsync(c->mm.p, c->new_mm.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ea.p, c->new_ea.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(launch_reduce_mean(s, c->msum_partial.p, nb, 1.0 / (double)n, c->scal.p + 1, c->series_counter.p,
                                  nullptr));
        HIPCHK(launch_step_count(s, c->step_counter.p));
    }
    *t += h;
    if (dt_used) *dt_used = h;
    if (sync || mz_out) {
        if ((st = read_flags(c)) != MDS_OK) return st;
        if (mz_out) HIPCHK(hipMemcpy(mz_out, c->scal.p + 1, sizeof(double), hipMemcpyDeviceToHost));
    }
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
    mds_ctx* c = new mds_ctx();
    c->cfg = *cfg;
    c->device = device;
    if (cfg->solver < 0 || cfg->solver > 2) {
        delete c;
        return MDS_ECONFIG;
    }
    if (cfg->precision != MDS_PREC_FP64 && cfg->precision != MDS_PREC_MIXED) {
        delete c;
        return MDS_EINVAL;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MDS_EDEVICE;
    }
    if (c->step_counter.ensure(1) != hipSuccess || c->pair_counter.ensure(1) != hipSuccess ||
        c->series_counter.ensure(1) != hipSuccess || c->scal.ensure(8) != hipSuccess ||
        c->msum.ensure(3) != hipSuccess || c->n_events.ensure(1) != hipSuccess || c->err_flag.ensure(1) != hipSuccess ||
        c->max_disp.ensure(1) != hipSuccess || c->nwork.ensure(1) != hipSuccess || c->n_moved.ensure(1) != hipSuccess) {
        delete c;
        return MDS_ENOMEM;
    }
    (void)hipMemset(c->step_counter.p, 0, sizeof(uint64_t));
    (void)hipMemset(c->pair_counter.p, 0, sizeof(uint64_t));
    (void)hipMemset(c->series_counter.p, 0, sizeof(int32_t));
    (void)hipMemset(c->err_flag.p, 0, sizeof(int32_t));
    (void)hipMemset(c->n_events.p, 0, sizeof(int32_t));
    (void)hipMemset(c->max_disp.p, 0, sizeof(double));
    (void)hipMemset(c->n_moved.p, 0, sizeof(int32_t));
    *out = c;
    return MDS_OK;
}

void mds_destroy(mds_ctx* c) { delete c; }

const char* mds_last_error(const mds_ctx* c) { return c ? c->err.c_str() : "null context"; }

mds_status mds_set_particles(mds_ctx* c, int64_t n, const double* r_mag, const double* r_hydr, double l_box) {
    if (!c || n <= 0 || !r_mag) return fail(c, MDS_EINVAL, "bad particle arguments");
    if (n > (int64_t)1 << 25) return fail(c, MDS_EINVAL, "at most 2^25 particles per context");
    HIPCHK(hipSetDevice(c->device));
    std::vector<double> rh(n), vm(n), mu(n), zr(n), zt(n), c1(n), c2(n);
    mds_status st = mds_particle_props(&c->cfg, n, r_mag, rh.data(), vm.data(), mu.data(), zr.data(), zt.data());
    if (st != MDS_OK) return fail(c, st, "particle properties");
    if (r_hydr) {
        // hydrodynamic radii given: friction coefficients follow them
        for (int64_t i = 0; i < n; ++i) {
            rh[i] = r_hydr[i];
            zr[i] = 8.0 * c->cfg.my_pi * c->cfg.vis * std::pow(rh[i], 3.0);
            zt[i] = 6.0 * c->cfg.my_pi * c->cfg.vis * rh[i];
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        c1[i] = 2.0 * c->cfg.anis_en * (vm[i] / zr[i]);  // velEaConst1, src/simulation.cpp:209
        c2[i] = 1.0 / zr[i];                               // velEaConst2, 210
    }
    c->n = n;
    c->L = l_box > 0 ? l_box : mds_box_length(&c->cfg, n, rh.data());
    st = mds_cutoff_sr(&c->cfg, n, r_mag, rh.data(), mu.data(), &c->rc_sr);
    if (st != MDS_OK) return fail(c, st, "short-range cut-off (InitCutOffSR)");
    const size_t n3 = 3 * (size_t)n;
    HIPCHK(c->r_mag.ensure(n));
    HIPCHK(c->mag_mom.ensure(n));
    HIPCHK(c->zeta_rot.ensure(n));
    HIPCHK(c->zeta_trans.ensure(n));
    HIPCHK(c->vol_mag.ensure(n));
    HIPCHK(c->c1.ensure(n));
    HIPCHK(c->c2.ensure(n));
    for (auto* b : {&c->pos, &c->pos0, &c->pos_cur, &c->mm, &c->ea, &c->n_t, &c->n_b, &c->n_f, &c->force, &c->field,
                    &c->mom, &c->new_mm, &c->new_ea})
        HIPCHK(b->ensure(n3));
    HIPCHK(c->pt[0].ensure(2 * n3));
    HIPCHK(c->pt[1].ensure(2 * n3));
    HIPCHK(c->kmm.ensure(7 * n3));
    HIPCHK(c->kea.ensure(7 * n3));
    HIPCHK(c->rec_pos.ensure(n));
    HIPCHK(c->rec_pos32.ensure(n));
    c->mref = *std::max_element(mu.begin(), mu.end());
    c->h_mag_mom = mu;
    c->h_r_hydr = rh;
    HIPCHK(c->rec_mom.ensure(n));
    for (auto* b : {&c->key_main, &c->key_fine, &c->sorted_main, &c->sorted_fine, &c->inv_perm, &c->cand_count,
                    &c->ev_head, &c->ev_tail})
        HIPCHK(b->ensure(n));
    HIPCHK(c->cand_off.ensure(n + 1));
    const int nb = blocks_for(n);
    HIPCHK(c->msum_partial.ensure(3 * (size_t)nb));
    HIPCHK(c->err_partial.ensure(nb));
    HIPCHK(hipMemcpy(c->r_mag.p, r_mag, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->mag_mom.p, mu.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->zeta_rot.p, zr.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->zeta_trans.p, zt.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->vol_mag.p, vm.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->c1.p, c1.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->c2.p, c2.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(launch_fill_f64(c->stream, c->n_t.p, n3, 0.0));
    HIPCHK(launch_fill_f64(c->stream, c->n_b.p, n3, 0.0));
    HIPCHK(launch_fill_f64(c->stream, c->n_f.p, n3, 0.0));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ewald_set = false;
    c->state_set = false;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status mds_set_ewald(mds_ctx* c, double alpha, double r_cut, int32_t nc) {
    if (!c || c->n <= 0) return fail(c, MDS_EINVAL, "set particles first");
    if (!(alpha > 0) || !(r_cut > 0) || nc < 0 || nc > 64) return fail(c, MDS_EINVAL, "bad Ewald parameters");
    if (r_cut > 0.5 * c->L * (1 + 1e-12)) return fail(c, MDS_EINVAL, "r_c above L/2 breaks the minimum image");
    HIPCHK(hipSetDevice(c->device));
    c->alpha = alpha;
    c->rc = r_cut;
    c->nc = nc;
    mds_status st = plan_grids(c);
    if (st != MDS_OK) return st;
    if ((st = plan_kspace(c)) != MDS_OK) return st;
    c->ewald_set = true;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status mds_set_tiling(mds_ctx* c, int32_t cell_reach) {
    if (!c || cell_reach < 1 || cell_reach > 8) return fail(c, MDS_EINVAL, "cell reach must be 1..8");
    c->reach = cell_reach;
    c->geometry_ready = false;
    if (c->ewald_set) return plan_grids(c);
    return MDS_OK;
}

mds_status mds_get_tiling(const mds_ctx* c, int32_t* cell_reach, int32_t* ncell, int32_t* nstencil) {
    if (!c) return MDS_EINVAL;
    if (cell_reach) *cell_reach = c->reach;
    if (ncell) *ncell = c->ncell;
    if (nstencil) *nstencil = c->nsten;
    return MDS_OK;
}

mds_status mds_autotune(mds_ctx* c, const double* rc_fracs, int32_t n_rc, const int32_t* reaches, int32_t n_reach,
                        int32_t reps, double* best_rc_frac, int32_t* best_reach, double* best_ms) {
    // CompEwaldParams (src/interaction.cpp:164-324) picks r_c in [0.2, 0.5] L
    // by timing the field sums on the CPU; here every candidate (r_c, cell
    // reach) is timed on the device over one force and one field evaluation,
    // with alpha and k_c from the same error model, so each candidate has the
    // accuracy the configuration asks for.
    if (!c || c->n <= 0 || !c->state_set || !rc_fracs || n_rc <= 0 || !reaches || n_reach <= 0)
        return fail(c, MDS_EINVAL, "autotune needs particles, state and candidates");
    HIPCHK(hipSetDevice(c->device));
    if (reps < 1) reps = 1;
    double best = 1e300, brc = -1;
    int32_t breach = c->reach;
    double b_alpha = 0;
    int32_t b_nc = 0;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool timing = c->timing;
    c->timing = false;
    mds_status st = MDS_OK;
    for (int a = 0; a < n_rc && st == MDS_OK; ++a) {
        const double rc = rc_fracs[a] * c->L;
        double alpha = 0, kc = 0;
        if (mds_ewald_params_for_rc(&c->cfg, c->n, c->h_mag_mom.data(), c->h_r_hydr.data(), c->L, rc, &alpha, &kc) !=
            MDS_OK)
            continue;
        const int32_t nc = (int32_t)std::floor(kc + 0.5);  // src/interaction.cpp:304
        if (nc > 64 || rc > 0.5 * c->L) continue;
        if ((st = mds_set_ewald(c, alpha, rc, nc)) != MDS_OK) break;
        for (int b = 0; b < n_reach && st == MDS_OK; ++b) {
            c->reach = reaches[b];
            if ((st = plan_grids(c)) != MDS_OK) break;
            if ((st = prepare_geometry(c, true)) != MDS_OK) break;
            float best_ms_here = 1e30f;
            for (int r = 0; r < reps + 1 && st == MDS_OK; ++r) {
                HIPCHK(hipEventRecord(e0, c->stream));
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
                b_alpha = alpha;
                b_nc = nc;
            }
        }
    }
    c->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != MDS_OK) return st;
    if (brc < 0) return fail(c, MDS_ENOCONV, "no admissible Ewald candidate");
    c->reach = breach;
    if ((st = mds_set_ewald(c, b_alpha, brc * c->L, b_nc)) != MDS_OK) return st;
    if (best_rc_frac) *best_rc_frac = brc;
    if (best_reach) *best_reach = breach;
    if (best_ms) *best_ms = best;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status mds_get_params(const mds_ctx* c, double* l_box, double* r_cut_sr, double* alpha, double* r_cut,
                          int32_t* nc) {
    if (!c) return MDS_EINVAL;
    if (l_box) *l_box = c->L;
    if (r_cut_sr) *r_cut_sr = c->rc_sr;
    if (alpha) *alpha = c->alpha;
    if (r_cut) *r_cut = c->rc;
    if (nc) *nc = c->nc;
    return MDS_OK;
}

mds_status mds_set_state(mds_ctx* c, const double* pos, const double* pos_mm, const double* pos_ea) {
    if (!c || c->n <= 0 || !pos || !pos_mm || !pos_ea) return fail(c, MDS_EINVAL, "bad state");
    const size_t n3 = 3 * (size_t)c->n;
    for (size_t k = 0; k < n3; ++k)
        if (!std::isfinite(pos[k]) || !std::isfinite(pos_mm[k]) || !std::isfinite(pos_ea[k]))
            return fail(c, MDS_ENONFINITE, "non-finite state");
    mds_status st;
    if ((st = upload(c, c->pos.p, pos, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->mm.p, pos_mm, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->ea.p, pos_ea, n3)) != MDS_OK) return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->state_set = true;
    c->geometry_ready = false;
    return MDS_OK;
}

mds_status mds_get_state(mds_ctx* c, double* pos, double* pos_mm, double* pos_ea) {
    if (!c || c->n <= 0) return MDS_EINVAL;
    const size_t n3 = 3 * (size_t)c->n;
    mds_status st;
    if (pos && (st = download(c, pos, c->pos.p, n3)) != MDS_OK) return st;
    if (pos_mm && (st = download(c, pos_mm, c->mm.p, n3)) != MDS_OK) return st;
    if (pos_ea && (st = download(c, pos_ea, c->ea.p, n3)) != MDS_OK) return st;
    return MDS_OK;
}

mds_status mds_set_noise(mds_ctx* c, const double* tt, const double* tb, const double* tf) {
    if (!c || c->n <= 0) return MDS_EINVAL;
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
    return MDS_OK;
}

mds_status mds_use_philox(mds_ctx* c, uint64_t seed, uint64_t counter) {
    if (!c) return MDS_EINVAL;
    c->philox = true;
    c->seed = seed;
    HIPCHK(hipMemcpy(c->step_counter.p, &counter, sizeof(counter), hipMemcpyHostToDevice));
    return MDS_OK;
}

mds_status mds_comp_interactions(mds_ctx* c, double* force, double* field, double* press, double* pos_after) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    HIPCHK(hipSetDevice(c->device));
    if ((st = prepare_geometry(c, true)) != MDS_OK) return st;
    if ((st = moments_of(c, c->mm.p)) != MDS_OK) return st;
    if ((st = evaluate(c, true)) != MDS_OK) return st;
    if ((st = read_flags(c)) != MDS_OK) return st;
    const size_t n3 = 3 * (size_t)c->n;
    if (force && (st = download(c, force, c->force.p, n3)) != MDS_OK) return st;
    if (field && (st = download(c, field, c->field.p, n3)) != MDS_OK) return st;
    if (pos_after && (st = download(c, pos_after, c->pos_cur.p, n3)) != MDS_OK) return st;
    if (press) {
        std::vector<double> pp(3 * (size_t)c->work_cap);
        int32_t nw = 0;
        HIPCHK(hipMemcpy(&nw, c->nwork.p, sizeof(nw), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(pp.data(), c->press_partial.p, 3 * (size_t)nw * sizeof(double), hipMemcpyDeviceToHost));
        press[0] = press[1] = press[2] = 0;
        for (int32_t w = 0; w < nw; ++w)
            for (int d = 0; d < 3; ++d) press[d] += pp[3 * (size_t)w + d];
        double pm[3];
        HIPCHK(hipMemcpy(pm, c->press_partial.p + 3 * (size_t)c->work_cap, sizeof(pm), hipMemcpyDeviceToHost));
        for (int d = 0; d < 3; ++d) press[d] += pm[d];
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
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    mds_status st;
    if ((st = upload(c, c->pt[0].p, pos_mm, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->pt[0].p + n3, pos_ea, n3)) != MDS_OK) return st;
    if ((st = upload(c, c->field.p, field, n3)) != MDS_OK) return st;
    const double bext = (t < c->cfg.t_mag) ? c->cfg.mag_flux_dens : 0.0;
    const StageParams sp = stage_params(c, 0.0, bext);
    const double zero[8] = {0};
    // a stage of a one-stage tableau: k only (the accepted-state outputs go to scratch)
    HIPCHK(launch_stage(c->stream, sp, 0, 1, c->pt[0].p, c->pt[0].p + n3, c->pt[0].p, c->pt[0].p + n3, c->field.p,
                        c->n_t.p, c->n_b.p, c->kmm.p, c->kea.p, nullptr, zero, zero, c->pt[1].p, c->pt[1].p + n3,
                        c->mom.p, c->rec_mom.p, c->msum_partial.p, c->err_partial.p, c->new_mm.p, c->new_ea.p,
                        blocks_for(n)));
    if (k_mm && (st = download(c, k_mm, c->kmm.p, n3)) != MDS_OK) return st;
    if (k_ea && (st = download(c, k_ea, c->kea.p, n3)) != MDS_OK) return st;
    return MDS_OK;
}

mds_status mds_step(mds_ctx* c, double* t, double* dt, double* dt_used, double* mz) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if (!t || !dt) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    return step_impl(c, t, dt, dt_used, mz, true);
}

mds_status mds_run(mds_ctx* c, int64_t steps, double* t, double* dt, double* mz_series) {
    mds_status st = check_ready(c);
    if (st != MDS_OK) return st;
    if (!t || !dt || steps < 0) return MDS_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    for (int64_t k = 0; k < steps; ++k) {
        double mz = 0;
        const bool want = mz_series != nullptr;
        st = step_impl(c, t, dt, nullptr, want ? &mz : nullptr, k == 0);
        if (st != MDS_OK) return st;
        if (want) mz_series[k] = mz;
    }
    return read_flags(c);
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
    if (kvecs) *kvecs = c->nk;
    if (overlap_events) *overlap_events = nev / 2;
    return MDS_OK;
}

}  // extern "C"
