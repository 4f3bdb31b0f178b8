// kernels.h - launchers of the gfx950 kernels (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mds {

// One k-space row: all n = (nx, ny, nz) with the same (ny, nz), nx in
// [-nxmax, nxmax] (nx in [1, nxmax] for the ny = nz = 0 row). poff is the
// index of the row's first entry in the packed scatter table (one entry per
// |nx| = 0..nxmax, padded with a zero entry to an even count).
struct KRow {
    int32_t ny, nz, nxmax, poff;
    int32_t koff;  // index of the row's first k vector in the S(k) arrays
};

// Packed scatter table entry for the k pair n = (+a, ny, nz), (-a, ny, nz):
// the indices of both in the S(k) arrays (-1: absent, as for a = 0's mirror
// or the ny = nz = 0 row's negative side) and e(k) = 2 exp(-w n^2) / n^2
// (the same for both). k_kpack turns it into (U+ + U-, U+ - U-, V+ + V-,
// V+ - V-) with U = e Re S, V = e Im S.
// t32: the entry's slot in the replica's FP32 tile-major copy for the
// matrix-core scatter, ((row / 16) (n_c + 2) + a) 16 + row % 16.
struct KPack {
    int32_t kp, km;
    double e;
    int32_t t32, pad_;
};

// FP32 sin/cos tables of the matrix-core kernels: per particle and axis the
// (cos, sin) pairs of n = 0..n_c, zero padded to a multiple of four floats.
constexpr int tw4_of(int nc) { return (2 * (nc + 1) + 3) & ~3; }

// A segment of a k-space row for the structure-factor build: |nx| in
// [alo, ahi] (both signs unless the row is ny = nz = 0), at most KSEG values.
constexpr int KSEG = 8;
// The matrix-core S(k) build (FP32 path) takes 16 rows and KBM_NX values of
// |nx| per wavefront and pass.
constexpr int KBM_ROWS = 16, KBM_NX = 24;
struct KSeg {
    int32_t ny, nz, alo, ahi, koff, nxmax;
};

// Parameters of one replica (an independent periodic system) in a context
// that holds several; particles, cells, k vectors and tables of all replicas
// are concatenated and every kernel finds its replica through this table.
struct Rep {
    int64_t off, n;              // particle range [off, off + n)
    double L, cell, fcell, margin;
    int32_t ncell, fncell, cell_off, fcell_off;  // grids; global cell id = cell_off + local
    int32_t sten_off, nsten;     // slice of the stencil list: the cell itself, the
    int32_t nhalf, pad3_;        // offsets positive in (z, y, x) order (nhalf with
                                 // the cell itself), then their negatives
    double alpha, alpha2, alpha4, rc, rc_sr;
    double ell, inv_ell, inv_mref, bscale, fscale, pscale;  // units of the FP32 pair path
    double fx_s[2], fx_is[2];    // fixed-point unit s1 of the pair sums (field, force) and 1/s1
    double px_s, px_is;          // fixed-point unit of the moved pairs' pressure sums (SI) and 1/unit
    float al_s, al2_s, al4_s, pad_;
    int32_t nc, row_off, nrows, seg_off, nseg, k_off, nk, kchunk, nchunks, kb_off, nsegw, nrowg;
    int32_t kbm_off, kbm_pass;   // matrix-core S(k) build: first work item, passes of KBM_NX |nx| values
    int32_t ksm_off;             // matrix-core scatter: first block (64 particles each)
    int32_t kmode;               // k space of this replica on the matrix cores: bit 0 S(k) build, bit 1 scatter
    int32_t pk_off, npk;         // slice of the packed scatter table
    int32_t red_off, nred;       // slice of the reduction chunk table
    int32_t ksplit[4];           // first row of each quarter of the scatter table (by entries)
    int64_t tab_off, part_off;   // offsets (doubles) into the sin/cos tables and S(k) partials
    int64_t tab32_off, pk32_off; // matrix-core path: offsets into the FP32 tables (floats) and the
                                 // FP32 tile-major scatter entries (float4s)
    double c_force, c_flux, c_self, c_surf, two_pi_over_L;
    double field_z;              // magFluxDens of this replica
    uint64_t seed;               // Philox key of this replica
    uint64_t stop_at;            // the replica stops stepping when its step counter reaches this
    // pair list (list_cap = 0: no list, the pair kernels search the cells
    // every evaluation): the list holds every pair within rlist = r_c + skin
    // of the build positions and stays valid while no particle has moved more
    // than skin / 2 from its build position (skin_half2 = (skin / 2)^2)
    double rlist, skin_half2;
    int64_t list_base, list_cap;  // the replica's region of the list (int32 words)
};

// Pair-list modes of a replica for the current step (list_mode): keep the
// list, build it from the current positions, or build it again from the
// stored build positions (after a change of layout or a resume: the same
// list as before, since a list is a function of its build positions); BAD:
// the list did not fit its region (flag 128), the replica's pairs are not
// evaluated this step. list_state: INVALID (no build positions yet), VALID,
// RESTORE (build positions valid, list to be rebuilt from them).
constexpr int32_t LIST_KEEP = 0, LIST_FRESH = 1, LIST_RESTORE = 2, LIST_BAD = 3;
constexpr int32_t LIST_INVALID = 0, LIST_VALID = 1, LIST_STALE = 2;
// A list window: two header words (first sorted slot j0 of the 64-particle
// block of a neighbour cell; E << 13 | count << 6 | image code) and E queue
// entries (j << 6) | i lane, each candidate's hits padded to whole quads.
constexpr int LIST_HDR = 2;

// Per-step geometry of the real-space sum (cell list + overlap versions).
struct Geometry {
    int64_t n;                   // all particles of all replicas
    double L, inv_L;             // set per replica inside the kernels
    const Rep* reps;
    const int32_t* rep_of;       // [n] replica of a particle
    const int32_t* cell_rep;     // [cells] replica of a main-grid cell
    // main cell grid
    int32_t ncell;        // cells per dimension
    int32_t reach;        // stencil half-width in cells
    double cell;          // cell edge
    const int32_t* cell_start;   // [ncell^3]
    const int32_t* cell_end;     // [ncell^3]
    const int32_t* sorted_idx;   // [n] original index of sorted slot
    const double4* rec_pos;      // [n] (x, y, z, tag) in sorted order, step-start positions
    const int32_t* work;         // [nwork] packed (cell << 8) | chunk
    const int32_t* nwork;        // device scalar
    // overlap-correction versions
    const int32_t* ev_head;      // [n] first event of particle or -1
    const int32_t* ev_next;      // [cap]
    const uint64_t* ev_key;      // [cap] lex key lo*n+hi of the pair that moved it
    const double* ev_pos;        // [cap*3] position after that event
    const int32_t* n_events;     // device scalar
    const double* max_disp;      // device scalar
    const int32_t* moved_list;   // original indices of moved particles
    const int32_t* n_moved;      // device scalar
    const int32_t* inv_perm;     // sorted slot of an original index
    const float4* rec_pos32;     // FP32 copy of rec_pos (candidate search)
    const float4* rec_mom32;     // FP32 moments / mref of the replica, sorted order (mixed path)
    // pair list (null when off)
    const float4* rec_build32;   // [n] FP32 build positions, sorted order (list build)
    const int32_t* list_mode;    // [replicas] LIST_KEEP / FRESH / RESTORE / BAD of this step
    const int32_t* lcount;       // [4n] words of the stream of (chunk start slot, wavefront)
    const int32_t* loff;         // [4n] its offset in the replica's region
    const int32_t* list;         // the streams
};

struct PairOut {
    double* force;   // [3n] SoA, written (not accumulated); may be null (field only)
    double* field;   // [3n] SoA, written
    double* press_partial;  // [3 * work items * PAIR_SLICES], then 3 per replica (moved pairs)
    // [12n] cell-sorted pair sums as fixed-point word pairs: component c
    // (force 0..2, field 3..5), word w at (2c + w) n + k
    unsigned long long* acc;
    int32_t* err;    // flag 64: a pair sum beyond the fixed-point range or not finite
};

// The pair kernel splits each work item's half stencil over this many
// wavefronts (one block each).
constexpr int PAIR_SLICES = 4;

struct PairParams {
    double alpha, alpha2, alpha4, ir_pi, fac, rc, rc_sr;
    // FP32 path: lengths in units of ell, moments in units of mref; the sums
    // are rescaled by bscale (field), fscale (force), pscale (pressure) when
    // written (FP64 path: ell = mref = 1, scales = facEwald)
    double ell, inv_ell, inv_mref, bscale, fscale, pscale;
    float al_s, al2_s, al4_s;
    // short-range (coating) constants
    int32_t coating;
    double eff_len, ster_fac, el_fac, kappa, hamaker;
    const double* r_mag;   // original order
};

// Launchers: each launches its kernel(s) on stream s and returns the launch status.
// skip (nullable): a device flag; when it is set the kernels that take it
// return at once (k_list_decide sets it when every replica keeps its list)
hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t count, int32_t v, const int32_t* skip = nullptr);
// peak probe: blocks x 256 threads, each `iters` x 8 FMAs; out holds one
// value per thread (double or float)
hipError_t launch_peak_fma(hipStream_t s, bool fp64, int32_t blocks, int32_t iters, void* out);
hipError_t launch_fill_f64(hipStream_t s, double* p, int64_t count, double v);
hipError_t launch_scale(hipStream_t s, double* p, int64_t count, double f);

// step setup: step-start positions and cell keys (global cell ids). With a
// pair list (pos_build not null) the main-grid keys are left to
// launch_list_keys, and each particle's squared displacement from its build
// position goes into the replica's maximum (list_maxd2, double bits).
hipError_t launch_prepare(hipStream_t s, int64_t n, const double* pos, double* pos0, double* pos_cur, const Rep* reps,
                          const int32_t* rep_of, int32_t* key_main, int32_t* key_fine, const double* pos_build,
                          const int32_t* list_state, unsigned long long* list_maxd2);
// Pair list: this step's mode per replica (from its state and maximum
// displacement; resets the maximum, marks the state valid), then the build
// positions of the replicas built afresh (<- pos0) and every particle's
// main-grid key from its build position.
// sort_skip: set to 1 when every replica keeps its list (the main grid's
// keys, sort, order and work items are then those of the last step).
hipError_t launch_list_decide(hipStream_t s, int32_t nrep, const Rep* reps, int32_t* list_state,
                              unsigned long long* list_maxd2, int32_t* list_mode, unsigned long long* list_builds,
                              int32_t* sort_skip);
hipError_t launch_list_keys(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of,
                            const int32_t* list_mode, const double* pos0, double* pos_build, int32_t* key_main,
                            const int32_t* skip = nullptr);
// Count pass, per-replica offsets, fill pass of the lists of the replicas
// whose mode is FRESH or RESTORE (grid: max_work blocks). list_used[r]: the
// replica's words; a list beyond its region sets flag 128, mode BAD and
// state INVALID.
hipError_t launch_list_build(hipStream_t s, const Geometry& g, int32_t nrep, int32_t max_work, const int4* stencil,
                             int32_t* lcount, int32_t* loff, int32_t* list, int32_t* list_mode,
                             int32_t* list_state, long long* list_used, int32_t* err);
// scan_tmp of scan_tiles(cells) ints
int32_t scan_tiles(int64_t m);
hipError_t launch_cell_sort(hipStream_t s, int64_t n, const int32_t* key, int32_t ncells, int32_t* count,
                            int32_t* start, int32_t* end, int32_t* cursor, int32_t* sorted_idx, int32_t* scan_tmp,
                            const int32_t* skip = nullptr);
// rec_pos[k].w carries a tag: the original index in the low 32 bits. With a
// pair list (pos_build not null) each record is the step-start position in
// the periodic image nearest the build position (so a list window's image
// shift stays right), rec_build32 gets the FP32 build position, and the
// stream lengths lcount of the replicas being built are cleared.
hipError_t launch_build_records(hipStream_t s, int64_t n, const int32_t* sorted_idx, const double* pos0,
                                double4* rec_pos, float4* rec_pos32, int32_t* inv_perm, const Rep* reps,
                                const int32_t* rep_of, const double* pos_build, const int32_t* list_mode,
                                float4* rec_build32, int32_t* lcount);
// after the overlap correction: flag the moved particles' tags
hipError_t launch_mark_moved(hipStream_t s, const int32_t* n_moved, const int32_t* moved_list,
                             const int32_t* inv_perm, double4* rec_pos);
hipError_t launch_build_work(hipStream_t s, int32_t ncells, const int32_t* start, const int32_t* end,
                             int32_t* work, int32_t* nwork, int32_t cap, int32_t* err, const int32_t* skip = nullptr);
// contact search + sequential overlap resolution
// frec: [n] scratch for the positions and radii in fine-cell order;
// cand_base: [n + 1] (each particle's run in cand_tmp, then the append
// counter); cand_tmp: [cand_cap] the runs in append order
hipError_t launch_contacts(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const double* pos0,
                           const double* r_mag, const int32_t* fstart, const int32_t* fend, const int32_t* fsorted,
                           double4* frec, int32_t* cand_count, int32_t* cand_off, int32_t* cand_j, int32_t cand_cap,
                           int32_t* cand_base, int32_t* cand_tmp, int32_t* err, int32_t* scan_tmp);
// one thread per replica walks its candidates in order
hipError_t launch_resolve(hipStream_t s, int32_t nrep, int64_t n, const Rep* reps, const double* r_mag,
                          const int32_t* cand_off, const int32_t* cand_j, double* pos_cur, const double* pos0,
                          int32_t* ev_head, int32_t* ev_tail, int32_t* ev_next, uint64_t* ev_key, double* ev_pos,
                          int32_t ev_cap, int32_t* n_events, double* max_disp, double disp_limit, int32_t* moved_list,
                          int32_t* n_moved, int32_t* err);
// clears the per-particle event lists and the event/moved counters
hipError_t launch_clear_events(hipStream_t s, int64_t n, int32_t* ev_head, int32_t* ev_tail, int32_t* n_events,
                               int32_t* n_moved, double* max_disp);

// moments m_i = mm_i * mu_i (original and sorted order; mom null: sorted
// only, inv_perm null: original only); per-replica sums
hipError_t launch_moments(hipStream_t s, int64_t n, const double* mm, const double* mag_mom, const int32_t* inv_perm,
                          const Rep* reps, const int32_t* rep_of, double* mom, double4* rec_mom, float4* rec_mom32);
// chunks: (replica, first particle) of RED_CHUNK-particle pieces that never
// straddle a replica; part: 3 doubles per chunk
constexpr int RED_CHUNK_PARTICLES = 2048;
hipError_t launch_rep_sum3(hipStream_t s, const Rep* reps, int32_t nrep, const int2* chunks, int32_t nch, int64_t n,
                           const double* v, double* part, double* out);
hipError_t launch_rep_mean_z(hipStream_t s, const Rep* reps, int32_t nrep, const int2* chunks, int32_t nch, int64_t n,
                             const double* v, double* part, double* out, const int32_t* counter, double* series,
                             int32_t series_cap);
// <m_z> of one system of n particles (moments column-major N x 3 on the
// device) into out[0], in the chunk order of launch_rep_mean_z, with the
// chunk partials in the library's static device array (no allocation).
constexpr int64_t NOALLOC_MAX_PARTICLES = (int64_t)1 << 27;
hipError_t launch_mean_z_noalloc(hipStream_t s, int64_t n, const double* mm, double* out);

// real space: one block per work item (max_work blocks, surplus blocks exit)
// Real-space pair sums (force pass: force, field, pressure; else field):
// the cell-tiled pair kernel and the pairs with an endpoint moved by the
// overlap correction (their pressure as fixed-point words at press_partial
// slot moved_press_off, 6 words per replica), into the fixed-point sums.
// launch_pairs_out then adds those sums to out.force / out.field, which the
// k-space scatter has written (it writes, the pair sums are added).
hipError_t launch_pairs(hipStream_t s, bool force, int32_t precision, const Geometry& g, const double4* rec_mom,
                        const PairParams& pp, const PairOut& out, int32_t max_work, const int4* stencil,
                        int32_t moved_press_off, int64_t* cnt_pairs);
// The same sums from the pair lists (no cell search): every pair of a list
// is evaluated with the exact r < r_c test on the records; the coating
// force pass keeps the searching kernel.
hipError_t launch_pairs_list(hipStream_t s, bool force, int32_t precision, const Geometry& g, const double4* rec_mom,
                             const PairParams& pp, const PairOut& out, int32_t max_work, const int4* stencil,
                             int32_t moved_press_off, int64_t* cnt_pairs);
hipError_t launch_pairs_out(hipStream_t s, const Geometry& g, const PairOut& out);
// EvalTransCoeffs per step (src/postprocessing.cpp:167-187): unwrap the
// true positions, per replica t, MSD x/y/z from pos_ref and the force pass's
// Pxy, Pxz, Pyz over the box volume into obs[7 r ..] and the series at the
// running step index. d2: [3n] scratch; part: the chunk sums buffer.
hipError_t launch_observables(hipStream_t s, const Rep* reps, int32_t nrep, const int32_t* rep_of,
                              const int2* chunks, int32_t nch, int64_t n, const double* pos, double* true_pos,
                              const double* pos_ref, double* d2, double* part, const double* rep_t,
                              const int32_t* work, const int32_t* nwork, const int32_t* cell_rep,
                              const double* press_partial, int64_t moved_press_off, double* obs,
                              const int32_t* counter, double* series, int32_t series_cap);
// Within each cell, the entries k_place left in `placed` (atomic order)
// ordered by value (original index) into sorted_idx, so the pair sums' order
// does not depend on scheduling; key: each particle's cell (as sorted).
hipError_t launch_cell_order(hipStream_t s, int64_t n, int32_t ncells, const int32_t* key, const int32_t* start,
                             const int32_t* end, const int32_t* placed, int32_t* sorted_idx,
                             const int32_t* skip = nullptr);

// k space
// FP64 tables for the replicas with a vector k-space kernel, FP32 ones for
// the replicas with a matrix-core kernel
hipError_t launch_sincos_tables(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const double* pos0,
                                double* tab, float* tab32);
// max_blocks > 0: a grid of at most that many blocks strides over the items
// (the build then holds fewer CU slots beside the pair kernels)
hipError_t launch_kbuild(hipStream_t s, int64_t n, const Rep* reps, int32_t nrep, int32_t kb_items, const KSeg* segs,
                         const double* tab, const double* mom, double* partial, int32_t max_blocks);
// The same S(k) partials (same chunks and layout) from FP32 matrix-core
// products, for the mixed-precision path.
// Every replica needs n_c <= KBM_NCMAX (the block's LDS).
constexpr int KBM_NCMAX = 50;
hipError_t launch_kbuild_mfma(hipStream_t s, int64_t n, const Rep* reps, int32_t nrep, int32_t items,
                              int32_t nc_max, const KRow* rows, const float* tab32, const double* mom,
                              double* partial, int32_t max_blocks);
// S(k) partials -> packed scatter table (fixed summation order)
// The k-space scatter from FP32 matrix-core products (mixed-precision path,
// n_c <= KSM_NCMAX in every replica); blocks: ceil(n_r / 64) per replica.
constexpr int KSM_NCMAX = 32;
hipError_t launch_kscatter_mfma(hipStream_t s, bool force, int64_t n, const Rep* reps, int32_t nrep, int32_t items,
                                int32_t nc_max, const KRow* rows, const float4* pack32, const float* tab32,
                                const double* mom, const double* msum, const double* add_force, double* force_out,
                                double* field);
// also the FP32 tile-major copy (units of the replica's largest moment) for
// the replicas scattered on the matrix cores
hipError_t launch_kpack(hipStream_t s, int32_t npk_all, const Rep* reps, int32_t nrep, const KPack* desc,
                        const double* partial, double4* pack, float4* pack32);
// field (and force) of the k-space sum, self and surface terms (and the
// thermal force add_force), written to field/force_out in original order;
// launch_pairs_out adds the real-space sums afterwards
hipError_t launch_kscatter(hipStream_t s, bool force, int64_t n, const Rep* reps, const int32_t* rep_of,
                           const KRow* rows, const double4* pack, const double* tab, const double* mom,
                           const double* msum, const double* add_force, double* force_out, double* field);

// integrator
struct StageParams {
    int64_t n;
    double h;            // step
    double anis_const, vel_mm_const, mag_damp;
    int32_t mobil;
    const double* c1;    // velEaConst1
    const double* c2;    // velEaConst2
    const double* mag_mom;
    const int32_t* inv_perm;
    // applied field: replica r has field_z while its time rep_t[r] < t_mag
    const double* rep_t;
    double t_mag;
    const Rep* reps;
    const int32_t* rep_of;
    float4* rec_mom32;   // FP32 copy of the stage-point moments (mixed pair path)
};
// k_s = RHS(y_s); then the next stage point (coef of length s+1) or, when
// last, the error norm partials and the normalised accepted state.
hipError_t launch_stage(hipStream_t s, const StageParams& sp, int32_t stage, int32_t stages, const double* mm0,
                        const double* ea0, const double* ymm, const double* yea, const double* field,
                        const double* n_t, const double* n_b, double* kmm, double* kea, const double* a_next,
                        const double* b_acc, const double* e_err, double* ymm_out, double* yea_out, double* mom,
                        double4* rec_mom, double* err_partial, double* new_mm, double* new_ea, int32_t nblk);
hipError_t launch_reduce_max(hipStream_t s, const double* partial, int32_t count, double scale, double* out);
hipError_t launch_translate(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const uint64_t* rep_step,
                            double h,
                            const double* pos_cur, const double* force, const double* zeta_trans, double* pos,
                            const double* new_mm, const double* new_ea, double* mm, double* ea);
hipError_t launch_noise(hipStream_t s, int64_t n, const Rep* reps, const int32_t* rep_of, const uint64_t* rep_step,
                        double dt, const double* zeta_rot, const double* vol_mag, const double* zeta_trans,
                        double kT, double alpha_g, double gyro, double sat_mag, int32_t en_t, int32_t en_b,
                        int32_t en_f, double* n_t, double* n_b, double* n_f);
hipError_t launch_step_count(hipStream_t s, int32_t nrep, const Rep* reps, uint64_t* rep_step, double* rep_t, double h,
                             int32_t* series_counter);

}  // namespace mds
