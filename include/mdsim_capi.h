/*
 * mdsim_capi.h - C ABI of the MI355X (gfx950) Brownian/LLG step for magnetic
 * nanoparticle suspensions.
 *
 * The reference program has no plugin or FFI layer: its step is a set of free
 * C++ functions over Eigen arrays and a global configuration singleton
 * (src/simulation.h:28-66, src/interactions.h:27-72, src/configuration.h:98).
 * Each entry point below replaces one of them; the comment on each names the
 * reference function (file:line) it stands in for.
 *
 * Conventions
 *  - Every call returns an mds_status and never exits the process.
 *  - Arrays of N x 3 doubles are column-major, exactly as Eigen's ArrayXXd
 *    stores them in the reference: x[0..N), then y[0..N), then z[0..N).
 *  - Host pointers stay caller-owned and are read/written only during the call.
 *    Device memory is owned by the context.
 *  - One host thread per context; contexts on one device are independent.
 */
#ifndef MDSIM_CAPI_H_
#define MDSIM_CAPI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MDS_OK = 0,
    MDS_EINVAL = -1,      /* bad argument or state */
    MDS_ENOMEM = -2,      /* device or host allocation failed */
    MDS_EDEVICE = -3,     /* HIP runtime error / no device */
    MDS_ECONFIG = -4,     /* configuration file missing or malformed */
    MDS_ENOCONV = -5,     /* step rejected at deltaTmin (the reference loops forever here) */
    MDS_EOVERLAP = -6,    /* overlap correction moved a particle further than the contact search allows */
    MDS_ENONFINITE = -7,  /* NaN/Inf found in the state */
    MDS_ECAPACITY = -8    /* a fixed device capacity was exceeded */
} mds_status;

/* Solver ids of [adaptive_timestepping_solver] SOLVER (src/configuration.h:8-10). */
enum { MDS_HEUN_EULER = 0, MDS_BOGACKI_SHAMPINE = 1, MDS_DORMAND_PRINCE = 2 };
/* Arithmetic of the pair and k-space phase terms. */
enum { MDS_PREC_FP64 = 0, MDS_PREC_MIXED = 1 };

/*
 * Every key of the reference's configuration.ini (src/configuration.ini:1-79,
 * read by Configuration::Configuration, src/configuration.cpp:6-86) with the
 * same meaning, followed by the derived constants that constructor computes,
 * followed by keys this library adds (all optional; a value <= 0 / -1 keeps
 * the reference behaviour).
 */
typedef struct mds_config {
    /* [initial_configuration] */
    int32_t init_config_rot, init_config_trans, read_from_file;
    /* [size_distribution] */
    int32_t size_dist;
    /* [mobilization] */
    int32_t enable_mobilization;
    /* [thermal_fluctuations] */
    int32_t enable_therm_field, enable_therm_torque, enable_therm_force;
    int64_t seed;
    /* [interactions] */
    int32_t enable_coating_pot;
    /* [adaptive_timestepping_solver] */
    int32_t solver;
    /* [particle_properties] */
    int32_t np_dim;
    double mag_damp, temp, sat_mag, anis_en;
    /* [assemble_properties] */
    double r_mag_mean, d_shell, vol_frac;
    /* [steric_repulsion], [vdw_attraction], [electrostat_repulsion] */
    double dens_surf_mol, hamaker, z_valency, epsilon_r, gamma_0, conc;
    /* [time_settings], [external_magnetic_field_t] */
    double t_mag, t_relax, delta_t_init, mag_flux_dens;
    /* [solver_settings], [long_range_interactions], [short_range_interactions], [output_settings] */
    double abs_tol, delta_t_min, delta_t_max, err_tol_ewald, err_tol_sr, np_coords;
    /* [physical_constants], [faster_calculation] */
    double my_pi, kB, gyro_mr, avogadro, epsilon_0, el, fac_ewald;

    /* derived exactly as src/configuration.cpp:39,64,71-84 */
    int64_t numb_part;          /* npDim^3 unless numbPart is given */
    double mu0, eff_len_surf_mol, vel_mm_const, anis_const, ster_rep_fac, gamma_el, epsilon,
        kappa, electrostat_rep_fac, ir_pi, vis, sigma, t_end;

    /* added keys (section [mi355x]) */
    int64_t numb_part_override; /* numbPart: any N, not only npDim^3 */
    double viscosity;           /* viscosity: Pa s, overrides the T-derived value */
    double r_mag_median;        /* rMagMedian: lognormal median core radius */
    double sigma_log;           /* sigmaLog: lognormal sigma of ln(r) */
    int32_t fixed_dt;           /* fixedDt: 1 = Heun step at deltaTinit, controller bypassed */
    int32_t precision;          /* precision: MDS_PREC_FP64 or MDS_PREC_MIXED */
    double ewald_alpha_l;       /* ewaldAlphaL: alpha * L override */
    double ewald_rc_frac;       /* ewaldRc: r_c / L override */
    int32_t ewald_nc;           /* ewaldNc: k-space cut-off override */
} mds_config;

/* Parse an ini file with the reference's keys (Configuration::Configuration,
 * src/configuration.cpp:6-86; inih rules of src/INIReader.h). Missing keys keep
 * the reference's default of -1/false. Returns MDS_ECONFIG on a missing or
 * malformed file instead of printing and continuing. */
mds_status mds_config_from_ini(const char* path, mds_config* out);
/* Same, from ini text in memory. */
mds_status mds_config_from_string(const char* text, mds_config* out);
/* Recompute the derived constants after fields were edited. */
void mds_config_derive(mds_config* cfg);
/* sizeof(mds_config), for FFI bindings to check their struct layout. */
int64_t mds_config_sizeof(void);

/* Per-particle properties from core radii: InitParticleProps,
 * src/simulation.cpp:118-214 (radii are inputs because the reference draws
 * them from its own RNG). Outputs may be NULL. */
mds_status mds_particle_props(const mds_config* cfg, int64_t n, const double* r_mag, double* r_hydr,
                              double* vol_mag, double* mag_mom, double* zeta_rot, double* zeta_trans);
/* Box edge from sum of hydrodynamic volumes (src/simulation.cpp:56). */
double mds_box_length(const mds_config* cfg, int64_t n, const double* r_hydr);
/* Short-range cut-off: InitCutOffSR, src/interaction.cpp:20-57. */
mds_status mds_cutoff_sr(const mds_config* cfg, int64_t n, const double* r_mag, const double* r_hydr,
                         const double* mag_mom, double* r_cut_sr);
/* Closed-form rms error models, FunRealError / FunRezError
 * (src/interaction.cpp:330-347), for alpha = x (1/m) or k_c = x. */
double mds_ewald_real_error(const mds_config* cfg, int64_t n, double M, double f_al, double l_box,
                            double r_cut, double x);
double mds_ewald_rez_error(const mds_config* cfg, int64_t n, double M, double f_al, double l_box,
                           double alpha, double x);
/* alpha and k_c for a given r_c by the reference's root scans
 * (src/interaction.cpp:192-257); k_c is returned unrounded. */
mds_status mds_ewald_params_for_rc(const mds_config* cfg, int64_t n, const double* mag_mom,
                                   const double* r_hydr, double l_box, double r_cut, double* alpha,
                                   double* kc);

/* The rms force and torque errors the reference's error model predicts for
 * (alpha, r_c, n_c), relative to the contact force and torque of the
 * smallest particles: the estimates CompEwaldParams prints
 * (src/interaction.cpp:311-323). Outputs may be NULL. */
mds_status mds_ewald_predicted_errors(const mds_config* cfg, int64_t n, const double* mag_mom, const double* r_hydr,
                                      double l_box, double alpha, double r_cut, int32_t nc, double* force_rel,
                                      double* torque_rel);

/* Initial suspension: InitParticleProps' radii (src/simulation.cpp:132-214)
 * and InitCoordsFromSketch (src/simulation.cpp:220-351), drawn from this
 * library's seeded xoshiro256** stream instead of EigenRand's Mersenne
 * twister (same distributions, different numbers).
 * Radii: rMagMean for SIZE_DIST 0; for SIZE_DIST 1 the reference's lognormal
 * of mean rMagMean and standard deviation sigma; the added key sigmaLog > 0
 * selects a lognormal of median rMagMedian and sigma sigmaLog instead. */
mds_status mds_init_radii(const mds_config* cfg, int64_t n, uint64_t seed, double* r_mag);
/* Positions (N x 3): INIT_CONFIG_TRANS 2 is the npDim^3 grid; 1 is random
 * sequential addition in index order with the reference's rule (candidates in
 * [-(L/2 - r_i), L/2 - r_i)^3, rejected while nearer than r_i + r_j to an
 * earlier particle), found through a cell list. periodic != 0 draws over the
 * whole box and measures by minimum image, so no pair overlaps through the
 * faces either. MDS_ECAPACITY when 1e8 draws do not place all particles. */
mds_status mds_init_positions(const mds_config* cfg, int64_t n, const double* r_hydr, double l_box,
                              uint64_t seed, int32_t periodic, double* pos);
/* Easy axes uniform on the sphere by the reference's sampler (azimuth 2 pi u,
 * elevation 2 asin sqrt u) and moments per INIT_CONFIG_ROT (1: along z,
 * 2: on the easy axes, 3: easy axes and moments along z). */
mds_status mds_init_orientations(const mds_config* cfg, int64_t n, uint64_t seed, double* pos_mm,
                                 double* pos_ea);
/* The initial suspension drawn on HIP device `device`: radii as
 * mds_init_radii, the box from volFrac as the reference (src/simulation.cpp:
 * 56) unless *l_box > 0 is given (it returns the edge used), positions with
 * the acceptance rule of mds_init_positions (`periodic` as there;
 * INIT_CONFIG_TRANS 2: the npDim^3 grid) by parallel random sequential
 * addition (rounds in which every unplaced particle proposes a position and
 * is accepted when it overlaps no placed particle and no proposal of a lower
 * index), orientations as mds_init_orientations; all from the Philox4x32-10
 * stream keyed by `seed` (the host functions' distributions, other numbers).
 * Host outputs: r_mag[n] and pos, pos_mm, pos_ea (N x 3 column-major); any
 * but pos may be NULL. MDS_ECAPACITY when 20000 rounds leave particles
 * unplaced (a volume fraction too high for random addition). */
mds_status mds_init_suspension_device(const mds_config* cfg, int device, int64_t n, uint64_t seed, int32_t periodic,
                                      double* l_box, double* r_mag, double* pos, double* pos_mm, double* pos_ea);
/* READ_FROM_FILE: the state from a coordinates file (replaces
 * InitCoordsFromSim, src/simulation.cpp:357-382, reading with openData's
 * rules, src/postprocessing.cpp:37-75). The file is WriteCoords2TXT's N x 10
 * text (pos | posMm | posEa | tInt), one block or several separated by blank
 * lines; the last block is read into column-major N x 3 arrays, posMm from
 * columns 3-5 and posEa from 6-8 (the reference reads 6-8 and 9-11, past the
 * columns WriteCoords2TXT writes). t_int (may be NULL) gets the block's time.
 * MDS_EINVAL for an unreadable file or a last block that is not N x 10. The
 * text holds 6 significant digits. */
mds_status mds_init_coords_from_sim(const char* path, int64_t n, double* pos, double* pos_mm, double* pos_ea,
                                    double* t_int);

/* ----------------------------------------------------------------------- */
/* Device context                                                          */
/* ----------------------------------------------------------------------- */
typedef struct mds_ctx mds_ctx;

/* Create a context on HIP device `device`. A context owns its main HIP
 * stream, on which every call's work is ordered, and, while the k-space
 * overlap is on (the default; mds_set_kspace_overlap), two helper streams:
 * one takes the k-space chain beside the pair kernels, one the contact
 * search beside the cell sort (both joined by events). */
mds_status mds_create(const mds_config* cfg, int device, mds_ctx** out);
/* Destroy a context. Device buffers that contexts outgrow and the graph
 * executables they re-capture are not released while they run (releasing
 * waits for every stream of the device, which would stall the other
 * contexts on it); destroying any context of the device releases them. */
void mds_destroy(mds_ctx* ctx);
const char* mds_last_error(const mds_ctx* ctx);
/* Number of HIP devices visible (0 when none); no context needed. */
int mds_device_count(void);

/* Particle radii and box: InitParticleProps (src/simulation.cpp:118-214) and
 * the box edge (56), plus InitCutOffSR (src/interaction.cpp:20-57). l_box <= 0
 * derives it from the volume fraction as the reference does. */
mds_status mds_set_particles(mds_ctx* ctx, int64_t n, const double* r_mag, const double* r_hydr,
                             double l_box);
/* Ewald parameters: InitEwaldConst (src/interaction.cpp:353-374), with alpha,
 * r_c and n_c given explicitly. */
mds_status mds_set_ewald(mds_ctx* ctx, double alpha, double r_cut, int32_t nc);
/* Tiling of the real-space pair search (no counterpart in the reference,
 * whose pair scan is O(N^2), src/interaction.cpp:94-134): neighbour cells of
 * edge >= (r_c + margin) / cell_reach, searched over a culled stencil of
 * +-cell_reach cells. Default 3; chosen by the project's autotuner. 0 picks
 * per replica the largest reach (4..1) whose cells hold >= 32 particles on
 * average, so sparse replicas get larger cells. */
mds_status mds_set_tiling(mds_ctx* ctx, int32_t cell_reach);
mds_status mds_get_tiling(const mds_ctx* ctx, int32_t* cell_reach, int32_t* ncell, int32_t* nstencil);
/* Pair lists (no counterpart in the reference, whose CompInteractVecs
 * rebuilds the full pair list every call, src/interaction.cpp:94-134): each
 * replica keeps the pairs within r_l = r_c (1 + skin) of its build
 * positions, and both evaluations of a step read that list instead of
 * searching the cells. A replica rebuilds its list when one of its particles
 * has moved more than (r_l - r_c) / 2 from its build position (checked on the
 * device each step), so every pair within r_c is always in it; the pairs
 * evaluated, and their results, are those of the search. skin = 0 turns the
 * lists off (every evaluation searches the cells). Default 0.005. The result
 * of a replica depends only on its own trajectory (batches, resumes and
 * layout changes rebuild a list from the stored build positions), but it
 * differs from the search's in the order of the pair additions.
 * region_scale (default 1) sizes each replica's list region as a multiple of
 * the planner's estimate (mds_pair_list_words); a list that outgrows its
 * region raises flag 128 and the checked step that sees it grows the region
 * to the counted size and redoes the geometry (graph-replayed chunks are
 * redone that way). */
mds_status mds_set_pair_list(mds_ctx* ctx, double skin, double region_scale);
/* Builds of replica `index`'s list so far (the first included), words of
 * its last build and words of its region. */
mds_status mds_pair_list_stats(mds_ctx* ctx, int32_t index, int64_t* builds, int64_t* words, int64_t* capacity);
/* Words (int32) of the list region the planner gives a replica of n
 * particles in a box of edge l_box with cut-off r_cut and list skin `skin`
 * (0 when the arguments describe no list): for admission by footprint. */
int64_t mds_pair_list_words(int64_t n, double l_box, double r_cut, double skin);
/* Autotuner, in place of CompEwaldParams's wall-clock choice among 30 r_c in
 * [0.2, 0.5] L (src/interaction.cpp:164-324): for each r_c = rc_fracs[i] * L,
 * alpha and k_c come from the reference's error model (so accuracy is the
 * configured errTolEwald throughout) and each cell reach in `reaches` is
 * timed on the device over one force + one field evaluation of the current
 * state (`reps` timed repetitions after one warm-up). The fastest candidate
 * is installed (mds_set_ewald + mds_set_tiling) and returned. With several
 * replicas, rc_fracs[i] is r_c / L for a replica of the mean size and each
 * replica gets rc_fracs[i] * (N_mean / N_r)^(1/6), clamped to [0.1, 0.5] (the
 * real-space/k-space cost balance at fixed accuracy scales so), with its own
 * alpha and k_c from its own error model. A reach of 0 lets each replica
 * pick its own (see mds_set_tiling). */
mds_status mds_autotune(mds_ctx* ctx, const double* rc_fracs, int32_t n_rc, const int32_t* reaches,
                        int32_t n_reach, int32_t reps, double* best_rc_frac, int32_t* best_reach,
                        double* best_ms);
/* Read back box edge, r_c(SR), alpha, r_c, n_c. */
mds_status mds_get_params(const mds_ctx* ctx, double* l_box, double* r_cut_sr, double* alpha,
                          double* r_cut, int32_t* nc);
/* State: coordsTrans.pos, coordsRot.posMm, coordsRot.posEa (src/typedefs.h:37-47). */
mds_status mds_set_state(mds_ctx* ctx, const double* pos, const double* pos_mm, const double* pos_ea);
mds_status mds_get_state(mds_ctx* ctx, double* pos, double* pos_mm, double* pos_ea);

/* Thermal draws for the next step, as DefThermFluct leaves them
 * (src/simulation.cpp:563-603: already scaled by their sigma). Any NULL
 * channel is zero. mds_use_philox switches back to in-kernel Philox draws. */
mds_status mds_set_noise(mds_ctx* ctx, const double* therm_torque, const double* therm_field,
                         const double* therm_force);
mds_status mds_use_philox(mds_ctx* ctx, uint64_t seed, uint64_t counter);

/* The thermal draws of the last step (or the injected ones), scaled as
 * DefThermFluct leaves them in thermTorque, thermField, thermForce
 * (src/simulation.cpp:563-603; src/typedefs.h WorkingVar_S). Any output may
 * be NULL. */
mds_status mds_get_noise(mds_ctx* ctx, double* therm_torque, double* therm_field, double* therm_force);
/* Host copy of the in-kernel generator (no counterpart in the reference,
 * whose draws come from EigenRand's Mersenne twister, src/simulation.cpp:570,
 * 584, 596): one Philox4x32-10 block, out = Philox(ctr, key), and the nine
 * standard normals of particle `index` (within its replica) at the replica's
 * accepted step `step` (torque xyz, field xyz, force xyz before scaling by
 * sigma_T, sigma_B, sigma_F), computed by the same code as the noise kernel
 * (the device's sincospi may differ from the host's sin/cos by an ulp). */
void mds_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void mds_philox_normals(uint64_t seed, int64_t index, uint64_t step, double* z9);

/* CompInteractions (src/simulation.cpp:628-680) at time t: moves overlapping
 * pairs, returns force (incl. thermal force), demag field, pressure sums
 * (Pxy, Pxz, Pyz) and the corrected positions. Outputs may be NULL. */
mds_status mds_comp_interactions(mds_ctx* ctx, double* force, double* field, double* press,
                                 double* pos_after);
/* CompDemagField (src/simulation.cpp:688-697) for stage moments pos_mm,
 * reusing the pair geometry of the last mds_comp_interactions. */
mds_status mds_comp_demag_field(mds_ctx* ctx, const double* pos_mm, double* field);
/* RKCoeffs (src/simulation.cpp:876-902) for the given orientations, demag
 * field and the current thermal draws; t selects the external field. */
mds_status mds_rk_coeffs(mds_ctx* ctx, double t, const double* pos_mm, const double* pos_ea,
                         const double* field, double* k_mm, double* k_ea);

/* One iteration of RunSimulation's loop (src/simulation.cpp:466-520):
 * DefExtField, DefThermFluct, EvalSinCos, CompInteractions, IntegrationRot
 * (with AdjustTimeStep), IntegrationTrans, ApplyBoundaryCondArr, <m_z>.
 * t and dt are updated as the reference updates tInt and deltaT. */
mds_status mds_step(mds_ctx* ctx, double* t, double* dt, double* dt_used, double* mz);
/* `steps` iterations; with fixedDt they are enqueued back to back on the
 * context's stream (captured once into a HIP graph and replayed). mz_series
 * (may be NULL) receives <m_z> after each step. */
mds_status mds_run(mds_ctx* ctx, int64_t steps, double* t, double* dt, double* mz_series);

/* Graph-replayed runs (no counterpart in the reference, whose loop checks
 * every step in line): mds_run replays unchecked steps in chunks of at most
 * `steps` (default 128) and reads the flags after each chunk. A chunk that
 * raised the overlap-margin or a capacity flag is rolled back to its start
 * (positions, orientations, clocks, Philox counters) and redone eagerly with
 * every step checked, widening the contact-search margin where a move needs
 * it; a step that still fails ends the run with MDS_EOVERLAP/MDS_ECAPACITY,
 * the state it started from, and t advanced to it. mds_chunks_redone counts
 * the chunks redone since the context was created. */
mds_status mds_set_check_interval(mds_ctx* ctx, int64_t steps);
int64_t mds_chunks_redone(const mds_ctx* ctx);
/* Transport observables of EvalTransCoeffs (src/postprocessing.cpp:167-187,
 * initialised at src/simulation.cpp:108-109). on != 0: the current positions
 * become the reference and the first unwrapped positions, and every later
 * step records per replica (t before the step, MSDx, MSDy, MSDz, Pxy, Pxz,
 * Pyz): the mean squared displacement of the unwrapped positions from the
 * reference per axis after the step's translation, and the pressure tensor
 * components of the step's force evaluation over the box volume (the values
 * the reference pushes to MSD* and P*Vec each step). Adding or removing
 * replicas or setting particles switches it off. */
mds_status mds_observe(mds_ctx* ctx, int32_t on);
/* mds_run that also returns obs[step][replica][7] in the order above. */
mds_status mds_run_observed(mds_ctx* ctx, int64_t steps, double* t, double* dt, double* mz_series, double* obs);
/* The last step's observables, obs[replica][7] (after mds_step). */
mds_status mds_last_observables(mds_ctx* ctx, double* obs);
/* Checkpoint/resume (no counterpart in the reference, whose only restart is
 * READ_FROM_FILE above). mds_save_checkpoint writes, to a binary file, every
 * replica's particles, box, field, seed, Ewald parameters, clock and step
 * counter (the Philox counter), the state, the cell tiling, the noise source
 * and, with mds_observe on, the reference and unwrapped positions, with the
 * caller's t and dt; all doubles exactly. mds_load_checkpoint replaces the
 * context's replicas with the file's (the context must be created with the
 * same precision, solver and fixedDt) and returns t and dt: a run continued
 * from it with Philox noise is bit for bit the run that never stopped.
 * Injected noise (mds_set_noise) is not part of a checkpoint. MDS_EINVAL for
 * an unreadable, truncated or corrupted file (FNV-1a 64 checksum). */
mds_status mds_save_checkpoint(mds_ctx* ctx, const char* path, double t, double dt);
mds_status mds_load_checkpoint(mds_ctx* ctx, const char* path, double* t, double* dt);
/* Allocation-free <m_z> (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). pos_mm: device address of one
 * system's moments, column-major N x 3; mz: device address of one double,
 * which gets the mean z component in the fixed chunk order of
 * mds_replica_mz (src/simulation.cpp:514-515), bit for bit. Enqueued on
 * `stream` (a hipStream_t; NULL: the null stream) of the current device,
 * asynchronously. The chunk partials live in a static 512 KiB device array
 * that each device holds until the process ends, so the call allocates
 * nothing; n is limited to 2^27 (MDS_ECAPACITY beyond, MDS_EINVAL for
 * n <= 0 or a NULL pointer). Calls that overlap on one device race on that
 * array. */
mds_status mds_mean_z_noalloc(const double* pos_mm, int64_t n, double* mz, void* stream);
/* Wait for queued work. */
mds_status mds_synchronize(mds_ctx* ctx);

/* Kernel timing by HIP events on the context's stream. Enabled contexts
 * record an event pair around each launch; mds_kernel_stats returns, per
 * kernel class id (0..count-1), launches and total milliseconds. */
mds_status mds_enable_timing(mds_ctx* ctx, int32_t on);
int32_t mds_kernel_class_count(void);
const char* mds_kernel_class_name(int32_t id);
mds_status mds_kernel_stats(mds_ctx* ctx, int64_t* launches, double* total_ms);
/* Work counters of the last evaluation: unique pairs inside r_c, candidate
 * pairs visited, k-vectors in use, overlap events. */
mds_status mds_work_counts(mds_ctx* ctx, int64_t* pairs, int64_t* candidates, int64_t* kvecs,
                           int64_t* overlap_events);

/* ----------------------------------------------------------------------- */
/* Replicas: several independent systems in one context                    */
/* ----------------------------------------------------------------------- */
/* No counterpart in the reference, which runs one system per process
 * (src/simulation.cpp:12-16 keeps it in static globals). A context holds an
 * ordered list of replicas; their particles are concatenated in every array
 * the other entry points take or return (replica r owns rows
 * [offset_r, offset_r + n_r) of each column), and every kernel launch covers
 * all of them. Each replica has its own box, Ewald parameters, applied field
 * (magFluxDens, switched off at tMag of its own clock), Philox seed, step
 * counter and time. mds_set_particles makes a context of one replica.
 * Several replicas need fixedDt (the adaptive controller's error norm would
 * couple them). */

/* Append a replica: radii as mds_set_particles (r_hydr may be NULL), box
 * edge (<= 0: from volFrac), field_z in T, Philox seed, and optionally its
 * state (all three N x 3 arrays or none). Its clock and step counter start at
 * 0; its Ewald parameters must be set before the next step
 * (mds_replica_set_ewald, mds_set_ewald or mds_autotune). Device state of the
 * other replicas is kept on the device. */
mds_status mds_add_replica(mds_ctx* ctx, int64_t n, const double* r_mag, const double* r_hydr, double l_box,
                           double field_z, uint64_t seed, const double* pos, const double* pos_mm,
                           const double* pos_ea, int32_t* index);
/* Append `count` replicas at once (one re-layout of the device arrays):
 * n[count] sizes; r_mag, r_hydr (may be NULL) over all added particles in
 * order; l_box, field_z, seeds per replica (NULL: box from volFrac, the
 * configured magFluxDens, seed 0); pos, pos_mm, pos_ea (all or none) N x 3
 * column-major over all added particles. *first_index gets the first new
 * index. */
mds_status mds_add_replicas(mds_ctx* ctx, int32_t count, const int64_t* n, const double* r_mag,
                            const double* r_hydr, const double* l_box, const double* field_z,
                            const uint64_t* seeds, const double* pos, const double* pos_mm,
                            const double* pos_ea, int32_t* first_index);
/* Remove replica `index`; later replicas move down by one index. */
mds_status mds_remove_replica(mds_ctx* ctx, int32_t index);
int32_t mds_replica_count(const mds_ctx* ctx);
/* InitEwaldConst (src/interaction.cpp:353-374) for one replica. */
mds_status mds_replica_set_ewald(mds_ctx* ctx, int32_t index, double alpha, double r_cut, int32_t nc);
/* Size, row offset, box, Ewald parameters, clock and accepted steps. Any
 * output may be NULL. */
mds_status mds_replica_info(mds_ctx* ctx, int32_t index, int64_t* n, int64_t* offset, double* l_box,
                            double* alpha, double* r_cut, int32_t* nc, double* t, uint64_t* steps);
/* State of one replica (N_r x 3, column-major). */
mds_status mds_replica_get_state(mds_ctx* ctx, int32_t index, double* pos, double* pos_mm, double* pos_ea);
/* Replica `index` stops stepping when its step counter (accepted steps
 * since it joined) reaches `step`: later steps of the context leave its
 * state, clock and Philox counter as they are, so a request that needs k
 * steps can share runs of any length with others and still end with its
 * k-step state. UINT64_MAX (the default) never stops. No counterpart in the
 * reference; not part of a checkpoint. */
mds_status mds_replica_stop_at(mds_ctx* ctx, int32_t index, uint64_t step);
/* <m_z> of every replica after the last step (src/simulation.cpp:513). With
 * several replicas, mds_run's mz_series is steps x replicas, row-major. */
mds_status mds_replica_mz(mds_ctx* ctx, double* mz);
/* Capacity for n_particles in total, so that adding and removing replicas
 * below it reallocates no per-particle buffer. */
mds_status mds_reserve(mds_ctx* ctx, int64_t n_particles);
/* With on != 0, mds_run (fixedDt) captures one step into a HIP graph and
 * replays it (re-captured after any change of replicas, parameters or dt). */
mds_status mds_set_graph(mds_ctx* ctx, int32_t on);
/* Build the k-space sums S(k) on a second stream of the context while the
 * real-space pair kernels run, and the contact search on a third beside the
 * cell sort (default on; off: one stream, in order, and the helper streams
 * are released). Results are identical either way; kernel timing
 * (mds_enable_timing) runs them in order. No counterpart in the reference. */
mds_status mds_set_kspace_overlap(mds_ctx* ctx, int32_t on);
/* Beside the pair kernels, the S(k) build takes at most `waves` wavefronts
 * per SIMD (a grid of blocks striding over its work items; 0: one block per
 * item; -1, the default: one wave per SIMD for the matrix-core build of the
 * mixed path, no cap for the FP64 vector build, the faster setting of each
 * at C3). Results do not depend on it. No counterpart in the reference. */
mds_status mds_set_kspace_waves(mds_ctx* ctx, int32_t waves);
/* Restrict the context's work to the compute units set in mask (words x 32
 * bits, bit i = CU i; words = 0 restores all CUs). Contexts that serve
 * different request classes on one device can so be kept from queueing
 * behind each other's kernels. */
mds_status mds_set_cu_mask(mds_ctx* ctx, const uint32_t* mask, int32_t words);
/* Compute units of HIP device `device` (0 when unknown). */
int32_t mds_device_cus(int device);
/* Measured peaks of a device (no counterpart in the reference; for the
 * roofline): vector FMA throughput in FP64 and FP32 (8 independent chains
 * per lane, 8 waves per SIMD; packed FP32 is disabled in this build, so the
 * FP32 figure is the unpacked rate) and the device-to-device copy rate of a
 * 1 GiB buffer, bytes read plus written. Any output may be NULL. */
mds_status mds_peak_probe(int device, double* fp64_tflops, double* fp32_tflops, double* copy_gbs);
/* Device memory held by the context, in bytes. */
int64_t mds_device_bytes(const mds_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MDSIM_CAPI_H_ */
