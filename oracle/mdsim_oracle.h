/*
 * mdsim_oracle.h - CPU restatement of the reference's Brownian/LLG step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, loaded by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg. It is never linked
 * into or called by the product library.
 *
 * Arrays of N x 3 doubles are column-major (x block, y block, z block), as the
 * reference's Eigen ArrayXXd.
 */
#ifndef MDSIM_ORACLE_H_
#define MDSIM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int64_t n;
    double l_box;
    /* Ewald: alpha, r_c, n_c (Params_S, src/typedefs.h:106-118) */
    double alpha, r_cut_lr, r_cut_sr;
    int32_t nc;
    /* per particle (PartProps_S, src/typedefs.h:21-34) */
    const double *r_mag, *r_hydr, *vol_mag, *mag_mom, *zeta_rot, *zeta_trans;
    /* configuration constants (src/configuration.cpp) */
    double my_pi, fac_ewald, kB, temp, gyro_mr, mag_damp, sat_mag, anis_en, anis_const,
        vel_mm_const, mag_flux_dens, t_mag, abs_tol, dt_min, dt_max;
    double eff_len_surf_mol, ster_rep_fac, electrostat_rep_fac, kappa, hamaker;
    int32_t enable_mobilization, enable_coating_pot, solver, fixed_dt;
    int32_t threads; /* OpenMP threads for the order-independent loops; <=0: 1 */
} orc_params;

/* Pair geometry of one CompInteractVecs call (src/interaction.cpp:88-137). */
typedef struct orc_pairs orc_pairs;

/* CompInteractVecs: scan all i<j, min image, list r<r_c, sequential overlap
 * correction that moves `pos` in place. Returns a new pair list. */
orc_pairs* orc_pairs_build(const orc_params* p, double* pos);
void orc_pairs_free(orc_pairs* pl);
int64_t orc_pairs_count(const orc_pairs* pl);
int64_t orc_pairs_overlaps(const orc_pairs* pl);
/* Copy out pair k's (i, j, r, ux, uy, uz, inSR). */
void orc_pairs_get(const orc_pairs* pl, int64_t* i, int64_t* j, double* r, double* u, int32_t* sr);

/* CompInteractions (src/simulation.cpp:628-680): force incl. therm_force,
 * demag field, pressure (Pxy,Pxz,Pyz). `pos` is moved by the overlap
 * correction; `pos_k` (positions for the k-space tables, EvalSinCos) is read
 * before any move. The pair list is returned in *pairs_out when not NULL. */
int orc_comp_interactions(const orc_params* p, double* pos, const double* pos_mm,
                          const double* therm_force, double* force, double* field, double* press,
                          orc_pairs** pairs_out);
/* CompDemagField (src/simulation.cpp:688-697) with a given pair list and the
 * k-space positions pos_k. */
int orc_comp_demag_field(const orc_params* p, const orc_pairs* pl, const double* pos_k,
                         const double* pos_mm, double* field);
/* k-space only, force + field (CompDipInteractEwaldF, src/interaction.cpp:429-514),
 * accumulated into force/field (self+surface term included). */
void orc_kspace(const orc_params* p, const double* pos_k, const double* mom, const double* mom_sum,
                double* force, double* field);
/* RKCoeffs (src/simulation.cpp:876-902). */
void orc_rk_coeffs(const orc_params* p, double bext_z, const double* pos_mm, const double* pos_ea,
                   const double* field, const double* therm_torque, const double* therm_field,
                   double* k_mm, double* k_ea);

/* One RunSimulation iteration (src/simulation.cpp:468-519) with the thermal
 * draws given as DefThermFluct leaves them (n_t, n_b, n_f; rescaled in place
 * on rejection as AdjustTimeStep does). Outputs: state in place, *t, *dt,
 * *dt_used, *mz; force/field/press of CompInteractions and the final demag
 * field (may be NULL). Returns 0, or -5 when a rejected step reached
 * deltaTmin (the reference would loop forever there), after max_tries tries. */
int orc_step(const orc_params* p, double* pos, double* pos_mm, double* pos_ea, double* t, double* dt,
             double* dt_used, double* mz, double* n_t, double* n_b, double* n_f, double* force,
             double* field, double* press, double* field_out, int32_t max_tries);

#ifdef __cplusplus
}
#endif

#endif
