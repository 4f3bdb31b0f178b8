// mdsim_oracle.cpp - CPU restatement of the reference's Brownian/LLG step.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared
// against; it is loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg, and never by the product library.
//
// It follows the reference's loop order and formulas (FP64 throughout), with
// 64-bit pair counts so it also runs where the reference's `int` pair count
// overflows (src/simulation.cpp:80, N > 46341). Each function cites the
// reference lines it restates. Arrays are column-major N x 3.
//
// Pinning (tests/test_oracle_*.py): every function here is compared with the
// golden vectors the reference itself wrote (tests/golden/*.npz), and, beyond
// those, with brute-force sums (real space, direct DFT), invariants (Newton's
// third law, alpha-invariance of the Ewald total, |m| = 1, m . k_Mm = 0) and
// closed forms (single-spin precession, cubic-lattice zero field).

#include "mdsim_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_pairs {
    std::vector<int64_t> i, j;
    std::vector<double> r, ux, uy, uz;
    std::vector<uint8_t> sr;
    int64_t overlaps = 0;
    // Streaming form (threads > 1 and no overlap move in this call): pairs
    // are recomputed row by row from these positions instead of stored.
    bool stream = false;
    std::vector<double> pos;
    int64_t count = 0;
};

namespace {

inline double& X(double* a, int64_t n, int64_t i, int d) { return a[d * n + i]; }
inline double X(const double* a, int64_t n, int64_t i, int d) { return a[d * n + i]; }

// ApplyBoundaryCondVec, src/interaction.cpp:145-155: one shift each way.
inline void min_image(double v[3], double L) {
    for (int d = 0; d < 3; ++d) {
        if (v[d] > L / 2) v[d] -= L;
        if (v[d] < -L / 2) v[d] += L;
    }
}

inline void set_threads(const orc_params* p) {
#ifdef _OPENMP
    omp_set_num_threads(p->threads > 0 ? p->threads : 1);
#else
    (void)p;
#endif
}

// magMomVec = posMm .* magMom, magMomSum = column sums
// (src/simulation.cpp:638-639, 691-692).
void moments(const orc_params* p, const double* mm, std::vector<double>& mom, double msum[3]) {
    const int64_t n = p->n;
    mom.assign(3 * n, 0.0);
    for (int d = 0; d < 3; ++d) {
        double s = 0;
        for (int64_t i = 0; i < n; ++i) {
            mom[d * n + i] = mm[d * n + i] * p->mag_mom[i];
            s += mom[d * n + i];
        }
        msum[d] = s;
    }
}

// Ewald constants of InitEwaldConst, src/interaction.cpp:353-374.
struct EwaldConst {
    double nc2, w, cFlux, cForce, cSurf, cSelf, alpha2, alpha4, irPi;
};
EwaldConst ewald_const(const orc_params* p) {
    EwaldConst c;
    const double pi = p->my_pi, L = p->l_box, f = p->fac_ewald;
    c.nc2 = (double)p->nc * p->nc;
    c.w = std::pow(pi / (L * p->alpha), 2.0);
    c.cFlux = 4.0 * pi * f / std::pow(L, 3.0);
    c.cForce = 8.0 * pi * pi * f / std::pow(L, 4.0);
    c.cSurf = f * 4.0 * pi / (3 * std::pow(L, 3.0));
    c.cSelf = f * 4.0 * std::pow(p->alpha, 3.0) / (3.0 * std::sqrt(pi));
    c.alpha2 = p->alpha * p->alpha;
    c.alpha4 = c.alpha2 * c.alpha2;
    c.irPi = 1 / std::sqrt(pi);  // src/configuration.cpp:80
    return c;
}

// EvalSinCos, src/interaction.cpp:521-548: cos/sin(2 pi n x_d / L) for
// n = 0..nc by the Rapaport recurrence. Layout: tab[(n*3 + d)*N + i].
void sincos_tables(const orc_params* p, const double* pos, std::vector<double>& tc, std::vector<double>& ts) {
    const int64_t n = p->n;
    const int nc = p->nc;
    tc.assign((size_t)(nc + 1) * 3 * n, 0.0);
    ts.assign((size_t)(nc + 1) * 3 * n, 0.0);
    const double t = 2 * p->my_pi / p->l_box;
    auto C = [&](int k, int d, int64_t i) -> double& { return tc[((size_t)k * 3 + d) * n + i]; };
    auto S = [&](int k, int d, int64_t i) -> double& { return ts[((size_t)k * 3 + d) * n + i]; };
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            const double tt = t * X(pos, n, i, d);
            C(0, d, i) = 1;
            S(0, d, i) = 0;
            C(1, d, i) = std::cos(tt);
            S(1, d, i) = std::sin(tt);
            const double u = 2.0 * C(1, d, i);
            if (nc >= 2) {
                C(2, d, i) = u * C(1, d, i) - 1;
                S(2, d, i) = u * S(1, d, i);
            }
            for (int k = 3; k <= nc; ++k) {
                C(k, d, i) = u * C(k - 1, d, i) - C(k - 2, d, i);
                S(k, d, i) = u * S(k - 1, d, i) - S(k - 2, d, i);
            }
        }
    }
}

// Real-space B/C/D coefficients, src/interaction.cpp:396-401.
inline void bcd(double r, double alpha, const EwaldConst& c, double& a1, double& a2, double& a3) {
    const double r2 = r * r;
    const double t = 2.0 * alpha * std::exp(-c.alpha2 * r2) * c.irPi / r2;
    a1 = std::erfc(alpha * r) / (r2 * r) + t;
    a2 = 3.0 * a1 / r2 + 2.0 * c.alpha2 * t;
    a3 = 5.0 * a2 / r2 + 4.0 * c.alpha4 * t;
}

// One pair of the streaming form: the same r, u and SR flag as
// CompInteractVecs computes (src/interaction.cpp:99-118) for i < j.
inline bool stream_pair(const orc_params* p, const double* pos, int64_t i, int64_t j, double& r, double u[3]) {
    const int64_t n = p->n;
    double v[3] = {X(pos, n, j, 0) - X(pos, n, i, 0), X(pos, n, j, 1) - X(pos, n, i, 1),
                   X(pos, n, j, 2) - X(pos, n, i, 2)};
    min_image(v, p->l_box);
    r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(r < p->r_cut_lr)) return false;
    u[0] = v[0] / r;
    u[1] = v[1] / r;
    u[2] = v[2] / r;
    return true;
}

// The per-pair body of CompDipInteractEwaldR / CompDipFieldEwaldR.
inline void pair_terms(const orc_params* p, const EwaldConst& c, double f, int64_t i, int64_t j, double r,
                       const double u[3], const double* mom, double* force, double* field, double* press) {
    const int64_t n = p->n;
    {
        const double r2 = r * r;
        double a1, a2, a3;
        bcd(r, p->alpha, c, a1, a2, a3);
        const double mi[3] = {X(mom, n, i, 0), X(mom, n, i, 1), X(mom, n, i, 2)};
        const double mj[3] = {X(mom, n, j, 0), X(mom, n, j, 1), X(mom, n, j, 2)};
        const double dot1 = mi[0] * u[0] + mi[1] * u[1] + mi[2] * u[2];
        const double dot2 = mj[0] * u[0] + mj[1] * u[1] + mj[2] * u[2];
        if (force) {
            const double mimj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
            double F[3];
            for (int d = 0; d < 3; ++d)
                F[d] = f * ((a2 * mimj - a3 * dot1 * dot2 * r2) * u[d] * r + a2 * r * (dot2 * mi[d] + dot1 * mj[d]));
            for (int d = 0; d < 3; ++d) {
                X(force, n, i, d) -= F[d];
                X(force, n, j, d) += F[d];
            }
            if (press) {
                press[0] += r * u[0] * F[1];
                press[1] += r * u[0] * F[2];
                press[2] += r * u[1] * F[2];
            }
        }
        for (int d = 0; d < 3; ++d) {
            X(field, n, i, d) += f * (a2 * dot2 * u[d] * r2 - a1 * mj[d]);
            X(field, n, j, d) += f * (a2 * dot1 * u[d] * r2 - a1 * mi[d]);
        }
    }
}

// CompDipInteractEwaldR (src/interaction.cpp:383-422) with EvalPressTensLR
// (src/postprocessing.cpp:195-200); with force == nullptr it is
// CompDipFieldEwaldR (src/interaction.cpp:555-582).
void real_space(const orc_params* p, const orc_pairs& pl, const double* mom, double* force, double* field,
                double* press) {
    const int64_t n = p->n;
    const EwaldConst c = ewald_const(p);
    const double f = p->fac_ewald;
    if (pl.stream) {
        // rows in parallel; each thread accumulates privately, reduced in thread order
        set_threads(p);
        const int T = std::max(1, (int)p->threads);
        std::vector<std::vector<double>> tf(T), tb(T);
        std::vector<std::array<double, 3>> tp(T, {0, 0, 0});
#pragma omp parallel num_threads(T)
        {
#ifdef _OPENMP
            const int t = omp_get_thread_num();
#else
            const int t = 0;
#endif
            std::vector<double>& Fl = tf[t];
            std::vector<double>& Bl = tb[t];
            if (force) Fl.assign(3 * n, 0.0);
            Bl.assign(3 * n, 0.0);
#pragma omp for schedule(dynamic, 16)
            for (int64_t i = 0; i < n - 1; ++i) {
                for (int64_t j = i + 1; j < n; ++j) {
                    double r, u[3];
                    if (!stream_pair(p, pl.pos.data(), i, j, r, u)) continue;
                    pair_terms(p, c, f, i, j, r, u, mom, force ? Fl.data() : nullptr, Bl.data(),
                               force ? tp[t].data() : nullptr);
                }
            }
        }
        for (int t = 0; t < T; ++t) {
            for (int64_t k = 0; k < 3 * n; ++k) {
                if (force) force[k] += tf[t][k];
                field[k] += tb[t][k];
            }
            if (force && press)
                for (int d = 0; d < 3; ++d) press[d] += tp[t][d];
        }
        return;
    }
    const int64_t P = (int64_t)pl.i.size();
    for (int64_t k = 0; k < P; ++k) {
        const int64_t i = pl.i[k], j = pl.j[k];
        const double r = pl.r[k];
        const double u[3] = {pl.ux[k], pl.uy[k], pl.uz[k]};
        pair_terms(p, c, f, i, j, r, u, mom, force, field, press);
    }
}

// CompDipInteractEwaldF (src/interaction.cpp:429-514) and, with force ==
// nullptr, CompDipFieldEwaldF (588-672): the n_z >= 0 half of the n-cube with
// the n_z = 0 plane in full at weight 1/2, two passes over the particles per
// k, then the self and vacuum-surface terms. The second pass is run particle
// by particle over all k; each particle still accumulates in k order, so the
// result is the same as the reference's k-outer loop.
void kspace(const orc_params* p, const double* pos_k, const double* mom, const double msum[3], double* force,
            double* field) {
    const int64_t n = p->n;
    const int nc = p->nc;
    const EwaldConst c = ewald_const(p);
    std::vector<double> tc, ts;
    sincos_tables(p, pos_k, tc, ts);
    struct K { int nx, ny, nz; double e; };
    std::vector<K> ks;
    for (int nz = 0; nz <= nc; nz++)
        for (int ny = -nc; ny <= nc; ny++)
            for (int nx = -nc; nx <= nc; nx++) {
                const double len = (double)nx * nx + (double)ny * ny + (double)nz * nz;
                if (len == 0 || len > c.nc2) continue;
                double e = 2.0 * std::exp(-c.w * len) / len;
                if (nz == 0) e *= 0.5;
                ks.push_back({nx, ny, nz, e});
            }
    const int64_t nk = (int64_t)ks.size();
    auto phase = [&](const K& k, int64_t i, double& pc, double& ps) {
        const int ax = std::abs(k.nx), ay = std::abs(k.ny), az = k.nz;
        const double cx = tc[((size_t)ax * 3 + 0) * n + i], cy = tc[((size_t)ay * 3 + 1) * n + i],
                     cz = tc[((size_t)az * 3 + 2) * n + i];
        double sx = ts[((size_t)ax * 3 + 0) * n + i], sy = ts[((size_t)ay * 3 + 1) * n + i];
        const double sz = ts[((size_t)az * 3 + 2) * n + i];
        if (k.nx < 0) sx = -sx;
        if (k.ny < 0) sy = -sy;
        pc = cx * cy * cz - cx * sy * sz - sx * cy * sz - sx * sy * cz;
        ps = sx * cy * cz + cx * sy * cz + cx * cy * sz - sx * sy * sz;
    };
    std::vector<double> sumC(nk), sumS(nk);
    set_threads(p);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t q = 0; q < nk; ++q) {
        double sc = 0, ss = 0;
        const K& k = ks[q];
        for (int64_t i = 0; i < n; ++i) {
            double pc, ps;
            phase(k, i, pc, ps);
            const double pre = k.nx * X(mom, n, i, 0) + k.ny * X(mom, n, i, 1) + k.nz * X(mom, n, i, 2);
            sc += pre * pc;
            ss += pre * ps;
        }
        sumC[q] = sc;
        sumS[q] = ss;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t q = 0; q < nk; ++q) {
            const K& k = ks[q];
            double pc, ps;
            phase(k, i, pc, ps);
            const double nv[3] = {(double)k.nx, (double)k.ny, (double)k.nz};
            if (force) {
                const double pre = nv[0] * X(mom, n, i, 0) + nv[1] * X(mom, n, i, 1) + nv[2] * X(mom, n, i, 2);
                const double s = c.cForce * k.e * pre * (sumC[q] * ps - sumS[q] * pc);
                for (int d = 0; d < 3; ++d) X(force, n, i, d) += s * nv[d];
            }
            const double b = c.cFlux * k.e * (sumC[q] * pc + sumS[q] * ps);
            for (int d = 0; d < 3; ++d) X(field, n, i, d) -= b * nv[d];
        }
        for (int d = 0; d < 3; ++d) X(field, n, i, d) += c.cSelf * X(mom, n, i, d) - c.cSurf * msum[d];
    }
}

// Short-range loop of CompInteractions (src/simulation.cpp:648-676):
// CompSterRep (src/interaction.cpp:766-788), CompElectrostatRep (796-809),
// CompForceVdW (817-832), EvalPressTensSR (src/postprocessing.cpp:208-213).
// The flag is r < rCutSR of this call (the reference's list is never reset,
// src/interaction.cpp:115-118; see the project's design notes).
void short_range(const orc_params* p, const orc_pairs& pl, double* force, double* press) {
    const int64_t n = p->n;
    const int64_t P = (int64_t)pl.i.size();
    for (int64_t k = 0; k < P; ++k) {
        if (!pl.sr[k]) continue;
        const int64_t i = pl.i[k], j = pl.j[k];
        const double r = pl.r[k];
        const double u[3] = {pl.ux[k], pl.uy[k], pl.uz[k]};
        const double ri = p->r_mag[i], rj = p->r_mag[j];
        double v2[3] = {0, 0, 0};
        // steric (uses r_i for the shell diameter and r_i^2)
        const double dDelta = 2 * (ri + p->eff_len_surf_mol);
        double s = 0;
        if (r <= dDelta) s = (2 * rj / (ri + rj)) * p->ster_rep_fac * ri * ri * std::log(dDelta / r);
        for (int d = 0; d < 3; ++d) {
            const double v = s * u[d];
            X(force, n, i, d) -= v;
            X(force, n, j, d) += v;
            v2[d] += v;
        }
        // electrostatic
        s = (ri * rj / (ri + rj)) * p->electrostat_rep_fac * std::exp(-p->kappa * (r - ri - rj));
        for (int d = 0; d < 3; ++d) {
            const double v = s * u[d];
            X(force, n, i, d) -= v;
            X(force, n, j, d) += v;
            v2[d] += v;
        }
        // van der Waals
        s = -p->hamaker * r * std::pow(4 * ri * rj, 3.0) /
            (6 * std::pow((r - ri - rj) * (r + ri + rj), 2.0) * std::pow(r * r - std::pow(ri - rj, 2.0), 2.0));
        for (int d = 0; d < 3; ++d) {
            const double v = s * u[d];
            X(force, n, i, d) -= v;
            X(force, n, j, d) += v;
            v2[d] += v;
        }
        press[0] += r * u[0] * v2[1];
        press[1] += r * u[0] * v2[2];
        press[2] += r * u[1] * v2[2];
    }
}

inline void cross(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// Butcher tableaux of IntegrationRot's three solvers (src/simulation.cpp:721-862).
struct Tableau {
    int stages;
    double a[7][7];
    double b[7];    // accepted combination (the last stage point for BS/DP)
    double e[7];    // error combination of k_Mm
    double power;   // exponent of AdjustTimeStep
    bool fsal;      // BS: y <- coords4 (the 4th stage point)
};
Tableau tableau(int solver) {
    Tableau t;
    std::memset(&t, 0, sizeof(t));
    if (solver == 0) {  // Heun-Euler, 721-750
        t.stages = 2;
        t.a[1][0] = 1.0;
        t.b[0] = 0.5;
        t.b[1] = 0.5;
        t.e[0] = 0.5;
        t.e[1] = -0.5;
        t.power = 1.0 / 2.0;
    } else if (solver == 1) {  // Bogacki-Shampine, 751-795
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
        t.fsal = true;
    } else {  // Dormand-Prince, 796-862
        t.stages = 7;
        t.a[1][0] = 1.0 / 5.0;
        t.a[2][0] = 3.0 / 40.0;
        t.a[2][1] = 9.0 / 40.0;
        t.a[3][0] = 44.0 / 45.0;
        t.a[3][1] = -56.0 / 15.0;
        t.a[3][2] = 32.0 / 9.0;
        t.a[4][0] = 19372.0 / 6561.0;
        t.a[4][1] = -25360.0 / 2187.0;
        t.a[4][2] = 64448.0 / 6561.0;
        t.a[4][3] = -212.0 / 729.0;
        t.a[5][0] = 9017.0 / 3168.0;
        t.a[5][1] = -355.0 / 33.0;
        t.a[5][2] = 46732.0 / 5247.0;
        t.a[5][3] = 49.0 / 176.0;
        t.a[5][4] = -5103.0 / 18656.0;
        t.a[6][0] = 35.0 / 384.0;
        t.a[6][2] = 500.0 / 1113.0;
        t.a[6][3] = 125.0 / 192.0;
        t.a[6][4] = -2187.0 / 6784.0;
        t.a[6][5] = 11.0 / 84.0;
        for (int s = 0; s < 6; ++s) t.b[s] = t.a[6][s];
        t.e[0] = 71.0 / 57600.0;
        t.e[2] = -71.0 / 16695.0;
        t.e[3] = 71.0 / 1920.0;
        t.e[4] = -17253.0 / 339200.0;
        t.e[5] = 22.0 / 525.0;
        t.e[6] = -1.0 / 40.0;
        t.power = 1.0 / 5.0;
    }
    return t;
}

}  // namespace

extern "C" {

orc_pairs* orc_pairs_build(const orc_params* p, double* pos) {
    // CompInteractVecs, src/interaction.cpp:88-137.
    const int64_t n = p->n;
    const double L = p->l_box;
    orc_pairs* pl = new orc_pairs();
    if (p->threads > 1 && !p->enable_coating_pot) {
        // Parallel pre-scan: with no pair inside contact distance the
        // sequential correction moves nothing, and the pair list is fully
        // determined by the positions, so it can be recomputed on the fly.
        int64_t overl = 0, cnt = 0;
        set_threads(p);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : overl, cnt)
        for (int64_t i = 0; i < n - 1; ++i)
            for (int64_t j = i + 1; j < n; ++j) {
                double r, u[3];
                if (!stream_pair(p, pos, i, j, r, u)) continue;
                ++cnt;
                if (r < p->r_mag[i] + p->r_mag[j] + 1e-10) ++overl;
            }
        if (overl == 0) {
            pl->stream = true;
            pl->count = cnt;
            pl->pos.assign(pos, pos + 3 * n);
            return pl;
        }
    }
    for (int64_t i = 0; i + 1 < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            double v[3] = {X(pos, n, j, 0) - X(pos, n, i, 0), X(pos, n, j, 1) - X(pos, n, i, 1),
                           X(pos, n, j, 2) - X(pos, n, i, 2)};
            min_image(v, L);
            const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (!(r < p->r_cut_lr)) continue;
            const double u[3] = {v[0] / r, v[1] / r, v[2] / r};
            double rs = r;
            const double dCont = p->r_mag[i] + p->r_mag[j] + 1e-10;
            if (r < dCont) {
                // move both particles half the overlap along the pre-move axis
                const double d = dCont - r;
                for (int k = 0; k < 3; ++k) {
                    X(pos, n, i, k) -= 0.5 * d * u[k];
                    X(pos, n, j, k) += 0.5 * d * u[k];
                }
                rs = dCont;
                pl->overlaps++;
            }
            pl->i.push_back(i);
            pl->j.push_back(j);
            pl->r.push_back(rs);
            pl->ux.push_back(u[0]);
            pl->uy.push_back(u[1]);
            pl->uz.push_back(u[2]);
            pl->sr.push_back(r < p->r_cut_sr ? 1 : 0);
        }
    }
    return pl;
}

void orc_pairs_free(orc_pairs* pl) { delete pl; }
int64_t orc_pairs_count(const orc_pairs* pl) { return pl->stream ? pl->count : (int64_t)pl->i.size(); }
int64_t orc_pairs_overlaps(const orc_pairs* pl) { return pl->overlaps; }
void orc_pairs_get(const orc_pairs* pl, int64_t* i, int64_t* j, double* r, double* u, int32_t* sr) {
    const size_t P = pl->i.size();
    for (size_t k = 0; k < P; ++k) {
        if (i) i[k] = pl->i[k];
        if (j) j[k] = pl->j[k];
        if (r) r[k] = pl->r[k];
        if (u) {
            u[k] = pl->ux[k];
            u[P + k] = pl->uy[k];
            u[2 * P + k] = pl->uz[k];
        }
        if (sr) sr[k] = pl->sr[k];
    }
}

void orc_kspace(const orc_params* p, const double* pos_k, const double* mom, const double* mom_sum, double* force,
                double* field) {
    kspace(p, pos_k, mom, mom_sum, force, field);
}

int orc_comp_interactions(const orc_params* p, double* pos, const double* pos_mm, const double* therm_force,
                          double* force, double* field, double* press, orc_pairs** pairs_out) {
    // CompInteractions, src/simulation.cpp:628-680 (EvalSinCos runs on the
    // positions before the overlap moves, src/simulation.cpp:475-476).
    const int64_t n = p->n;
    std::vector<double> pos_k(pos, pos + 3 * n);
    std::fill(force, force + 3 * n, 0.0);
    std::fill(field, field + 3 * n, 0.0);
    double pr[3] = {0, 0, 0};
    std::vector<double> mom;
    double msum[3];
    moments(p, pos_mm, mom, msum);
    orc_pairs* pl = orc_pairs_build(p, pos);
    real_space(p, *pl, mom.data(), force, field, pr);
    kspace(p, pos_k.data(), mom.data(), msum, force, field);
    if (p->enable_coating_pot) short_range(p, *pl, force, pr);
    if (therm_force)
        for (int64_t k = 0; k < 3 * n; ++k) force[k] += therm_force[k];
    if (press) std::memcpy(press, pr, sizeof(pr));
    if (pairs_out)
        *pairs_out = pl;
    else
        delete pl;
    return 0;
}

int orc_comp_demag_field(const orc_params* p, const orc_pairs* pl, const double* pos_k, const double* pos_mm,
                         double* field) {
    // CompDemagField, src/simulation.cpp:688-697.
    const int64_t n = p->n;
    std::fill(field, field + 3 * n, 0.0);
    std::vector<double> mom;
    double msum[3];
    moments(p, pos_mm, mom, msum);
    real_space(p, *pl, mom.data(), nullptr, field, nullptr);
    kspace(p, pos_k, mom.data(), msum, nullptr, field);
    return 0;
}

void orc_rk_coeffs(const orc_params* p, double bext_z, const double* pos_mm, const double* pos_ea,
                   const double* field, const double* therm_torque, const double* therm_field, double* k_mm,
                   double* k_ea) {
    // RKCoeffs, src/simulation.cpp:876-902; velEaConst1/2 from
    // src/simulation.cpp:166-167, 209-210.
    const int64_t n = p->n;
    for (int64_t i = 0; i < n; ++i) {
        const double m[3] = {X(pos_mm, n, i, 0), X(pos_mm, n, i, 1), X(pos_mm, n, i, 2)};
        const double e[3] = {X(pos_ea, n, i, 0), X(pos_ea, n, i, 1), X(pos_ea, n, i, 2)};
        const double dt = m[0] * e[0] + m[1] * e[1] + m[2] * e[2];
        double B[3];
        for (int d = 0; d < 3; ++d) {
            const double bext = d == 2 ? bext_z : 0.0;
            B[d] = bext + X(field, n, i, d) + p->anis_const * (e[d] * dt) + X(therm_field, n, i, d);
        }
        if (p->enable_mobilization) {
            const double c1 = 2.0 * p->anis_en * (p->vol_mag[i] / p->zeta_rot[i]);
            const double c2 = 1 / p->zeta_rot[i];
            const double T[3] = {X(therm_torque, n, i, 0), X(therm_torque, n, i, 1), X(therm_torque, n, i, 2)};
            double Te[3];
            cross(T, e, Te);
            for (int d = 0; d < 3; ++d) X(k_ea, n, i, d) = (m[d] - e[d] * dt) * (c1 * dt) + Te[d] * c2;
        } else {
            for (int d = 0; d < 3; ++d) X(k_ea, n, i, d) = 0.0;
        }
        double b2[3], b3[3];
        cross(m, B, b2);
        cross(m, b2, b3);
        for (int d = 0; d < 3; ++d) X(k_mm, n, i, d) = -p->vel_mm_const * (b2[d] + p->mag_damp * b3[d]);
    }
}

int orc_step(const orc_params* p, double* pos, double* pos_mm, double* pos_ea, double* t, double* dt,
             double* dt_used, double* mz, double* n_t, double* n_b, double* n_f, double* force_o, double* field_o,
             double* press_o, double* field_out_o, int32_t max_tries) {
    const int64_t n = p->n;
    const size_t n3 = 3 * (size_t)n;
    // DefExtField, src/simulation.cpp:546-556.
    const double bext_z = (*t < p->t_mag) ? p->mag_flux_dens : 0.0;
    // EvalSinCos + CompInteractions (src/simulation.cpp:475-476).
    std::vector<double> pos_k(pos, pos + n3);
    std::vector<double> force(n3), field(n3);
    double press[3];
    orc_pairs* pl = nullptr;
    orc_comp_interactions(p, pos, pos_mm, n_f, force.data(), field.data(), press, &pl);
    if (force_o) std::memcpy(force_o, force.data(), n3 * sizeof(double));
    if (field_o) std::memcpy(field_o, field.data(), n3 * sizeof(double));
    if (press_o) std::memcpy(press_o, press, sizeof(press));

    // IntegrationRot, src/simulation.cpp:707-868, with AdjustTimeStep 912-951.
    const Tableau tb = tableau(p->solver);
    std::vector<std::vector<double>> kMm(tb.stages, std::vector<double>(n3)), kEa(tb.stages, std::vector<double>(n3));
    std::vector<double> ymm(n3), yea(n3);
    bool cond = true, noFailed = true;
    double h = 0, used = 0;
    int tries = 0;
    int status = 0;
    while (cond) {
        h = *dt;
        // k1 uses whatever demag field is current: CompInteractions' on the
        // first try, the previous try's last stage on a retry (729, 734).
        orc_rk_coeffs(p, bext_z, pos_mm, pos_ea, field.data(), n_t, n_b, kMm[0].data(), kEa[0].data());
        for (int s = 1; s < tb.stages; ++s) {
            for (size_t q = 0; q < n3; ++q) {
                double am = 0, ae = 0;
                for (int r = 0; r < s; ++r) {
                    am += tb.a[s][r] * kMm[r][q];
                    ae += tb.a[s][r] * kEa[r][q];
                }
                ymm[q] = pos_mm[q] + h * am;
                yea[q] = pos_ea[q] + h * ae;
            }
            orc_comp_demag_field(p, pl, pos_k.data(), ymm.data(), field.data());
            orc_rk_coeffs(p, bext_z, ymm.data(), yea.data(), field.data(), n_t, n_b, kMm[s].data(), kEa[s].data());
        }
        double err = 0;
        for (int64_t i = 0; i < n; ++i) {
            double s2 = 0;
            for (int d = 0; d < 3; ++d) {
                double v = 0;
                for (int s = 0; s < tb.stages; ++s) v += tb.e[s] * kMm[s][d * n + i];
                s2 += v * v;
            }
            err = std::max(err, std::sqrt(s2));
        }
        err *= h;
        used = h;
        if (p->fixed_dt) break;
        // AdjustTimeStep
        double hn;
        if (err > p->abs_tol) {
            if (noFailed) {
                noFailed = false;
                hn = std::max(p->dt_min, h * std::max(0.1, 0.9 * std::pow(p->abs_tol / err, tb.power)));
            } else {
                hn = std::max(p->dt_min, h * 0.5);
            }
            const double sc = std::sqrt(hn / h);
            for (size_t q = 0; q < n3; ++q) {
                n_b[q] *= sc;
                n_t[q] *= sc;
                n_f[q] *= sc;
            }
            if (++tries >= max_tries) {
                status = -5;
                cond = false;
            }
        } else {
            hn = noFailed ? h * std::min(std::max(0.9 * std::pow(p->abs_tol / err, tb.power), 0.2), 5.0) : h;
            cond = false;
        }
        // the min with deltaTmax is overwritten (949-950)
        *dt = std::max(hn, p->dt_min);
    }
    // accept: y += dt_used * sum b_s k_s (BS takes the 4th stage point, 791-792)
    for (size_t q = 0; q < n3; ++q) {
        double am = 0, ae = 0;
        for (int s = 0; s < tb.stages; ++s) {
            am += tb.b[s] * kMm[s][q];
            ae += tb.b[s] * kEa[s][q];
        }
        pos_mm[q] += used * am;
        pos_ea[q] += used * ae;
    }
    // renormalise (866-867)
    for (int64_t i = 0; i < n; ++i) {
        double a = 0, b = 0;
        for (int d = 0; d < 3; ++d) {
            a += pos_mm[d * n + i] * pos_mm[d * n + i];
            b += pos_ea[d * n + i] * pos_ea[d * n + i];
        }
        a = std::sqrt(a);
        b = std::sqrt(b);
        for (int d = 0; d < 3; ++d) {
            pos_mm[d * n + i] /= a;
            pos_ea[d * n + i] /= b;
        }
    }
    if (field_out_o) std::memcpy(field_out_o, field.data(), n3 * sizeof(double));
    // IntegrationTrans (959-962) and ApplyBoundaryCondArr (610-620).
    const double L = p->l_box;
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < n; ++i) {
            double& x = pos[d * n + i];
            x += used * (force[d * n + i] / p->zeta_trans[i]);
            if (x > L / 2) x -= L;
            if (x < -L / 2) x += L;
        }
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += pos_mm[2 * n + i];
    *mz = s / n;
    *dt_used = used;
    *t += used;
    delete pl;
    return status;
}

}  // extern "C"
