// config.cpp - configuration.ini reader, derived constants, per-particle
// properties and the closed-form Ewald error models. Host code.
//
// The keys, defaults and derived formulas are those of the reference's
// Configuration class (src/configuration.cpp:6-86) and inih reader
// (src/INIReader.h); the parser here is this project's own.

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "mdsim_capi.h"

namespace {

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string rstrip(const std::string& s) {
    size_t e = s.size();
    while (e > 0 && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(0, e);
}
std::string lstrip(const std::string& s) {
    size_t b = 0;
    while (b < s.size() && std::isspace((unsigned char)s[b])) ++b;
    return s.substr(b);
}
// Position of the first of `chars`, or of an inline ';' comment that follows
// whitespace (inih's find_chars_or_comment), or npos.
size_t find_chars_or_comment(const std::string& s, const char* chars) {
    bool was_space = false;
    for (size_t i = 0; i < s.size(); ++i) {
        if (chars && std::strchr(chars, s[i])) return i;
        if (was_space && s[i] == ';') return i;
        was_space = std::isspace((unsigned char)s[i]) != 0;
    }
    return std::string::npos;
}

// inih rules: sections, name=value or name:value, full-line ';'/'#' comments,
// inline ';' comments after whitespace, continuation lines, case-insensitive
// keys, repeated keys joined with '\n'. Returns 0 or the first bad line.
int parse_ini(const std::string& text, std::map<std::string, std::string>& kv) {
    std::istringstream in(text);
    std::string raw, section, prev;
    int lineno = 0, error = 0;
    while (std::getline(in, raw)) {
        ++lineno;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        if (lineno == 1 && raw.size() >= 3 && (unsigned char)raw[0] == 0xEF && (unsigned char)raw[1] == 0xBB &&
            (unsigned char)raw[2] == 0xBF)
            raw = raw.substr(3);
        std::string start = lstrip(rstrip(raw));
        const bool indented = !raw.empty() && std::isspace((unsigned char)raw[0]);
        if (start.empty() || start[0] == ';' || start[0] == '#') continue;
        if (!prev.empty() && indented) {
            // continuation of the previous value
            size_t c = find_chars_or_comment(start, nullptr);
            std::string v = rstrip(start.substr(0, c));
            std::string& dst = kv[lower(section + "=" + prev)];
            if (!dst.empty()) dst += "\n";
            dst += v;
        } else if (start[0] == '[') {
            size_t e = find_chars_or_comment(start.substr(1), "]");
            if (e != std::string::npos && start[1 + e] == ']') {
                section = start.substr(1, e);
                prev.clear();
            } else if (!error) {
                error = lineno;
            }
        } else {
            size_t e = find_chars_or_comment(start, "=:");
            if (e != std::string::npos && (start[e] == '=' || start[e] == ':')) {
                std::string name = rstrip(start.substr(0, e));
                std::string value = lstrip(start.substr(e + 1));
                size_t c = find_chars_or_comment(value, nullptr);
                if (c != std::string::npos) value = value.substr(0, c);
                value = rstrip(value);
                std::string& dst = kv[lower(section + "=" + name)];
                if (!dst.empty()) dst += "\n";
                dst += value;
                prev = name;
            } else if (!error) {
                error = lineno;
            }
        }
    }
    return error;
}

struct Reader {
    std::map<std::string, std::string> kv;
    std::string get(const char* s, const char* n) const {
        auto it = kv.find(lower(std::string(s) + "=" + n));
        return it == kv.end() ? std::string() : it->second;
    }
    bool has(const char* s, const char* n) const { return kv.count(lower(std::string(s) + "=" + n)) != 0; }
    long integer(const char* s, const char* n, long def) const {
        std::string v = get(s, n);
        char* end;
        long x = std::strtol(v.c_str(), &end, 0);
        return end > v.c_str() ? x : def;
    }
    double real(const char* s, const char* n, double def) const {
        std::string v = get(s, n);
        char* end;
        double x = std::strtod(v.c_str(), &end);
        return end > v.c_str() ? x : def;
    }
    bool boolean(const char* s, const char* n, bool def) const {
        std::string v = lower(get(s, n));
        if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
        if (v == "false" || v == "no" || v == "off" || v == "0") return false;
        return def;
    }
};

void fill(const Reader& r, mds_config* c) {
    std::memset(c, 0, sizeof(*c));
    // src/configuration.cpp:11-72
    c->init_config_rot = (int32_t)r.integer("initial_configuration", "INIT_CONFIG_ROT", -1);
    c->init_config_trans = (int32_t)r.integer("initial_configuration", "INIT_CONFIG_TRANS", -1);
    c->read_from_file = r.boolean("initial_configuration", "READ_FROM_FILE", false);
    c->size_dist = (int32_t)r.integer("size_distribution", "SIZE_DIST", -1);
    c->enable_mobilization = r.boolean("mobilization", "ENABLE_MOBILIZATION", false);
    c->enable_therm_field = r.boolean("thermal_fluctuations", "ENABLE_THERM_FIELD", false);
    c->enable_therm_torque = r.boolean("thermal_fluctuations", "ENABLE_THERM_TORQUE", false);
    c->enable_therm_force = r.boolean("thermal_fluctuations", "ENABLE_THERM_FORCE", false);
    c->seed = r.integer("thermal_fluctuations", "seed", -1);
    c->enable_coating_pot = r.boolean("interactions", "ENABLE_COATING_POT", false);
    c->solver = (int32_t)r.integer("adaptive_timestepping_solver", "SOLVER", -1);
    c->np_dim = (int32_t)r.integer("particle_properties", "npDim", -1);
    c->mag_damp = r.real("particle_properties", "magDamp", -1);
    c->temp = r.real("particle_properties", "temp", -1);
    c->sat_mag = r.real("particle_properties", "satMag", -1);
    c->anis_en = r.real("particle_properties", "anisEn", -1);
    c->r_mag_mean = r.real("assemble_properties", "rMagMean", -1);
    c->d_shell = r.real("assemble_properties", "dShell", -1);
    c->vol_frac = r.real("assemble_properties", "volFrac", -1);
    c->dens_surf_mol = r.real("steric_repulsion", "densSurfMol", -1);
    c->hamaker = r.real("vdw_attraction", "hamaker", -1);
    c->z_valency = r.real("electrostat_repulsion", "z_valency", -1);
    c->epsilon_r = r.real("electrostat_repulsion", "epsilon_r", -1);
    c->gamma_0 = r.real("electrostat_repulsion", "gamma_0", -1);
    c->conc = r.real("electrostat_repulsion", "conc", -1);
    c->t_mag = r.real("time_settings", "tMag", -1);
    c->t_relax = r.real("time_settings", "tRelax", -1);
    c->delta_t_init = r.real("time_settings", "deltaTinit", -1);
    c->mag_flux_dens = r.real("external_magnetic_field_t", "magFluxDens", -1);
    c->abs_tol = r.real("solver_settings", "absTol", -1);
    c->delta_t_min = r.real("solver_settings", "deltaTmin", -1);
    c->delta_t_max = r.real("solver_settings", "deltaTmax", -1);
    c->err_tol_ewald = r.real("long_range_interactions", "errTolEwald", -1);
    c->err_tol_sr = r.real("short_range_interactions", "errTolSR", -1);
    c->np_coords = r.real("output_settings", "npCoords", -1);
    c->my_pi = r.real("physical_constants", "my_pi", -1);
    c->kB = r.real("physical_constants", "kB", -1);
    c->gyro_mr = r.real("physical_constants", "gyroMr", -1);
    c->avogadro = r.real("physical_constants", "avogadro", -1);
    c->epsilon_0 = r.real("physical_constants", "epsilon_0", -1);
    c->el = r.real("physical_constants", "el", -1);
    c->fac_ewald = r.real("faster_calculation", "facEwald", -1);
    // keys this library adds
    c->numb_part_override = r.integer("mi355x", "numbPart", -1);
    c->viscosity = r.real("mi355x", "viscosity", -1);
    c->r_mag_median = r.real("mi355x", "rMagMedian", -1);
    c->sigma_log = r.real("mi355x", "sigmaLog", -1);
    c->fixed_dt = r.boolean("mi355x", "fixedDt", false);
    std::string prec = lower(r.get("mi355x", "precision"));
    c->precision = (prec == "mixed" || prec == "1") ? MDS_PREC_MIXED : MDS_PREC_FP64;
    c->ewald_alpha_l = r.real("mi355x", "ewaldAlphaL", -1);
    c->ewald_rc_frac = r.real("mi355x", "ewaldRc", -1);
    c->ewald_nc = (int32_t)r.integer("mi355x", "ewaldNc", -1);
    mds_config_derive(c);
}

inline double sq(double x) { return x * x; }

// IntegralEffDiameter / funEffDiameter, src/mathFun.cpp:27-54: Simpson rule,
// 1000 panels, for the hard-sphere-equivalent diameter of the steric shell.
double fun_eff_diameter(const mds_config* c, double dMag, double r) {
    const double dDelta = dMag + 2 * c->eff_len_surf_mol;
    return 1 - std::exp(-c->my_pi * dMag * dMag * c->dens_surf_mol / (2 * c->eff_len_surf_mol) *
                        (dDelta - r * (std::log(dDelta / (r + 1e-20)) + 1)));
}
double integral_eff_diameter(const mds_config* c, double dMag) {
    const double hi = dMag + 2 * c->eff_len_surf_mol;
    const int steps = 1000;
    const double h = hi / steps;
    double area = 0;
    for (int i = 0; i < steps; i++)
        area += h / 6 * (fun_eff_diameter(c, dMag, i * h) + 4 * fun_eff_diameter(c, dMag, (i + 0.5) * h) +
                         fun_eff_diameter(c, dMag, (i + 1) * h));
    return area;
}

}  // namespace

extern "C" {

void mds_config_derive(mds_config* c) {
    // src/configuration.cpp:39, 64, 71-84
    c->eff_len_surf_mol = 0.2 * c->r_mag_mean;
    c->mu0 = 4.0 * c->my_pi * std::pow(10, -7);
    c->numb_part = c->numb_part_override > 0 ? c->numb_part_override
                                             : (int64_t)c->np_dim * c->np_dim * c->np_dim;
    c->vel_mm_const = c->gyro_mr / (1.0 + std::pow(c->mag_damp, 2.0));
    c->anis_const = 2.0 * c->anis_en / c->sat_mag;
    c->ster_rep_fac = 4 * c->kB * c->temp * c->my_pi * c->dens_surf_mol / (2 * c->eff_len_surf_mol);
    c->gamma_el = std::tanh(c->z_valency * c->el * c->gamma_0 / (4 * c->kB * c->temp));
    c->epsilon = c->epsilon_0 * c->epsilon_r;
    c->kappa = std::sqrt(c->el * c->el * c->avogadro * c->conc * c->z_valency * c->z_valency /
                         (c->epsilon * c->kB * c->temp));
    c->electrostat_rep_fac = c->kappa * 64 * c->my_pi * c->epsilon *
                             std::pow(c->kB * c->temp / (c->z_valency * c->el), 2.0) * c->gamma_el * c->gamma_el;
    c->ir_pi = 1 / std::sqrt(c->my_pi);
    c->vis = c->viscosity > 0 ? c->viscosity : (2.414e-5) * std::pow(10, 247.8 / (c->temp - 140));
    c->sigma = std::sqrt(0.1) * 2.0 * c->r_mag_mean;
    c->t_end = c->t_mag + c->t_relax;
}

int64_t mds_config_sizeof(void) { return (int64_t)sizeof(mds_config); }

mds_status mds_config_from_string(const char* text, mds_config* out) {
    if (!text || !out) return MDS_EINVAL;
    Reader r;
    if (parse_ini(text, r.kv) != 0) return MDS_ECONFIG;
    fill(r, out);
    return MDS_OK;
}

mds_status mds_config_from_ini(const char* path, mds_config* out) {
    if (!path || !out) return MDS_EINVAL;
    std::ifstream f(path, std::ios::binary);
    if (!f) return MDS_ECONFIG;
    std::stringstream ss;
    ss << f.rdbuf();
    return mds_config_from_string(ss.str().c_str(), out);
}

mds_status mds_particle_props(const mds_config* c, int64_t n, const double* r_mag, double* r_hydr, double* vol_mag,
                              double* mag_mom, double* zeta_rot, double* zeta_trans) {
    // InitParticleProps, src/simulation.cpp:135-211, for given core radii.
    if (!c || n <= 0 || !r_mag) return MDS_EINVAL;
    const double pi = c->my_pi;
    for (int64_t i = 0; i < n; ++i) {
        const double r = r_mag[i];
        double rh;
        if (c->enable_coating_pot) {
            const double dH = integral_eff_diameter(c, 2 * r);
            rh = dH >= 2 * r ? dH / 2.0 : r;
        } else {
            rh = r + c->d_shell;
        }
        const double vm = 4.0 / 3.0 * pi * std::pow(r, 3.0);
        if (r_hydr) r_hydr[i] = rh;
        if (vol_mag) vol_mag[i] = vm;
        if (mag_mom) mag_mom[i] = c->sat_mag * vm;
        if (zeta_rot) zeta_rot[i] = 8.0 * pi * c->vis * std::pow(rh, 3.0);
        if (zeta_trans) zeta_trans[i] = 6.0 * pi * c->vis * rh;
    }
    return MDS_OK;
}

double mds_box_length(const mds_config* c, int64_t n, const double* r_hydr) {
    // src/simulation.cpp:56 (volHydr from src/simulation.cpp:162, 205)
    double v = 0;
    for (int64_t i = 0; i < n; ++i) v += 4.0 / 3.0 * c->my_pi * std::pow(r_hydr[i], 3.0);
    return std::pow(v / c->vol_frac, 1.0 / 3.0);
}

mds_status mds_cutoff_sr(const mds_config* c, int64_t n, const double* r_mag, const double* r_hydr,
                         const double* mag_mom, double* r_cut_sr) {
    // InitCutOffSR, src/interaction.cpp:20-57; FunCutOffError 63-77.
    if (!c || n <= 0 || !r_mag || !r_cut_sr) return MDS_EINVAL;
    double rMax = r_mag[0], mMax = mag_mom ? mag_mom[0] : 0, hMax = r_hydr ? r_hydr[0] : 0;
    for (int64_t i = 1; i < n; ++i) {
        rMax = std::max(rMax, r_mag[i]);
        if (mag_mom) mMax = std::max(mMax, mag_mom[i]);
        if (r_hydr) hMax = std::max(hMax, r_hydr[i]);
    }
    if (!c->enable_coating_pot) {
        *r_cut_sr = 2 * rMax;
        return MDS_OK;
    }
    const double c1 = 128 * c->hamaker * std::pow(rMax, 6.0) * c->my_pi / (18 * c->mu0 * mMax * mMax);
    const double delta = c->err_tol_sr;
    auto f = [&](double x) { return c1 * x / std::pow(x * x - 4 * rMax * rMax, 2.0) - delta; };
    auto df = [&](double x) { return c1 * (4 * rMax * rMax + 3 * x * x) / std::pow(4 * rMax * rMax - x * x, 3.0); };
    double x = 1.1 * 2 * rMax;
    double e = 100;
    int it = 0;
    do {
        x -= f(x) / df(x);
        e = std::fabs(f(x));
        if (++it > 10000 || !std::isfinite(x)) return MDS_ENOCONV;
    } while (e > 1e-7);
    if (x < 2 * hMax) return MDS_ECONFIG;  // the reference prints and exits here (47-51)
    *r_cut_sr = x;
    return MDS_OK;
}

double mds_ewald_real_error(const mds_config* c, int64_t n, double M, double f_al, double l_box, double r_cut,
                            double x) {
    // FunRealError, src/interaction.cpp:330-338
    const double a = x * r_cut;
    const double b = M / std::sqrt(std::pow(l_box, 3.0) * std::pow(r_cut, 9.0) * (double)n * std::pow(x, 4.0));
    const double cc = 4.0 * std::pow(a, 4.0) + 6.0 * std::pow(a, 2.0) + 3.0;
    const double d = 8.0 * std::pow(a, 6.0) + 20.0 * std::pow(a, 4.0) + 30.0 * std::pow(a, 2.0) + 15.0;
    return c->fac_ewald * b * std::sqrt(13.0 / 6.0 * cc * cc + 2.0 / 15.0 * d * d - 13.0 / 15.0 * cc * d) *
               std::exp(-a * a) / f_al -
           c->err_tol_ewald / std::sqrt(2.0);
}

double mds_ewald_rez_error(const mds_config* c, int64_t n, double M, double f_al, double l_box, double alpha,
                           double x) {
    // FunRezError, src/interaction.cpp:344-347
    return c->fac_ewald * 8.0 * M * c->my_pi / std::pow(l_box, 3.0) * alpha *
               std::sqrt(2.0 * c->my_pi * x * x * x / (15.0 * (double)n)) *
               std::exp(-std::pow(c->my_pi * x / (alpha * l_box), 2.0)) / f_al -
           c->err_tol_ewald / std::sqrt(2.0);
}

mds_status mds_ewald_predicted_errors(const mds_config* c, int64_t n, const double* mag_mom, const double* r_hydr,
                                      double l_box, double alpha, double r_cut, int32_t nc, double* force_rel,
                                      double* torque_rel) {
    // The error estimates CompEwaldParams prints for the parameters it chose
    // (src/interaction.cpp:311-323), relative to the contact force and torque
    // of the smallest particles (178-180).
    if (!c || n <= 0 || !mag_mom || !r_hydr || l_box <= 0 || alpha <= 0 || r_cut <= 0) return MDS_EINVAL;
    double M = 0, mMin = mag_mom[0], hMin = r_hydr[0];
    for (int64_t i = 0; i < n; ++i) {
        M += std::pow(mag_mom[i], 2.0);
        mMin = std::min(mMin, mag_mom[i]);
        hMin = std::min(hMin, r_hydr[i]);
    }
    const double pi = c->my_pi, f = c->fac_ewald, N = (double)n, k = (double)nc;
    const double fAl = 3 * c->mu0 * std::pow(mMin, 2.0) / (2.0 * pi * std::pow(2.0 * hMin, 4.0));
    const double tAl = std::pow(mMin, 2.0) * c->mu0 / (16.0 * pi * std::pow(hMin, 3.0));
    const double ar = alpha * r_cut;
    const double cc = 4.0 * std::pow(ar, 4.0) + 6.0 * std::pow(ar, 2.0) + 3.0;
    const double dc = 8.0 * std::pow(ar, 6.0) + 20.0 * std::pow(ar, 4.0) + 30.0 * std::pow(ar, 2.0) + 15.0;
    const double bc = 2.0 * std::pow(ar, 2.0) + 1.0;
    const double gk = std::exp(-std::pow(pi * k / (alpha * l_box), 2.0));
    const double eRF = f * M / std::sqrt(std::pow(l_box, 3.0) * std::pow(alpha, 4.0) * std::pow(r_cut, 9.0) * N) *
                       std::sqrt(13.0 / 6.0 * cc * cc + 2.0 / 15.0 * dc * dc - 13.0 / 15.0 * cc * dc) *
                       std::exp(-ar * ar);
    const double eKF = f * 8.0 * pi * M / std::pow(l_box, 3.0) * alpha * std::sqrt(2.0 * pi * k * k * k / (15.0 * N)) * gk;
    const double eRT = f * M / std::sqrt(std::pow(l_box, 3.0) * std::pow(alpha, 4.0) * std::pow(r_cut, 7.0) * N) *
                       std::sqrt(0.5 * bc * bc + 0.2 * cc * cc) * std::exp(-ar * ar);
    const double eKT = f * 4.0 * M / std::pow(l_box, 2.0) * alpha * std::sqrt(pi * k / (5.0 * N)) * gk;
    if (force_rel) *force_rel = std::sqrt(eKF * eKF + eRF * eRF) / fAl;
    if (torque_rel) *torque_rel = std::sqrt(eKT * eKT + eRT * eRT) / tAl;
    return MDS_OK;
}

mds_status mds_ewald_params_for_rc(const mds_config* c, int64_t n, const double* mag_mom, const double* r_hydr,
                                   double l_box, double r_cut, double* alpha, double* kc) {
    // The deterministic part of CompEwaldParams (src/interaction.cpp:179-257):
    // alpha by a linear scan of FunRealError over [0, 2e8] (10000 points) and
    // linear interpolation of the sign change; k_c by a scan over [0, 30]
    // starting at the maximum of FunRezError; k_c >= 2.
    if (!c || n <= 0 || !mag_mom || !r_hydr || !alpha || !kc) return MDS_EINVAL;
    double M = 0, mMin = mag_mom[0], hMin = r_hydr[0];
    for (int64_t i = 0; i < n; ++i) {
        M += std::pow(mag_mom[i], 2.0);
        mMin = std::min(mMin, mag_mom[i]);
        hMin = std::min(hMin, r_hydr[i]);
    }
    const double fAl = 3 * c->mu0 * std::pow(mMin, 2.0) / (2.0 * c->my_pi * std::pow(2.0 * hMin, 4.0));
    const int steps = 10000;
    // Eigen's setLinSpaced: low + i * step, with the last point exactly high
    auto lin = [&](double lo, double hi, int i) {
        return i == steps - 1 ? hi : lo + (double)i * ((hi - lo) / (steps - 1));
    };
    int i1 = 0, i2 = 0;
    double prev = 0, cur = 0;
    for (int i = 0; i < steps; i++) {
        cur = mds_ewald_real_error(c, n, M, fAl, l_box, r_cut, lin(0, 2e8, i));
        if (cur < 0) {
            i2 = i;
            i1 = i - 1;
            break;
        }
        prev = cur;
    }
    if (i1 <= 0 || i2 <= 0) return MDS_ENOCONV;
    {
        const double x1 = lin(0, 2e8, i1), x2 = lin(0, 2e8, i2);
        *alpha = -prev / (cur - prev) * (x2 - x1) + x1;
    }
    // The reference scans k_c over [0, 30] only; where that range holds no
    // root (small r_c, large N) it would read stale indices, so the scan is
    // widened to [0, 120] there.
    double k = -1;
    for (double kmax : {30.0, 120.0}) {
        double best = -1e300;
        int maxIdx = 0;
        for (int i = 0; i < steps; i++) {
            double v = mds_ewald_rez_error(c, n, M, fAl, l_box, *alpha, lin(0, kmax, i));
            if (v > best) {
                best = v;
                maxIdx = i;
            }
        }
        if (best < 0) return MDS_ENOCONV;  // the reference exits here (230-236)
        double fprev = 0, fcur = 0;
        i1 = i2 = 0;
        for (int i = maxIdx; i < steps; i++) {
            fcur = mds_ewald_rez_error(c, n, M, fAl, l_box, *alpha, lin(0, kmax, i));
            if (fcur < 0) {
                i2 = i;
                i1 = i - 1;
                fprev = mds_ewald_rez_error(c, n, M, fAl, l_box, *alpha, lin(0, kmax, i1));
                break;
            }
        }
        if (i2 == 0) continue;
        const double x1 = lin(0, kmax, i1), x2 = lin(0, kmax, i2);
        k = -fprev / (fcur - fprev) * (x2 - x1) + x1;
        break;
    }
    if (k < 0) return MDS_ENOCONV;
    *kc = k < 2.0 ? 2.0 : k;
    return MDS_OK;
}

}  // extern "C"
