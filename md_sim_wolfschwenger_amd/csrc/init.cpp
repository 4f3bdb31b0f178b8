// init.cpp - initial suspensions (host code).
//
// InitParticleProps' radii (src/simulation.cpp:132-214) and
// InitCoordsFromSketch (src/simulation.cpp:220-351) with this library's own
// seeded generator, xoshiro256** (Blackman and Vigna, "Scrambled linear
// pseudorandom number generators", ACM TOMS 47, 2021) seeded through
// splitmix64, in place of EigenRand's Mersenne twister. The random
// placement keeps the reference's acceptance rule but finds the particles a
// candidate could overlap through a cell list instead of scanning all
// earlier particles (src/simulation.cpp:241-264 is O(N^2) per accepted
// particle).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mdsim_capi.h"

namespace {

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            w = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double balanced() { return 2.0 * uniform() - 1.0; }       // [-1, 1)
    double normal() {
        // Box-Muller, one of the pair
        const double u1 = ((next() >> 11) + 1) * 0x1.0p-53;  // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

}  // namespace

extern "C" {

mds_status mds_init_radii(const mds_config* cfg, int64_t n, uint64_t seed, double* r_mag) {
    if (!cfg || n <= 0 || !r_mag) return MDS_EINVAL;
    Xoshiro g(seed);
    if (cfg->sigma_log > 0) {
        // lognormal by median and sigma of ln r (added keys rMagMedian, sigmaLog)
        const double med = cfg->r_mag_median > 0 ? cfg->r_mag_median : cfg->r_mag_mean;
        for (int64_t i = 0; i < n; ++i) r_mag[i] = med * std::exp(cfg->sigma_log * g.normal());
    } else if (cfg->size_dist == 1) {
        // SIZE_DIST_LOG: mean and standard deviation of r (src/simulation.cpp:180-183)
        const double m = cfg->r_mag_mean, s = cfg->sigma;
        if (!(m > 0) || !(s > 0)) return MDS_ECONFIG;
        const double mu = std::log(m * m / std::sqrt(m * m + s * s));
        const double sl = std::sqrt(std::log(1.0 + (s / m) * (s / m)));
        for (int64_t i = 0; i < n; ++i) r_mag[i] = std::exp(mu + sl * g.normal());
    } else {
        if (!(cfg->r_mag_mean > 0)) return MDS_ECONFIG;
        for (int64_t i = 0; i < n; ++i) r_mag[i] = cfg->r_mag_mean;
    }
    return MDS_OK;
}

mds_status mds_init_positions(const mds_config* cfg, int64_t n, const double* r_hydr, double l_box, uint64_t seed,
                              int32_t periodic, double* pos) {
    if (!cfg || n <= 0 || !r_hydr || !pos || !(l_box > 0)) return MDS_EINVAL;
    const double L = l_box;
    if (cfg->init_config_trans == 2) {
        // grid of npDim^3 (src/simulation.cpp:268-289): z fastest
        const int64_t d = cfg->np_dim;
        if (d <= 0 || d * d * d != n) return MDS_EINVAL;
        const double gap = L / (double)d;
        int64_t ip = 0;
        for (int64_t ix = 0; ix < d; ++ix)
            for (int64_t iy = 0; iy < d; ++iy)
                for (int64_t iz = 0; iz < d; ++iz, ++ip) {
                    pos[ip] = (ix + 0.5) * gap - L / 2.0;
                    pos[n + ip] = (iy + 0.5) * gap - L / 2.0;
                    pos[2 * n + ip] = (iz + 0.5) * gap - L / 2.0;
                }
        return MDS_OK;
    }
    // random sequential addition in index order (src/simulation.cpp:238-265):
    // candidate x_d = u_d (L/2 - r_i), u uniform in [-1, 1); rejected while
    // closer than r_i + r_j to an earlier particle. periodic != 0 draws over
    // the whole box and measures distances by minimum image instead (no
    // overlap through the faces).
    double rmax = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(r_hydr[i] > 0) || !(r_hydr[i] < L / 2)) return MDS_EINVAL;
        rmax = std::max(rmax, r_hydr[i]);
    }
    int64_t nc = (int64_t)std::floor(L / (2 * rmax));
    nc = std::max<int64_t>(1, std::min<int64_t>(nc, (int64_t)std::cbrt(8.0 * (double)n) + 1));
    if (periodic && nc < 3) nc = 1;
    const double cell = L / (double)nc;
    std::vector<int64_t> head((size_t)(nc * nc * nc), -1), next((size_t)n, -1);
    auto coord = [&](double x) {
        int64_t c = (int64_t)std::floor((x + L / 2) / cell);
        return std::min(std::max<int64_t>(c, 0), nc - 1);
    };
    Xoshiro g(seed);
    const int64_t max_tries = 100000000;
    int64_t tries = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double ri = r_hydr[i];
        const double ext = periodic ? L / 2 : L / 2 - ri;
        for (;;) {
            if (++tries > max_tries) return MDS_ECAPACITY;  // volume fraction too high for sequential addition
            double p[3];
            for (double& v : p) v = g.balanced() * ext;
            const int64_t cx = coord(p[0]), cy = coord(p[1]), cz = coord(p[2]);
            bool ok = true;
            const int64_t R = nc == 1 ? 0 : 1;
            for (int64_t dz = -R; dz <= R && ok; ++dz)
                for (int64_t dy = -R; dy <= R && ok; ++dy)
                    for (int64_t dx = -R; dx <= R && ok; ++dx) {
                        int64_t jx = cx + dx, jy = cy + dy, jz = cz + dz;
                        if (periodic) {
                            jx = (jx + nc) % nc;
                            jy = (jy + nc) % nc;
                            jz = (jz + nc) % nc;
                        } else if (jx < 0 || jy < 0 || jz < 0 || jx >= nc || jy >= nc || jz >= nc) {
                            continue;
                        }
                        for (int64_t j = head[(size_t)((jz * nc + jy) * nc + jx)]; j >= 0; j = next[(size_t)j]) {
                            double d2 = 0;
                            for (int k = 0; k < 3; ++k) {
                                double v = pos[k * n + j] - p[k];
                                if (periodic) v -= L * std::nearbyint(v / L);
                                d2 += v * v;
                            }
                            const double lim = ri + r_hydr[j];
                            if (d2 < lim * lim) {
                                ok = false;
                                break;
                            }
                        }
                    }
            if (!ok) continue;
            for (int k = 0; k < 3; ++k) pos[k * n + i] = p[k];
            const size_t c = (size_t)((cz * nc + cy) * nc + cx);
            next[(size_t)i] = head[c];
            head[c] = i;
            break;
        }
    }
    return MDS_OK;
}

mds_status mds_init_orientations(const mds_config* cfg, int64_t n, uint64_t seed, double* pos_mm, double* pos_ea) {
    if (!cfg || n <= 0 || !pos_mm || !pos_ea) return MDS_EINVAL;
    // azimuth 2 pi u, elevation 2 asin(sqrt(u)) (src/simulation.cpp:296-312);
    // all azimuths are drawn before the elevations, as there
    Xoshiro g(seed);
    std::vector<double> az((size_t)n), el((size_t)n);
    for (auto& a : az) a = (g.balanced() + 1) * 0.5 * 2 * cfg->my_pi;
    for (auto& e : el) e = std::asin(std::sqrt((g.balanced() + 1) * 0.5)) * 2.0;
    const int rot = cfg->init_config_rot;
    for (int64_t i = 0; i < n; ++i) {
        double e[3];
        if (rot == 3) {
            e[0] = e[1] = 0;
            e[2] = 1;
        } else {
            e[0] = std::sin(el[i]) * std::cos(az[i]);
            e[1] = std::sin(el[i]) * std::sin(az[i]);
            e[2] = std::cos(el[i]);
        }
        for (int k = 0; k < 3; ++k) {
            pos_ea[k * n + i] = e[k];
            // INIT_CONFIG_ROT 1: moments along z; 2 and 3: along the easy axes
            pos_mm[k * n + i] = rot == 1 ? (k == 2 ? 1.0 : 0.0) : e[k];
        }
    }
    return MDS_OK;
}

// InitCoordsFromSim (src/simulation.cpp:357-382) over openData's format
// (src/postprocessing.cpp:37-75: whitespace-separated numbers, one matrix row
// per line). The file is what WriteCoords2TXT writes (src/postprocessing.cpp:
// 80-100: N x 10, pos | posMm | posEa | tInt in (0, 9)), one block or, as in
// coords.txt, a block per progress point separated by blank lines; the last
// block is read. The reference reads posMm from columns 6-8 and posEa from
// 9-11, which a 10-column file does not have; here they come from 3-5 and
// 6-8, where WriteCoords2TXT puts them.
mds_status mds_init_coords_from_sim(const char* path, int64_t n, double* pos, double* pos_mm, double* pos_ea,
                                    double* t_int) {
    if (!path || n <= 0 || !pos || !pos_mm || !pos_ea) return MDS_EINVAL;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return MDS_EINVAL;
    std::vector<std::vector<double>> block, last;
    std::string line;
    bool bad = false;
    auto end_block = [&]() {
        if (!block.empty()) last.swap(block);
        block.clear();
    };
    for (int ch = 0; !bad;) {
        line.clear();
        while ((ch = std::fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
        if (line.empty() && ch == EOF) break;
        std::vector<double> row;
        const char* q = line.c_str();
        for (;;) {
            while (*q == ' ' || *q == '\t' || *q == '\r') ++q;
            if (!*q) break;
            char* e = nullptr;
            const double v = std::strtod(q, &e);
            if (e == q) {
                bad = true;
                break;
            }
            row.push_back(v);
            q = e;
        }
        if (row.empty()) end_block(); else block.push_back(std::move(row));
        if (ch == EOF) break;
    }
    std::fclose(f);
    end_block();
    if (bad || (int64_t)last.size() != n) return MDS_EINVAL;
    for (const auto& row : last)
        if (row.size() != 10) return MDS_EINVAL;
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            pos[d * n + i] = last[i][d];
            pos_mm[d * n + i] = last[i][3 + d];
            pos_ea[d * n + i] = last[i][6 + d];
        }
    if (t_int) *t_int = last[0][9];
    return MDS_OK;
}

}  // extern "C"
