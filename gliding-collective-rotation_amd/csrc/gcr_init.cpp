// gcr_init.cpp -- host-side initial conditions: straight chains (or single
// beads) placed by random sequential addition without overlaps, from a seed.
// The reference instead grows dot-like chains (populations.py:530-557,
// control.py:153-219); this placement gives the "random non-overlapping init
// from a seed" of the benchmark workloads directly.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gcr.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014) as the placement's random source
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double wrapd(double d, double L)
{
    if (d < -0.5 * L) d += L;
    if (d >= 0.5 * L) d -= L;
    return d;
}

}  // namespace

extern "C" int gcr_init_chains(int32_t dim, double L, int64_t n_chains, const int32_t* chain_len,
                               double bond, double min_dist, double min_dist_floor, uint64_t seed,
                               int64_t max_tries, double* pos, double* tangent,
                               double* min_dist_used)
{
    if ((dim != 2 && dim != 3) || !(L > 0) || n_chains < 1 || !chain_len || !pos || !tangent ||
        !(min_dist > 0) || !(min_dist < 0.5 * L) || !(min_dist_floor <= min_dist))
        return GCR_ERR_ARG;
    int64_t n = 0;
    for (int64_t c = 0; c < n_chains; ++c) {
        if (chain_len[c] < 1) return GCR_ERR_ARG;
        n += chain_len[c];
    }
    int nc = (int)std::floor(L / min_dist);
    if (nc < 3) nc = 1;  // one cell holding every bead when a 3-cell stencil does not fit
    const double h = L / nc;
    const int64_t ncell = dim == 3 ? (int64_t)nc * nc * nc : (int64_t)nc * nc;
    std::vector<int64_t> head(ncell, -1), next(n, -1);
    Rng rng{seed ^ 0x6A09E667F3BCC909ull};
    // the exclusion distance in force; lowered by 10% steps (not below
    // min_dist_floor) whenever a chain finds no free place in max_tries
    double cur = min_dist;
    double md2 = cur * cur;
    auto cell_of = [&](const double* x) {
        int g[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) {
            int v = (int)(x[i] / h);
            g[i] = v < 0 ? 0 : (v >= nc ? nc - 1 : v);
        }
        return ((int64_t)g[2] * nc + g[1]) * nc + g[0];
    };
    const int span = (nc >= 3) ? 1 : 0;
    int64_t b0 = 0;
    std::vector<double> cand;
    for (int64_t c = 0; c < n_chains; ++c) {
        const int len = chain_len[c];
        cand.assign((size_t)len * dim, 0.0);
        double t[3] = {0, 0, 0};
        bool placed = false;
        for (int64_t tries = 0; tries < max_tries && !placed; ++tries) {
            double ctr[3];
            for (int i = 0; i < dim; ++i) ctr[i] = rng.uniform() * L;
            if (dim == 2) {
                double th = 2.0 * M_PI * rng.uniform();
                t[0] = std::cos(th);
                t[1] = std::sin(th);
            } else {
                double z = 2.0 * rng.uniform() - 1.0, ph = 2.0 * M_PI * rng.uniform();
                double s = std::sqrt(1.0 - z * z);
                t[0] = s * std::cos(ph);
                t[1] = s * std::sin(ph);
                t[2] = z;
            }
            // bead 0 leads: x_j = centre + ((len-1)/2 - j) * bond * t
            bool ok = true;
            for (int j = 0; j < len && ok; ++j) {
                double* x = &cand[(size_t)j * dim];
                for (int i = 0; i < dim; ++i) {
                    double v = std::fmod(ctr[i] + (0.5 * (len - 1) - j) * bond * t[i], L);
                    if (v < 0) v += L;
                    if (v >= L) v -= L;
                    x[i] = v;
                }
                int64_t cc = cell_of(x);
                int g[3] = {(int)(cc % nc), (int)((cc / nc) % nc), (int)(cc / ((int64_t)nc * nc))};
                for (int dz = (dim == 3 ? -span : 0); dz <= (dim == 3 ? span : 0) && ok; ++dz)
                    for (int dy = -span; dy <= span && ok; ++dy)
                        for (int dx = -span; dx <= span && ok; ++dx) {
                            int gx = (g[0] + dx + nc) % nc, gy = (g[1] + dy + nc) % nc;
                            int gz = dim == 3 ? (g[2] + dz + nc) % nc : 0;
                            for (int64_t q = head[((int64_t)gz * nc + gy) * nc + gx]; q >= 0;
                                 q = next[q]) {
                                double d2 = 0;
                                for (int i = 0; i < dim; ++i) {
                                    double d = wrapd(pos[q * dim + i] - x[i], L);
                                    d2 += d * d;
                                }
                                if (d2 < md2) {
                                    ok = false;
                                    break;
                                }
                            }
                        }
            }
            if (!ok) continue;
            placed = true;
            for (int j = 0; j < len; ++j) {
                int64_t b = b0 + j;
                std::memcpy(&pos[b * dim], &cand[(size_t)j * dim], sizeof(double) * dim);
                for (int i = 0; i < dim; ++i) tangent[b * dim + i] = t[i];
                int64_t cc = cell_of(&pos[b * dim]);
                next[b] = head[cc];
                head[cc] = b;
            }
        }
        if (!placed) {
            if (cur * 0.9 < min_dist_floor) return GCR_ERR_STATE;
            cur *= 0.9;
            md2 = cur * cur;
            --c;  // retry this chain at the smaller distance
            continue;
        }
        b0 += len;
    }
    if (min_dist_used) *min_dist_used = cur;
    return GCR_OK;
}
