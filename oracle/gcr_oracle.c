/*
 * gcr_oracle.c -- CPU FP64 restatement of the reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline). Nothing in the
 * product library links this file. See gcr_oracle.h for the interface and the
 * reference lines each function restates.
 */
#include "gcr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double orc_pymod(double a, double b)
{
    /* CPython float_rem: fmod, then move into the divisor's sign. */
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

void orc_wca_params(double r, double eps, double shift, double cutoff_factor, double out[4])
{
    /* populations.py:48-58 */
    double cut = cutoff_factor * r - shift;
    double s6 = pow(r, 6.0);
    double fac = 12.0 * s6 * eps;
    out[0] = fac;
    out[1] = fac * s6;
    out[2] = shift;
    out[3] = cut;
}

double orc_wca_coeff(double ds, const double p[4])
{
    /* forces.py:13-15 */
    return pow(ds, -4.0) * (p[0] - p[1] * pow(ds, -3.0));
}

double orc_inbounds_diff(int dim, const double* a, const double* b, double hb, double L,
                         double* out)
{
    /* forces.py:44-54 */
    double ds = 0.0;
    for (int i = 0; i < dim; ++i) {
        out[i] = a[i] - b[i];
        if (out[i] < -hb)
            out[i] += L;
        else if (out[i] >= hb)
            out[i] -= L;
        ds += out[i] * out[i];
    }
    return ds;
}

static double sgn(double x) { return (x > 0) - (x < 0); }

void orc_chain_forces(int dim, int n, const double* pos, const int64_t* chain_id, double L,
                      const double sp[6], const double* rest_lengths, double* tangent,
                      double* F)
{
    /* forces.py:285-402 (Hookean, or FENE when sp[3] == 1: forces.py:331-341
     * with the maximal stretch sp[5] * rest length, forces.py:311-312, and
     * the reference's guard, which compares stretch^2 with that length
     * itself). The reference walks beads in index
     * order, keeps the two most recent normalised bond vectors in a ring
     * (ti = bond b1->b1+1, tj = bond b1-1->b1) and adds each term to the
     * beads it touches as soon as it is known. */
    double half = L / 2.0;
    int bend = (sp[1] != 0.0);
    int orient = (sp[4] != 0.0);
    double tg[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int ti = 0, tj = 1, new_chain = 1;
    int64_t ci = 0;
    double a = 0.0;
    for (int b1 = 0; b1 < n; ++b1) {
        if (new_chain) {
            ti = 0;
            tj = 1;
            ci = chain_id[b1];
            a = rest_lengths[ci];
        }
        int end = (b1 == n - 1) || (chain_id[b1 + 1] != ci);
        if (!end) {
            double ds = orc_inbounds_diff(dim, pos + b1 * dim, pos + (b1 + 1) * dim, half, L,
                                          tg[ti]);
            if (ds != 0.0) {
                double inv = pow(ds, -0.5);
                double mag;
                if (sp[3] != 0.0) {
                    double m = sp[5] * a;
                    double st = sqrt(ds) - a;
                    double den = (st * st != m && m != 0.0) ? fabs(st * st / (m * m) - 1.0) : 0.0;
                    mag = -sp[0] * inv * st * 1.0 / den;
                } else {
                    mag = sp[0] * (a * inv - 1.0);
                }
                for (int i = 0; i < dim; ++i) {
                    double f = tg[ti][i] * mag;
                    F[b1 * dim + i] += f;
                    F[(b1 + 1) * dim + i] -= f;
                    tg[ti][i] *= inv;
                }
            }
        }
        if (orient || bend) {
            if (new_chain) {
                if (orient)
                    for (int i = 0; i < dim; ++i) tangent[b1 * dim + i] = tg[ti][i];
            } else if (end) {
                if (orient)
                    for (int i = 0; i < dim; ++i) tangent[b1 * dim + i] = tg[tj][i];
            } else {
                double dot = 0.0;
                for (int i = 0; i < dim; ++i) dot += tg[ti][i] * tg[tj][i];
                if (-1.0 < dot && dot < 1.0) {
                    if (bend) {
                        double phi = acos(dot);
                        if (sp[2] != 0.0)
                            phi += sp[2] * sgn(-tg[ti][0] * tg[tj][1] + tg[ti][1] * tg[tj][0]);
                        phi *= sp[1] / sqrt(1.0 - dot * dot);
                        for (int i = 0; i < dim; ++i) {
                            double ff = phi * (dot * tg[tj][i] - tg[ti][i]);
                            double fb = phi * (tg[tj][i] - dot * tg[ti][i]);
                            F[b1 * dim + i] += ff + fb;
                            F[(b1 - 1) * dim + i] -= ff;
                            F[(b1 + 1) * dim + i] -= fb;
                        }
                    }
                    if (orient) {
                        double c = pow(2.0 * (1.0 + dot), -0.5);
                        for (int i = 0; i < dim; ++i)
                            tangent[b1 * dim + i] = c * (tg[ti][i] + tg[tj][i]);
                    }
                } else if (orient) {
                    for (int i = 0; i < dim; ++i) tangent[b1 * dim + i] = tg[ti][i];
                }
            }
        }
        tj = ti;
        ti = 1 - ti;
        new_chain = end;
    }
}

/* The pair term for one ordered pair, forces.py:100-133. Returns 1 and the
 * force on b1 in f[] when the pair is inside the cutoff. */
static int pair_term(int dim, const double* x1, const double* x2, double L, double half,
                     const double p[4], double rc2, double* f)
{
    double seps[3];
    double ds = 0.0;
    for (int i = 0; i < dim; ++i) {
        double sep = x2[i] - x1[i];
        if (sep < -half)
            sep += L;
        else if (sep >= half)
            sep -= L;
        seps[i] = sep;
        ds += sep * sep;
    }
    if (!(0.0 < ds && ds <= rc2)) return 0;
    if (p[2] != 0.0) {
        double mag = sqrt(ds);
        double nm = mag + p[2];
        if (nm < 0.0) nm = 0.0;
        double rescale = nm / mag;
        for (int i = 0; i < dim; ++i) seps[i] *= rescale;
        ds = nm * nm;
    }
    double c = orc_wca_coeff(ds, p);
    for (int i = 0; i < dim; ++i) f[i] = seps[i] * c;
    return 1;
}

static int excluded(const int64_t* chain_id, int b1, int b2, int n_nearby)
{
    /* updates.py:645-650 */
    if (chain_id[b1] != chain_id[b2]) return 0;
    if (n_nearby < 0) return 1;
    return abs(b1 - b2) <= n_nearby;
}

void orc_pair_forces_brute(int dim, int n, const double* pos, const int64_t* chain_id,
                           double L, const double p[4], int n_nearby, double* F)
{
    double half = L / 2.0, rc2 = p[3] * p[3];
    for (int b1 = 0; b1 < n; ++b1) {
        for (int b2 = 0; b2 < n; ++b2) {
            if (excluded(chain_id, b1, b2, n_nearby)) continue;
            double f[3];
            if (pair_term(dim, pos + b1 * dim, pos + b2 * dim, L, half, p, rc2, f))
                for (int i = 0; i < dim; ++i) F[b1 * dim + i] += f[i];
        }
    }
}

void orc_sliding(int dim, int n, const double* pos, const int64_t* chain_id, double L,
                 const double* t, int n_nearby, double maxdist, double v0, const int32_t* chunk,
                 double* F, int32_t* wa)
{
    double half = L / 2.0, md2 = maxdist * maxdist;
    for (int b1 = 0; b1 < n; ++b1)
        for (int b2 = 0; b2 < n; ++b2) {
            if (excluded(chain_id, b1, b2, n_nearby)) continue;
            if ((chunk ? chunk[b1] == chunk[b2] : 1) && b2 >= b1) continue;
            double ds = 0.0, dot = 0.0;
            for (int i = 0; i < dim; ++i) {
                double sep = pos[b2 * dim + i] - pos[b1 * dim + i];
                if (sep < -half)
                    sep += L;
                else if (sep >= half)
                    sep -= L;
                ds += sep * sep;
                dot += t[b1 * dim + i] * t[b2 * dim + i];
            }
            if (ds < md2 && dot < 0) {
                for (int i = 0; i < dim; ++i) F[b1 * dim + i] += v0 * t[b1 * dim + i];
                if (wa) wa[b1] += 1;
            }
        }
}

void orc_self_propel(int dim, int n, const int64_t* chain_id, double v0, const double* vtan,
                     const double* t, double* F)
{
    /* forces.py:206-223, angle looked up per chain. */
    for (int b = 0; b < n; ++b) {
        double ang = vtan ? vtan[chain_id[b]] : 0.0;
        if (ang == 0.0 || dim != 2) {
            for (int i = 0; i < dim; ++i) F[b * dim + i] += v0 * t[b * dim + i];
        } else {
            double sn = sin(ang), cs = cos(ang);
            F[b * dim + 0] += v0 * (cs * t[b * dim + 0] - sn * t[b * dim + 1]);
            F[b * dim + 1] += v0 * (sn * t[b * dim + 0] + cs * t[b * dim + 1]);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Verlet list on a cell grid (updates.py:502-669), rebuilt on a cadence,
 * with the reference's thread chunking (updates.py:323-344) and Newton-3
 * inside a chunk (forces.py:95-99, 134-143). Used by the CPU baseline. */

typedef struct verlet {
    int n, dim, n_threads;
    int* start;     /* CSR row starts, n+1 */
    int* list;      /* neighbour indices */
    size_t cap;
    int* chunk;     /* n_threads+1 chain-aligned chunk bounds */
    int n_grid, window;
    double radius2;
    int* cell_head; /* cell -> first bead (linked list) */
    int* cell_next;
} verlet;

static void chunk_chain(int T, int n, const int64_t* chain_id, int* bounds)
{
    /* updates.py:323-344: cut after roughly n/T beads, moved forward to the
     * next chain boundary. */
    int chunklen = (n + T - 1) / T;
    bounds[0] = 0;
    int b = 0;
    for (int t = 1; t < T; ++t) {
        int test = (t * chunklen - 1 > b) ? t * chunklen - 1 : b;
        if (test > n - 1) test = n - 1;
        int64_t c = chain_id[test];
        while (test < n && chain_id[test] == c) ++test;
        b = test;
        bounds[t] = b;
    }
    bounds[T] = n;
}

static void verlet_build(verlet* v, int dim, int n, const double* pos, const int64_t* chain_id,
                         double L, int n_nearby)
{
    int ng = v->n_grid, w = v->window;
    int ncell = ng * ng * (dim == 3 ? ng : 1);
    double h = L / ng, half = L / 2.0;
    for (int c = 0; c < ncell; ++c) v->cell_head[c] = -1;
    /* insert in descending order so each cell lists beads ascending, as
     * update_grid_list appends them (updates.py:572-582). */
    for (int b = n - 1; b >= 0; --b) {
        int g[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) {
            int gi = (int)floor(pos[b * dim + i] / h);
            if (gi >= ng) gi = ng - 1;
            if (gi < 0) gi = 0;
            g[i] = gi;
        }
        int c = (g[2] * ng + g[1]) * ng + g[0];
        v->cell_next[b] = v->cell_head[c];
        v->cell_head[c] = b;
    }
    int dgs = -(w < ng / 2 ? w : ng / 2);
    int dge = (1 + w < ng + dgs) ? 1 + w : ng + dgs;
    size_t cnt = 0;
    v->start[0] = 0;
    for (int b1 = 0; b1 < n; ++b1) {
        int g[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) {
            int gi = (int)floor(pos[b1 * dim + i] / h);
            if (gi >= ng) gi = ng - 1;
            if (gi < 0) gi = 0;
            g[i] = gi;
        }
        int zs = (dim == 3) ? dgs : 0, ze = (dim == 3) ? dge : 1;
        for (int dx = dgs; dx < dge; ++dx)
            for (int dy = dgs; dy < dge; ++dy)
                for (int dz = zs; dz < ze; ++dz) {
                    int gx = ((g[0] + dx) % ng + ng) % ng;
                    int gy = ((g[1] + dy) % ng + ng) % ng;
                    int gz = (dim == 3) ? ((g[2] + dz) % ng + ng) % ng : 0;
                    for (int b2 = v->cell_head[(gz * ng + gy) * ng + gx]; b2 >= 0;
                         b2 = v->cell_next[b2]) {
                        if (excluded(chain_id, b1, b2, n_nearby)) continue;
                        double d2 = 0.0;
                        for (int i = 0; i < dim; ++i) {
                            double s = pos[b2 * dim + i] - pos[b1 * dim + i];
                            if (s < -half)
                                s += L;
                            else if (s >= half)
                                s -= L;
                            d2 += s * s;
                        }
                        if (d2 <= v->radius2) {
                            if (cnt == v->cap) {
                                v->cap *= 2;
                                v->list = (int*)realloc(v->list, v->cap * sizeof(int));
                            }
                            v->list[cnt++] = b2;
                        }
                    }
                }
        v->start[b1 + 1] = (int)cnt;
    }
}

static void verlet_forces(const verlet* v, int dim, const double* pos, double L,
                          const double p[4], double* F)
{
    double half = L / 2.0, rc2 = p[3] * p[3];
#pragma omp parallel for schedule(static, 1) num_threads(v->n_threads)
    for (int t = 0; t < v->n_threads; ++t) {
        int s0 = v->chunk[t], s1 = v->chunk[t + 1];
        for (int b1 = s0; b1 < s1; ++b1) {
            for (int k = v->start[b1]; k < v->start[b1 + 1]; ++k) {
                int b2 = v->list[k];
                if (b1 <= b2 && b2 < s1) continue;
                int same = (b1 > b2 && b2 >= s0);
                double f[3];
                if (!pair_term(dim, pos + b1 * dim, pos + b2 * dim, L, half, p, rc2, f))
                    continue;
                for (int i = 0; i < dim; ++i) {
                    F[b1 * dim + i] += f[i];
                    if (same) F[b2 * dim + i] -= f[i];
                }
            }
        }
    }
}

/* Active sliding over the Verlet list, where the reference computes it
 * (inside central_pair_force's neighbour loop, forces.py:79-83, 114-120):
 * the rules of orc_sliding, whose count it equals while the sliding range
 * lies within the list radius. Each thread writes its own beads only. */
static void verlet_sliding(const verlet* v, int dim, const double* pos, double L, const double* t,
                           double maxdist, double v0, const int32_t* chunk, double* F,
                           int32_t* wa)
{
    double half = L / 2.0, md2 = maxdist * maxdist;
#pragma omp parallel for schedule(static, 1) num_threads(v->n_threads)
    for (int th = 0; th < v->n_threads; ++th)
        for (int b1 = v->chunk[th]; b1 < v->chunk[th + 1]; ++b1)
            for (int k = v->start[b1]; k < v->start[b1 + 1]; ++k) {
                int b2 = v->list[k];
                if ((chunk ? chunk[b1] == chunk[b2] : 1) && b2 >= b1) continue;
                double ds = 0.0, dot = 0.0;
                for (int i = 0; i < dim; ++i) {
                    double sep = pos[b2 * dim + i] - pos[b1 * dim + i];
                    if (sep < -half)
                        sep += L;
                    else if (sep >= half)
                        sep -= L;
                    ds += sep * sep;
                    dot += t[b1 * dim + i] * t[b2 * dim + i];
                }
                if (ds < md2 && dot < 0) {
                    for (int i = 0; i < dim; ++i) F[b1 * dim + i] += v0 * t[b1 * dim + i];
                    if (wa) wa[b1] += 1;
                }
            }
}

void orc_walls(int dim, int n, const double* pos, double L, const double wca[4], int walls,
               double zone, double* F)
{
    double half = L / 2.0;
    for (int b = 0; b < n; ++b)
        for (int i = 0; i < dim; ++i) {
            if (!((walls >> i) & 1)) continue;
            double sep = pos[b * dim + i];
            if (sep >= half)
                sep -= (L - zone);
            else
                sep -= zone;
            double ds;
            if (wca[2] != 0) {
                /* forces.py:425-429: |sep| - shift, clipped at 0, sign kept */
                double m = fabs(sep) - wca[2];
                if (m < 0) m = 0;
                sep = (sep > 0) ? m : (sep < 0 ? -m : 0.0);
                ds = m * m;
            } else {
                ds = sep * sep;
            }
            /* forces.py:432: the squared distance against wca[3] itself */
            if (0 < ds && ds <= wca[3]) F[b * dim + i] += -sep * orc_wca_coeff(ds, wca);
        }
}

void orc_force_eval(const orc_sys* s, const double* pos, const int64_t* chain_id,
                    const double* rest_lengths, const double* vtan, double* tangent,
                    double* F, int pair_mode, void* vv, const double* wall_pos)
{
    int dim = s->dim, n = s->n;
    if (s->has_springs) {
        if (pair_mode == 1 && vv) {
            const verlet* v = (const verlet*)vv;
#pragma omp parallel for schedule(static, 1) num_threads(v->n_threads)
            for (int t = 0; t < v->n_threads; ++t) {
                int s0 = v->chunk[t], s1 = v->chunk[t + 1];
                if (s1 > s0)
                    orc_chain_forces(dim, s1 - s0, pos + s0 * dim, chain_id + s0, s->L,
                                     s->spring, rest_lengths, tangent + s0 * dim,
                                     F + s0 * dim);
            }
        } else {
            orc_chain_forces(dim, n, pos, chain_id, s->L, s->spring, rest_lengths, tangent, F);
        }
    }
    if (s->sliding_maxdist > 0) {
        if (pair_mode == 1 && vv)
            verlet_sliding((const verlet*)vv, dim, pos, s->L, tangent, s->sliding_maxdist,
                           s->sliding_v0, s->chunk, F, s->where_active);
        else
            orc_sliding(dim, n, pos, chain_id, s->L, tangent, s->n_nearby, s->sliding_maxdist,
                        s->sliding_v0, s->chunk, F, s->where_active);
    }
    if (s->has_wca) {
        if (pair_mode == 1 && vv)
            verlet_forces((const verlet*)vv, dim, pos, s->L, s->wca, F);
        else
            orc_pair_forces_brute(dim, n, pos, chain_id, s->L, s->wca, s->n_nearby, F);
    }
    if (s->has_propel) orc_self_propel(dim, n, chain_id, s->v0, vtan, tangent, F);
    if (s->walls) {
        double ww[4] = {s->wca[0], s->wca[1], s->wca[2],
                        s->wall_cutoff > 0 ? s->wall_cutoff : s->wca[3]};
        orc_walls(dim, n, wall_pos ? wall_pos : pos, s->L, ww, s->walls, s->wall_zone, F);
    }
}

/* ---------------------------------------------------------------------- */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11. Multipliers and Weyl key increments are the
 * published constants. */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform in (0,1) with 2^-32 resolution, never 0 or 1 */
static double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

static void box_muller(uint32_t a, uint32_t b, double* n0, double* n1)
{
    double r = sqrt(-2.0 * log(u01(a)));
    double th = 6.283185307179586 * u01(b);
    *n0 = r * cos(th);
    *n1 = r * sin(th);
}

void orc_bead_noise(int dim, uint64_t seed, uint32_t orig, int64_t step, double* xi)
{
    /* Counter = (bead, step lo, step hi, block); key = seed. Block 0 gives
     * four normals (2D: translational pair then rotational pair). 3D uses
     * block 0 for x,y,z + rot x and block 1 for rot y,z. */
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {orig, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), 0u};
    uint32_t o[4];
    double g[8];
    orc_philox4x32_10(ctr, key, o);
    box_muller(o[0], o[1], &g[0], &g[1]);
    box_muller(o[2], o[3], &g[2], &g[3]);
    if (dim == 3) {
        ctr[3] = 1u;
        orc_philox4x32_10(ctr, key, o);
        box_muller(o[0], o[1], &g[4], &g[5]);
        box_muller(o[2], o[3], &g[6], &g[7]);
    }
    for (int i = 0; i < 2 * dim; ++i) xi[i] = g[i];
}

/* ---------------------------------------------------------------------- */

void orc_rod_project(int dim, int n, double* rtn, const double* prev, const double* F,
                     const double* xi, int nd, const int64_t* chain_id, const double* rest_lengths,
                     double factor, double L, double dt, double D, double radius)
{
    /* updates.py:347-499, chain by chain (chains are runs of chain_id). */
    double half = 0.5 * L;
    double dt_over_r = dt / radius;
    double sq = sqrt(2.0 * D * dt);
    int b0 = 0;
    while (b0 < n) {
        int64_t ci = chain_id[b0];
        int cl = 0;
        while (b0 + cl < n && chain_id[b0 + cl] == ci) ++cl;
        double a = rest_lengths[ci];
        double half_clm1 = (cl - 1) / 2.0;
        double* upd = (double*)malloc(sizeof(double) * (size_t)cl * dim);
        double* dist = (double*)malloc(sizeof(double) * (size_t)cl);
        double moi = 0.0;
        for (int k = 0; k < cl; ++k) {
            dist[k] = a * (k - half_clm1);
            for (int i = 0; i < dim; ++i) moi += dist[k] * dist[k]; /* summed once per axis */
        }
        for (int k = 0; k < cl; ++k)
            for (int i = 0; i < dim; ++i) {
                upd[k * dim + i] = dt_over_r * F[(b0 + k) * dim + i];
                if (sq != 0.0) upd[k * dim + i] += sq * xi[(size_t)(b0 + k) * nd + i];
            }
        double mean[3] = {0, 0, 0}, com[3] = {0, 0, 0}, rh[3] = {0, 0, 0}, rmag = 0.0;
        for (int k = 0; k < cl; ++k)
            for (int i = 0; i < dim; ++i) mean[i] += upd[k * dim + i];
        for (int i = 0; i < dim; ++i) mean[i] /= cl;
        for (int k = 0; k < cl; ++k)
            for (int i = 0; i < dim; ++i) {
                double sep = prev[(b0 + k) * dim + i] - prev[b0 * dim + i];
                if (sep < -half)
                    sep += L;
                else if (sep > half)
                    sep -= L;
                com[i] += sep;
                if (k == cl - 1) {
                    rh[i] = sep;
                    rmag += sep * sep;
                }
            }
        for (int i = 0; i < dim; ++i) com[i] = orc_pymod(prev[b0 * dim + i] + com[i] / cl, L);
        if (rmag > 0) {
            rmag = sqrt(rmag);
            for (int i = 0; i < dim; ++i) rh[i] /= rmag;
        }
        double* ool = (double*)malloc(sizeof(double) * (size_t)cl * dim);
        for (int k = 0; k < cl; ++k)
            for (int i = 0; i < dim; ++i) {
                double sep = prev[(b0 + k) * dim + i] - (com[i] + rh[i] * dist[k]);
                if (sep < -half)
                    sep += L;
                else if (sep >= half)
                    sep -= L;
                ool[k * dim + i] = sep;
            }
        if (dim == 2) {
            double tq[2] = {0, 0};
            for (int k = 0; k < cl; ++k)
                for (int i = 0; i < 2; ++i) tq[i] += (2 * i - 1) * rh[1 - i] * upd[k * 2 + i] * dist[k];
            if (moi != 0)
                for (int i = 0; i < 2; ++i) tq[i] /= moi;
            for (int k = 0; k < cl; ++k)
                for (int i = 0; i < 2; ++i) {
                    double u = upd[k * 2 + i] * (1.0 - factor);
                    u += factor * (-ool[k * 2 + i] + mean[i] + (tq[i] * dist[k] * (2 * i - 1) * rh[1 - i]));
                    rtn[(b0 + k) * 2 + i] = orc_pymod(prev[(b0 + k) * 2 + i] + u, L);
                }
        } else {
            /* torque = |r| sum_k r_hat x u_k; update += (|r|/I) torque x r_hat */
            double tq[3] = {0, 0, 0};
            for (int k = 0; k < cl; ++k) {
                const double* u = upd + k * 3;
                tq[0] += rh[1] * u[2] - rh[2] * u[1];
                tq[1] += rh[2] * u[0] - rh[0] * u[2];
                tq[2] += rh[0] * u[1] - rh[1] * u[0];
            }
            for (int i = 0; i < 3; ++i) tq[i] *= rmag;
            double c[3] = {tq[1] * rh[2] - tq[2] * rh[1], tq[2] * rh[0] - tq[0] * rh[2],
                           tq[0] * rh[1] - tq[1] * rh[0]};
            double w = rmag / moi;
            for (int k = 0; k < cl; ++k)
                for (int i = 0; i < 3; ++i) {
                    double u = upd[k * 3 + i] * (1.0 - factor);
                    u += factor * (-ool[k * 3 + i] + mean[i] + w * c[i]);
                    rtn[(b0 + k) * 3 + i] = orc_pymod(prev[(b0 + k) * 3 + i] + u, L);
                }
        }
        free(upd);
        free(dist);
        free(ool);
        b0 += cl;
    }
}

/* The Verlet list of one population with the reference's grid constants
 * (config.grid_list_setup, config.py:147-193, and the in-loop window of
 * SRK.py:99-103), chunked for n_threads threads, built at pos. */
static void verlet_setup(verlet* vv, const orc_sys* s, const double* pos, const int64_t* chain_id,
                         const double* rest_lengths, int n_threads, double grid_factor)
{
    int n = s->n, dim = s->dim;
    double L = s->L;
    double maxrest = 0.0;
    if (s->has_springs) {
        int64_t nc = chain_id[n - 1] + 1;
        for (int64_t c = 0; c < nc; ++c)
            if (rest_lengths[c] > maxrest) maxrest = rest_lengths[c];
    }
    double r_int = 2.0 * maxrest > s->wca[3] ? 2.0 * maxrest : s->wca[3];
    double ngd = L / (grid_factor * r_int);
    /* at most about four cells per bead: beads grown from points have a
     * vanishing range (larger cells only lengthen the lists) */
    double ng_cap = floor(pow(4.0 * (n > 0 ? n : 1), 1.0 / dim));
    if (!(ngd <= ng_cap)) ngd = ng_cap;
    int ng = (int)ngd;
    if (ng < 1) ng = 1;
    double h = L / ng;
    int w = (int)ceil(grid_factor * s->wca[3] / h);
    if (w < 1) w = 1;
    vv->n = n;
    vv->dim = dim;
    vv->n_threads = n_threads;
    vv->n_grid = ng;
    vv->window = w;
    vv->radius2 = (w * h) * (w * h);
    vv->start = (int*)malloc(sizeof(int) * (n + 1));
    vv->cap = (size_t)n * 64 + 64;
    vv->list = (int*)malloc(sizeof(int) * vv->cap);
    vv->chunk = (int*)malloc(sizeof(int) * (n_threads + 1));
    int ncell = ng * ng * (dim == 3 ? ng : 1);
    vv->cell_head = (int*)malloc(sizeof(int) * ncell);
    vv->cell_next = (int*)malloc(sizeof(int) * n);
    chunk_chain(n_threads, n, chain_id, vv->chunk);
    verlet_build(vv, dim, n, pos, chain_id, L, s->n_nearby);
}

static void verlet_free(verlet* vv)
{
    free(vv->start);
    free(vv->list);
    free(vv->chunk);
    free(vv->cell_head);
    free(vv->cell_next);
    memset(vv, 0, sizeof *vv);
}

int64_t orc_srk2_run(const orc_sys* s, double* pos, double* pos1, double* tangent, double* F,
                     const int64_t* chain_id, const double* rest_lengths, const double* vtan,
                     const double tv[8], double status_out[4], const double* xi_in,
                     uint64_t seed, int64_t step0, int pair_mode, int n_threads,
                     double grid_factor)
{
    int dim = s->dim, n = s->n;
    double dt = tv[0], dtmax = tv[1], mdps = tv[2], cheat = tv[3], mdu = tv[4];
    double until = tv[5], n_upd = tv[6], max_steps = tv[7];
    double L = s->L, D = s->D, Dr = s->Dr, r = s->radius;
    int orient = s->has_springs && (s->spring[4] != 0.0);
    int nd = 2 * dim;
    if (n_threads < 1) n_threads = 1;
    /* SRK.py:364: projection when the factor is set and the longest chain
     * has more than two beads */
    int longest = 0;
    for (int b = 0, run = 0; b < n; ++b) {
        run = (b > 0 && chain_id[b] == chain_id[b - 1]) ? run + 1 : 1;
        if (run > longest) longest = run;
    }
    int rod = s->rod != 0.0 && longest > 2;
    double* xi_buf = (double*)malloc(sizeof(double) * (size_t)n * nd);
    double* tmpT = NULL;
    (void)tmpT;

    verlet vv;
    memset(&vv, 0, sizeof vv);
    if (pair_mode == 1 && s->has_wca)
        verlet_setup(&vv, s, pos, chain_id, rest_lengths, n_threads, grid_factor);
    void* vp = (pair_mode == 1 && s->has_wca) ? (void*)&vv : NULL;

    const int order = s->order == 4 ? 4 : 2;
    double* Faux = order == 4 ? (double*)malloc(sizeof(double) * (size_t)n * dim) : NULL;

    int64_t steps = 0;
    double elapsed = 0.0, m = 0.0, dt_for_update = dt;
    while ((steps < max_steps || max_steps < 0) && (elapsed < mdu || mdu < 0)) {
        if (s->where_active)  /* SRK.py:132 */
            for (int b = 0; b < n; ++b) s->where_active[b] = -1;
        /* noise for this step (SRK.py:134-142): drawn when D>0 / Dr>0 */
        const double* xi;
        if (xi_in) {
            xi = xi_in + (size_t)steps * n * nd;
        } else {
#pragma omp parallel for num_threads(n_threads)
            for (int b = 0; b < n; ++b)
                orc_bead_noise(dim, seed, (uint32_t)b, step0 + steps, xi_buf + (size_t)b * nd);
            xi = xi_buf;
        }
        memset(F, 0, sizeof(double) * (size_t)n * dim);
        for (int stage = 0; order == 4 && stage < 4; ++stage) {
            /* SRK4 (SRK.py:186-215): forces of each stage into a wiped
             * auxiliary array, at pos in the first stage and at the last
             * virtual positions after it; total = aux/6 + aux/3 + aux/3 +
             * aux/6 with the multipliers as doubles (SRK.py:297-318); virtual
             * updates into pos1 from pos with dt * 1/2, 1/2, 1 and the plain
             * sqrt(2 D dt) noise; the real update with the total and
             * dt_used, force prefactor 1 throughout (SRK.py:145, 355-358). */
            const double* at = stage == 0 ? pos : pos1;
            double* out = stage == 3 ? pos : pos1;
            memset(Faux, 0, sizeof(double) * (size_t)n * dim);
            orc_force_eval(s, at, chain_id, rest_lengths, vtan, tangent, Faux, pair_mode, vp, pos);
            const double mult = (stage == 0 || stage == 3) ? 1. / 6. : 1. / 3.;
            for (int k = 0; k < n * dim; ++k)
                F[k] = stage == 0 ? Faux[k] * mult : F[k] + Faux[k] * mult;
            const double* src;
            double resc;
            if (stage == 3) {
                double mx = 0.0;
                for (int b = 0; b < n; ++b) {
                    double q = 0.0;
                    for (int i = 0; i < dim; ++i) q += F[b * dim + i] * F[b * dim + i];
                    if (q > mx) mx = q;
                }
                m = sqrt(2.0 * D);
                if (n > 0) m += 1.0 * sqrt(mx) / r;
                steps += 1;
                if (m == 0.0)
                    dt = dtmax;
                else
                    dt = fmin(dtmax, mdps / m);
                double dt_used = cheat * dt;
                if (mdu > 0) dt_used = fmin(dt_used, mdu - elapsed);
                elapsed += dt_used;
                dt_for_update = dt_used;
                resc = 1.0;
                src = F;
            } else {
                dt_for_update = dt;
                resc = stage <= 1 ? 0.5 : 1.0;
                src = Faux;
            }
            double dt_eff = dt_for_update * resc * 1.0;
            double s2 = sqrt(2.0 * dt_for_update * D);
            if (rod) {
                orc_rod_project(dim, n, out, at, src, xi, nd, chain_id, rest_lengths, s->rod, L,
                                dt_eff, D, r);
                continue;
            }
            dt_eff /= r;
            for (int b = 0; b < n; ++b)
                for (int i = 0; i < dim; ++i) {
                    double tmp = pos[b * dim + i] + dt_eff * src[b * dim + i];
                    if (s2 != 0.0) tmp += s2 * xi[(size_t)b * nd + i];
                    out[b * dim + i] = orc_pymod(tmp, L);
                }
        }
        for (int stage = 0; order == 2 && stage < 2; ++stage) {
            double pref = stage == 0 ? 1.0 : 0.5;
            const double* at = stage == 0 ? pos : pos1;
            double* out = stage == 0 ? pos1 : pos;
            orc_force_eval(s, at, chain_id, rest_lengths, vtan, tangent, F, pair_mode, vp, pos);
            if (stage == 1) {
                /* SRK.py:324-353 */
                double mx = 0.0;
                for (int b = 0; b < n; ++b) {
                    double q = 0.0;
                    for (int i = 0; i < dim; ++i) q += F[b * dim + i] * F[b * dim + i];
                    if (q > mx) mx = q;
                }
                m = sqrt(2.0 * D);
                if (n > 0) m += pref * sqrt(mx) / r;
                steps += 1;
                if (m == 0.0)
                    dt = dtmax;
                else
                    dt = fmin(dtmax, mdps / m);
                double dt_used = cheat * dt;
                if (mdu > 0) dt_used = fmin(dt_used, mdu - elapsed);
                elapsed += dt_used;
                dt_for_update = dt_used;
            } else {
                dt_for_update = dt;
            }
            /* SRK.py:355-388 */
            double dt_eff = dt_for_update * pref;
            double s2 = sqrt(2.0 * dt_for_update * D);
            if (rod) {
                /* SRK.py:364-371: the projected update starts from the
                 * position the forces were evaluated at (pos1 in stage 1)
                 * and uses sqrt(2 D dt_eff) as noise amplitude */
                orc_rod_project(dim, n, out, at, F, xi, nd, chain_id, rest_lengths, s->rod, L,
                                dt_eff, D, r);
                continue;
            }
            dt_eff /= r;
#pragma omp parallel for num_threads(n_threads)
            for (int b = 0; b < n; ++b)
                for (int i = 0; i < dim; ++i) {
                    double tmp = pos[b * dim + i] + dt_eff * F[b * dim + i];
                    if (s2 != 0.0) tmp += s2 * xi[(size_t)b * nd + i];
                    out[b * dim + i] = orc_pymod(tmp, L);
                }
        }
        /* rotational diffusion, SRK.py:398-423 */
        if (Dr > 0 && !orient) {
            double sr = sqrt(2.0 * dt_for_update * Dr) * 3.141592653589793;
            for (int b = 0; b < n; ++b) {
                double nr[3], proj = 0.0;
                for (int i = 0; i < dim; ++i) nr[i] = xi[(size_t)b * nd + dim + i];
                for (int i = 0; i < dim; ++i) proj += nr[i] * tangent[b * dim + i];
                for (int i = 0; i < dim; ++i) nr[i] -= proj * tangent[b * dim + i];
                double mag = 0.0;
                for (int i = 0; i < dim; ++i) {
                    double t = tangent[b * dim + i] + sr * nr[i];
                    mag += t * t;
                    tangent[b * dim + i] = t;
                }
                mag = sqrt(mag);
                for (int i = 0; i < dim; ++i) tangent[b * dim + i] /= mag;
            }
        }
        /* list rebuild cadence, SRK.py:428-484 */
        if (vp && (double)steps >= until) {
            verlet_build(&vv, dim, n, pos, chain_id, L, s->n_nearby);
            until += n_upd;
        }
    }
    status_out[0] = dt;
    status_out[1] = elapsed;
    status_out[2] = (double)steps;
    status_out[3] = m;
    free(xi_buf);
    free(Faux);
    if (vp) {
        verlet_free(&vv);
    }
    return steps;
}

/* ---------------------------------------------------------------------- */
/* Several populations (updates.py:132-267, SRK.py:44, 287-293). */

void orc_cross_pair_forces(int dim, int nA, const double* posA, int nB, const double* posB,
                           double L, const double p[4], double* FA)
{
    double half = L / 2.0, rc2 = p[3] * p[3];
    for (int a = 0; a < nA; ++a)
        for (int b = 0; b < nB; ++b) {
            double f[3];
            if (pair_term(dim, posA + a * dim, posB + b * dim, L, half, p, rc2, f))
                for (int i = 0; i < dim; ++i) FA[a * dim + i] += f[i];
        }
}

double orc_motor_uniform(uint64_t seed, uint32_t motor, uint32_t bead, int64_t step, int stage,
                         int kind)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {motor, bead, (uint32_t)(uint64_t)step,
                       ((uint32_t)((uint64_t)step >> 32) << 8) | 0x80u | ((uint32_t)stage << 1) |
                           (uint32_t)kind};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
}

void orc_motor_push(int dim, int nb, const double* posb, const double* tb, int nm,
                    const double* posm, double L, const orc_pop* M, int64_t step, int stage,
                    double* F, int32_t* wab)
{
    double half = L / 2.0, reach2 = M->motor_reach * M->motor_reach;
    for (int b = 0; b < nb; ++b)
        for (int m = 0; m < nm; ++m) {
            double d2 = 0.0;
            for (int i = 0; i < dim; ++i) {
                double sep = posm[m * dim + i] - posb[b * dim + i];
                if (sep < -half)
                    sep += L;
                else if (sep >= half)
                    sep -= L;
                d2 += sep * sep;
            }
            if (d2 > reach2) continue;
            int32_t* mb = &M->motor_bead[m];
            if (*mb >= 0) {
                if (orc_motor_uniform(M->seed, (uint32_t)m, (uint32_t)b, step, stage, 0) > M->k_off)
                    continue;
                *mb = -1;
                if (wab) wab[b] -= 1;
            }
            if (*mb < 0 && orc_motor_uniform(M->seed, (uint32_t)m, (uint32_t)b, step, stage, 1) < M->k_on &&
                d2 <= M->motor_max_dist_sq) {
                if (wab) wab[b] += 1;
                *mb = b;
                for (int i = 0; i < dim; ++i) F[b * dim + i] += M->motor_imparts * tb[b * dim + i];
            }
        }
}

/* The beads of one population on a cell grid of edge >= reach (about four
 * beads per cell when that is larger), each cell listing its beads in
 * ascending index: the cell-list scans of the CPU
 * baseline between populations (the reference scans nbor_list[other],
 * updates.py:586-734, built the same way). */
typedef struct cgrid {
    int ng, dim, cap_cells, cap_beads;
    double L, h;
    int* head;
    int* next;
} cgrid;

static void cgrid_build(cgrid* g, int dim, int n, const double* pos, double L, double reach)
{
    /* at least the reach, and about four beads per cell when that is more */
    double edge = n > 0 ? pow(4.0 * pow(L, dim) / n, 1.0 / dim) : L;
    if (reach > edge) edge = reach;
    int ng = edge > 0 ? (int)(L / edge) : 1;
    int lim = dim == 3 ? 256 : 4096;
    if (ng < 1) ng = 1;
    if (ng > lim) ng = lim;
    int ncell = ng * ng * (dim == 3 ? ng : 1);
    if (ncell > g->cap_cells) {
        g->head = (int*)realloc(g->head, sizeof(int) * ncell);
        g->cap_cells = ncell;
    }
    if (n > g->cap_beads) {
        g->next = (int*)realloc(g->next, sizeof(int) * n);
        g->cap_beads = n;
    }
    g->ng = ng;
    g->dim = dim;
    g->L = L;
    g->h = L / ng;
    for (int c = 0; c < ncell; ++c) g->head[c] = -1;
    for (int b = n - 1; b >= 0; --b) {
        int c = 0;
        for (int i = dim - 1; i >= 0; --i) {
            int gi = (int)floor(pos[b * dim + i] / g->h);
            if (gi >= ng) gi = ng - 1;
            if (gi < 0) gi = 0;
            c = c * ng + gi;
        }
        g->next[b] = g->head[c];
        g->head[c] = b;
    }
}

/* The cells within one cell of x's (every cell when the grid has fewer
 * than three per side), into cells[]; returns their number. */
static int cgrid_near(const cgrid* g, const double* x, int* cells)
{
    int ng = g->ng, dim = g->dim, k = 0;
    if (ng < 3) {
        int ncell = ng * ng * (dim == 3 ? ng : 1);
        for (int c = 0; c < ncell; ++c) cells[k++] = c;
        return k;
    }
    int gc[3] = {0, 0, 0};
    for (int i = 0; i < dim; ++i) {
        int gi = (int)floor(x[i] / g->h);
        if (gi >= ng) gi = ng - 1;
        if (gi < 0) gi = 0;
        gc[i] = gi;
    }
    for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int cx = (gc[0] + dx + ng) % ng, cy = (gc[1] + dy + ng) % ng;
                int cz = dim == 3 ? (gc[2] + dz + ng) % ng : 0;
                cells[k++] = (cz * ng + cy) * ng + cx;
            }
    return k;
}

/* orc_cross_pair_forces over the cells near each bead of A (grid of B with
 * edge >= the law's cutoff), beads of A split over n_threads threads. */
static void cross_pair_forces_cells(int dim, int nA, const double* posA, const cgrid* g,
                                    const double* posB, double L, const double p[4], double* FA,
                                    int n_threads)
{
    double half = L / 2.0, rc2 = p[3] * p[3];
#pragma omp parallel for schedule(static) num_threads(n_threads)
    for (int a = 0; a < nA; ++a) {
        int cells[27];
        int nc = g->ng < 3 ? 0 : cgrid_near(g, posA + a * dim, cells);
        int all = g->ng < 3;
        int ncell = g->ng * g->ng * (dim == 3 ? g->ng : 1);
        for (int k = 0; k < (all ? ncell : nc); ++k)
            for (int b = g->head[all ? k : cells[k]]; b >= 0; b = g->next[b]) {
                double f[3];
                if (pair_term(dim, posA + a * dim, posB + b * dim, L, half, p, rc2, f))
                    for (int i = 0; i < dim; ++i) FA[a * dim + i] += f[i];
            }
    }
}

/* orc_motor_push over the motors in the cells near each bead (grid of the
 * motors with edge >= their reach). Beads in ascending index as there; each
 * motor then meets its beads in the same order, and a bead's pushes are
 * equal terms, so the result is the same. */
static void motor_push_cells(int dim, int nb, const double* posb, const double* tb,
                             const cgrid* g, const double* posm, double L, const orc_pop* M,
                             int64_t step, int stage, double* F, int32_t* wab)
{
    double half = L / 2.0, reach2 = M->motor_reach * M->motor_reach;
    int ncell = g->ng * g->ng * (dim == 3 ? g->ng : 1);
    for (int b = 0; b < nb; ++b) {
        int cells[27];
        int all = g->ng < 3;
        int nc = all ? ncell : cgrid_near(g, posb + b * dim, cells);
        for (int k = 0; k < nc; ++k)
            for (int m = g->head[all ? k : cells[k]]; m >= 0; m = g->next[m]) {
                double d2 = 0.0;
                for (int i = 0; i < dim; ++i) {
                    double sep = posm[m * dim + i] - posb[b * dim + i];
                    if (sep < -half)
                        sep += L;
                    else if (sep >= half)
                        sep -= L;
                    d2 += sep * sep;
                }
                if (d2 > reach2) continue;
                int32_t* mb = &M->motor_bead[m];
                if (*mb >= 0) {
                    if (orc_motor_uniform(M->seed, (uint32_t)m, (uint32_t)b, step, stage, 0) > M->k_off)
                        continue;
                    *mb = -1;
                    if (wab) wab[b] -= 1;
                }
                if (*mb < 0 && orc_motor_uniform(M->seed, (uint32_t)m, (uint32_t)b, step, stage, 1) < M->k_on &&
                    d2 <= M->motor_max_dist_sq) {
                    if (wab) wab[b] += 1;
                    *mb = b;
                    for (int i = 0; i < dim; ++i) F[b * dim + i] += M->motor_imparts * tb[b * dim + i];
                }
            }
    }
}

int64_t orc_multi_srk2_run(int npop, orc_pop* pops, const double* cross, const double tv[8],
                           double status_out[4], int64_t step0, int pair_mode, int n_threads,
                           double grid_factor)
{
    double dt = tv[0], dtmax = tv[1], mdps = tv[2], cheat = tv[3], mdu = tv[4];
    double until = tv[5], n_upd = tv[6], max_steps = tv[7];
    if (n_threads < 1) n_threads = 1;
    /* pair_mode 1 (the CPU baseline): each moving population's own pairs
     * over its Verlet list, rebuilt on the reference's cadence; the other
     * populations through a cell grid of their beads */
    verlet vv[16];
    memset(vv, 0, sizeof vv);
    cgrid cg;
    memset(&cg, 0, sizeof cg);
    for (int q = 0; q < npop && q < 16; ++q)
        if (pair_mode == 1 && !pops[q].fixed && pops[q].sys.has_wca)
            verlet_setup(&vv[q], &pops[q].sys, pops[q].pos, pops[q].chain_id,
                         pops[q].rest_lengths, n_threads, grid_factor);
    int dim = pops[0].sys.dim;
    double L = pops[0].sys.L;
    int nd = 2 * dim;
    double** xi_buf = (double**)calloc((size_t)npop, sizeof(double*));
    for (int q = 0; q < npop; ++q)
        xi_buf[q] = (double*)malloc(sizeof(double) * (size_t)(pops[q].sys.n > 0 ? pops[q].sys.n : 1) * nd);
    int64_t steps = 0;
    double elapsed = 0.0, m = 0.0, dt_for_update = dt;
    while ((steps < max_steps || max_steps < 0) && (elapsed < mdu || mdu < 0)) {
        for (int q = 0; q < npop; ++q) {  /* SRK.py:132, by every moving population */
            orc_pop* P = &pops[q];
            if (P->fixed) continue;
            if (P->sys.where_active)
                for (int b = 0; b < P->sys.n; ++b) P->sys.where_active[b] = -1;
            if (P->motor && P->motor_bead)
                for (int b = 0; b < P->sys.n; ++b) P->motor_bead[b] = -1;
        }
        const double* xi[16];
        for (int q = 0; q < npop && q < 16; ++q) {
            const orc_pop* P = &pops[q];
            if (P->xi) {
                xi[q] = P->xi + (size_t)steps * P->sys.n * nd;
            } else {
                for (int b = 0; b < P->sys.n; ++b)
                    orc_bead_noise(dim, P->seed, (uint32_t)b, step0 + steps,
                                   xi_buf[q] + (size_t)b * nd);
                xi[q] = xi_buf[q];
            }
        }
        for (int stage = 0; stage < 2; ++stage) {
            /* forces of every moving population, then (stage 1) one dt */
            for (int q = 0; q < npop; ++q) {
                orc_pop* P = &pops[q];
                if (P->fixed) continue;
                int n = P->sys.n;
                const double* at = stage == 0 ? P->pos : P->pos1;
                if (stage == 0) memset(P->forces, 0, sizeof(double) * (size_t)n * dim);
                void* vq = (q < 16 && vv[q].start) ? (void*)&vv[q] : NULL;
                orc_force_eval(&P->sys, at, P->chain_id, P->rest_lengths, P->vtan, P->tangent,
                               P->forces, vq ? 1 : 0, vq, P->pos);
                for (int ii = 1; ii < npop; ++ii) {
                    int o = (q + ii) % npop;
                    const double* c = cross + ((size_t)q * npop + o) * 5;
                    if (c[0] == 0.0) continue;
                    const orc_pop* Q = &pops[o];
                    const double* qpos = (stage == 0 || Q->fixed) ? Q->pos : Q->pos1;
                    if (pair_mode == 1) {
                        cgrid_build(&cg, dim, Q->sys.n, qpos, L,
                                    c[0] == 2.0 ? Q->motor_reach : c[4]);
                        if (c[0] == 2.0)
                            motor_push_cells(dim, n, at, P->tangent, &cg, qpos, L, Q,
                                             step0 + steps, stage, P->forces, P->sys.where_active);
                        else
                            cross_pair_forces_cells(dim, n, at, &cg, qpos, L, c + 1, P->forces,
                                                    n_threads);
                    } else if (c[0] == 2.0) {
                        orc_motor_push(dim, n, at, P->tangent, Q->sys.n, qpos, L, Q,
                                       step0 + steps, stage, P->forces, P->sys.where_active);
                    } else {
                        orc_cross_pair_forces(dim, n, at, Q->sys.n, qpos, L, c + 1, P->forces);
                    }
                }
            }
            if (stage == 1) {
                m = 0.0;
                for (int q = 0; q < npop; ++q) {
                    const orc_pop* P = &pops[q];
                    if (P->fixed) continue;
                    double mx = 0.0;
                    for (int b = 0; b < P->sys.n; ++b) {
                        double s2 = 0.0;
                        for (int i = 0; i < dim; ++i)
                            s2 += P->forces[b * dim + i] * P->forces[b * dim + i];
                        if (s2 > mx) mx = s2;
                    }
                    double mp = sqrt(2.0 * P->sys.D);
                    if (P->sys.n > 0) mp += 0.5 * sqrt(mx) / P->sys.radius;
                    if (mp > m) m = mp;
                }
                steps += 1;
                dt = (m == 0.0) ? dtmax : fmin(dtmax, mdps / m);
                double dt_used = cheat * dt;
                if (mdu > 0) dt_used = fmin(dt_used, mdu - elapsed);
                elapsed += dt_used;
                dt_for_update = dt_used;
            } else {
                dt_for_update = dt;
            }
            double pref = stage == 0 ? 1.0 : 0.5;
            for (int q = 0; q < npop; ++q) {
                orc_pop* P = &pops[q];
                if (P->fixed) continue;
                double* out = stage == 0 ? P->pos1 : P->pos;
                double dt_eff = dt_for_update * pref / P->sys.radius;
                double s2 = sqrt(2.0 * dt_for_update * P->sys.D);
                for (int b = 0; b < P->sys.n; ++b)
                    for (int i = 0; i < dim; ++i) {
                        /* SRK.py:389-396: a pushing motor stays where it is
                         * (its pos1 keeps the last value written) */
                        if (P->motor && P->motor_still && P->motor_bead[b] >= 0) break;
                        double tmp = P->pos[b * dim + i] + dt_eff * P->forces[b * dim + i];
                        if (s2 != 0.0) tmp += s2 * xi[q][(size_t)b * nd + i];
                        out[b * dim + i] = orc_pymod(tmp, L);
                    }
            }
        }
        for (int q = 0; q < npop; ++q) {  /* SRK.py:398-423 per population */
            orc_pop* P = &pops[q];
            int orient = P->sys.has_springs && P->sys.spring[4] != 0.0;
            if (P->fixed || !(P->sys.Dr > 0) || orient) continue;
            double sr = sqrt(2.0 * dt_for_update * P->sys.Dr) * 3.141592653589793;
            for (int b = 0; b < P->sys.n; ++b) {
                double nr[3], proj = 0.0, mag = 0.0;
                double* t = P->tangent + b * dim;
                for (int i = 0; i < dim; ++i) nr[i] = xi[q][(size_t)b * nd + dim + i];
                for (int i = 0; i < dim; ++i) proj += nr[i] * t[i];
                for (int i = 0; i < dim; ++i) nr[i] -= proj * t[i];
                for (int i = 0; i < dim; ++i) {
                    t[i] = t[i] + sr * nr[i];
                    mag += t[i] * t[i];
                }
                mag = sqrt(mag);
                for (int i = 0; i < dim; ++i) t[i] /= mag;
            }
        }
        if (pair_mode == 1 && (double)steps >= until) {  /* the lists' cadence */
            for (int q = 0; q < npop && q < 16; ++q)
                if (vv[q].start)
                    verlet_build(&vv[q], dim, pops[q].sys.n, pops[q].pos, pops[q].chain_id, L,
                                 pops[q].sys.n_nearby);
            until += n_upd;
        }
    }
    status_out[0] = dt;
    status_out[1] = elapsed;
    status_out[2] = (double)steps;
    status_out[3] = m;
    for (int q = 0; q < npop; ++q) free(xi_buf[q]);
    free(xi_buf);
    for (int q = 0; q < npop && q < 16; ++q)
        if (vv[q].start) verlet_free(&vv[q]);
    free(cg.head);
    free(cg.next);
    return steps;
}
