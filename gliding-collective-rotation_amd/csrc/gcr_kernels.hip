// gcr_kernels.hip -- CDNA4 (gfx950) kernels of the Brownian-dynamics hot path.
//
// One step of SRK2 (SRK.py:165-396) is four launches over all replicas of a
// pool (the force stages are the LDS tile kernels of gcr_tile2d.hip and
// gcr_tile3d.hip):
//   stage 0     forces at x   (springs, pair, propulsion) -> x1, F0
//   stage 1     forces at x1, F = F0 + F1, per-tile max |F|^2
//   k_finalize  per-replica max force -> dt, dt_used (SRK.py:323-353)
//   k_update    x <- x + dt_used/(2r) F + sqrt(2 D dt_used) xi, rot. diffusion,
//               the next step's noise
// Every `rebuild_every` steps the beads are re-sorted by cell (counting sort):
//   k_bin, k_scan_{chunks,reps,final}, k_scatter, k_gather, k_links, lists
// Force kernels write one output per bead and use no atomics on forces, so a
// replica's result does not depend on scheduling or on other replicas.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gcr_device.h"

#define BS 256  // threads per bead block
#define MOTOR_LANES 4  // lanes that scan the window of one motor together
#define SLIDE_LANES 4  // lanes that scan the window of one bead for sliding partners
// Lanes that scan the other populations' cells for one bead. One: the cross
// forces are FP64 pair arithmetic for most candidates (unlike the sliding and
// motor scans, which reject most of theirs on the range), and four lanes per
// bead measured slower on 400k beads in two populations (47 against 45 us per
// stage).
#define GROUP_LANES 1
#define ROD_MAX_CPW 21  // chains per wavefront of k_rod at most (chains of 3 beads)

namespace gcr {

// ------------------------------------------------------------------ helpers

// CPython float_rem (the `tmp %= Lx` at SRK.py:387), for b > 0. A position
// that moved by less than one box lies in (-b, 2b): there fmod is the
// identity or one subtraction that is exact (Sterbenz), and the sign fix-up
// is the same addition CPython makes, so the fast cases give CPython's bits
// without the cost of a general FP64 fmod.
__device__ __forceinline__ double pymod(double a, double b)
{
    if (a >= 0.0 && a < b) return a == 0.0 ? 0.0 : a;  // +0 for a signed zero, as copysign(0, b)
    if (a >= b && a < 2.0 * b) return a - b;
    if (a < 0.0 && a > -b) return a + b;
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

// One-shot periodic correction into [-L/2, L/2) (forces.py:44-54, 103-107).
__device__ __forceinline__ double wrap1(double s, double L, double halfL)
{
    if (s < -halfL)
        s += L;
    else if (s >= halfL)
        s -= L;
    return s;
}

// Philox4x32-10 (Salmon et al., SC'11), published multipliers / Weyl constants.
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& n0, double& n1)
{
    double r = sqrt(-2.0 * log(u01(a)));
    double th = 6.283185307179586 * u01(b);
    double s, c;
    sincos(th, &s, &c);
    n0 = r * c;
    n1 = r * s;
}

// Noise of one bead for one step: counter (orig, step lo, step hi, block),
// key = replica seed. Same definition as the oracle's orc_bead_noise.
template <int DIM>
__device__ __forceinline__ void bead_noise(uint64_t seed, uint32_t orig, int64_t step, bool need_rot,
                                           double* xi)
{
    uint32_t c[4] = {orig, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double g0, g1, g2 = 0.0, g3 = 0.0;
    box_muller(c[0], c[1], g0, g1);
    if (DIM == 3 || need_rot) box_muller(c[2], c[3], g2, g3);
    xi[0] = g0;
    xi[1] = g1;
    if (DIM == 2) {
        xi[2] = g2;
        xi[3] = g3;
    } else {
        xi[2] = g2;
        xi[3] = g3;
        if (need_rot) {
            uint32_t d[4] = {orig, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), 1u};
            philox(d, (uint32_t)seed, (uint32_t)(seed >> 32));
            double g4, g5;
            box_muller(d[0], d[1], g4, g5);
            xi[4] = g4;
            xi[5] = g5;
        } else {
            xi[4] = 0.0;
            xi[5] = 0.0;
        }
    }
}

// Noise of bead `orig_global` for the replica's step with call-local index
// `steps` and global index `step_global`: drawn in-kernel (keyed by the
// original bead index, so it does not depend on the slot layout), or read
// from the injected array (zeros past its end: that step never runs). The
// draw runs in both cases and the injected values replace it, so that no
// branch holding loads precedes the draw's arithmetic (the waits on memory
// after such a join would wait for every load in flight).
template <int DIM>
__device__ __forceinline__ void draw_noise(const RepParams& P, int orig_global, int64_t steps,
                                           int64_t step_global, double* xi)
{
    const int orig_local = orig_global - P.bead_off;
    bead_noise<DIM>(P.seed, (uint32_t)orig_local, step_global, (P.flags & GCR_F_ROTDIFF) != 0, xi);
    if (P.flags & GCR_F_INJECT) {
        if (steps < P.xi_steps) {
            // a global-memory load (a flat one would count against both waits)
            typedef const __attribute__((address_space(1))) double* gptr;
            const gptr src = (gptr)(P.xi_inject + ((size_t)steps * P.n + orig_local) * (2 * DIM));
#pragma unroll
            for (int i = 0; i < 2 * DIM; ++i) xi[i] = src[i];
        } else {
#pragma unroll
            for (int i = 0; i < 2 * DIM; ++i) xi[i] = 0.0;
        }
    }
}

// Per-block running condition of a replica for local step j of a batch.
// Every field is read whatever the outcome and the tests are combined without
// branches, so that the reads issue together: at the start of a kernel the
// replica's state is not in the scalar cache, and a test per read would wait
// on memory once per cache line, in sequence.
__device__ __forceinline__ bool step_runs(const RepState* S, int j)
{
    const int32_t frozen = S->frozen, fixed = S->fixed;
    const int64_t steps = S->steps, base = S->batch_base, max_steps = S->max_steps;
    const double mdu = S->mdu, elapsed = S->elapsed;
    const bool idle = (frozen | fixed) != 0;
    const bool due = steps == base + j;
    const bool a = (max_steps < 0) | (steps < max_steps);
    const bool b = (mdu < 0) | (elapsed < mdu);
    return !idle & due & a & b;
}

// A cell re-sort placed before local step j runs only when that step runs
// (j < 0: unconditional, for set-up and diagnostics), so the arrays of the
// last executed step (x1, forces) keep the layout they were computed in. The
// mid-step re-sort on x1 runs only for a replica whose predictor of this
// step left the skin. All reads issue together, as in step_runs.
__device__ __forceinline__ bool rebuild_runs(const PoolArgs& a, int r, int j)
{
    const RepState* S = a.S + r;
    const int32_t frozen = S->frozen, fixed = S->fixed, pred_far = S->pred_far, sort_due = S->sort_due;
    const bool runs = step_runs(S, j);
    if (frozen) return false;
    if (fixed) return j < 0 && !a.on_x1;  // a fixed population is sorted once
    if (a.on_x1) return runs && ((pred_far >> a.sub) & 1);
    // the safe batch re-sorts at its start and then wherever an update left
    // the skin (its lists are otherwise still complete)
    if (a.safe && j > 0) return runs && sort_due;
    return j < 0 || runs;
}

// A virtual position outside the skin: outside the safe batch the batch is
// replayed; inside it the replica is re-sorted on those positions before the
// next force evaluation.
__device__ __forceinline__ void flag_predictor(const PoolArgs& a, RepState* S, int sub)
{
    atomicOr(a.safe ? &S->pred_far : &S->violation, a.safe ? 1 << sub : 1);
}

// SRK order handling (SRK.py:143-215, 297-321, 355-360).
// Force prefactor of the real update and of the max force: the mean of two
// evaluations in SRK2, 1 in SRK4.
__device__ __forceinline__ double force_prefactor(const RepParams& P)
{
    return (P.flags & GCR_F_SRK4) ? 1.0 : 0.5;
}

// dt_rescale_factor of a virtual update: SRK4 stages 0 and 1 take half the
// step, stage 2 the whole step; SRK2's predictor the whole step.
__device__ __forceinline__ double virtual_rescale(const RepParams& P, int sub)
{
    return ((P.flags & GCR_F_SRK4) && sub <= 1) ? 0.5 : 1.0;
}

// Whether a launch of the later force stages (at X1) for SRK stage `sub`
// has work for this replica: SRK4 evaluates at stages 1, 2 and 3, SRK2 at 3.
__device__ __forceinline__ bool later_stage_applies(uint32_t flags, int sub)
{
    return (flags & GCR_F_SRK4) ? sub >= 1 : sub == 3;
}

template <int DIM>
__device__ __forceinline__ void load_vec(const double* A, int s, double* v)
{
    if (DIM == 2) {
        double2 t = *reinterpret_cast<const double2*>(A + 2 * s);
        v[0] = t.x;
        v[1] = t.y;
    } else {
        v[0] = A[3 * s];
        v[1] = A[3 * s + 1];
        v[2] = A[3 * s + 2];
    }
}

template <int DIM>
__device__ __forceinline__ void store_vec(double* A, int s, const double* v)
{
    if (DIM == 2) {
        *reinterpret_cast<double2*>(A + 2 * s) = make_double2(v[0], v[1]);
    } else {
        A[3 * s] = v[0];
        A[3 * s + 1] = v[1];
        A[3 * s + 2] = v[2];
    }
}

// Noise of the bead in slot s for the given step: translation into XI[s],
// rotation into XIR[s] when the update uses it.
template <int DIM>
__device__ __forceinline__ void store_noise(const PoolArgs& a, const RepParams& P, int s, int64_t steps,
                                            int64_t step_global)
{
    double xi[2 * DIM];
    draw_noise<DIM>(P, a.META[2 * s], steps, step_global, xi);
    store_vec<DIM>(a.XI, s, xi);
    if (P.flags & GCR_F_ROTDIFF) store_vec<DIM>(a.XIR, s, xi + DIM);
}

// Cell coordinate of a position along one axis, clamped into [0, nc).
__device__ __forceinline__ int cell_coord(double x, double inv_h, int nc)
{
    int c = (int)(x * inv_h);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

template <int DIM>
__device__ __forceinline__ int cell_of(const double* x, const RepParams& P)
{
    int cx = cell_coord(x[0], P.inv_h, P.nc);
    int cy = cell_coord(x[1], P.inv_h, P.nc);
    int cz = DIM == 3 ? cell_coord(x[2], P.inv_h, P.nc) : 0;
    return (cz * P.nc + cy) * P.nc + cx;
}

// How far x lies outside the (periodic) box of cell c, max over axes.
template <int DIM>
__device__ __forceinline__ double outside_cell(const double* x, int c, const RepParams& P)
{
    int cc[3] = {c % P.nc, (c / P.nc) % P.nc, DIM == 3 ? c / (P.nc * P.nc) : 0};
    double worst = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double d = wrap1(x[i] - (cc[i] + 0.5) * P.h, P.L, P.halfL);
        double o = fabs(d) - 0.5 * P.h;
        worst = o > worst ? o : worst;
    }
    return worst;
}

// Half-skin check of a new position of the bead in slot s: its displacement
// since the last re-sort, which keeps the neighbour lists (radius rc + skin)
// complete for every pair within rc, and keeps the bead within half a skin
// of its cell box, which keeps the cell scan complete.
template <int DIM>
__device__ __forceinline__ bool displaced_too_far(const RepParams& P, const double* r, const double* x)
{
    double d2 = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        const double d = wrap1(x[i] - r[i], P.L, P.halfL);
        d2 += d * d;
    }
    return d2 > P.skin_half * P.skin_half;
}

template <int DIM>
__device__ __forceinline__ bool moved_too_far(const PoolArgs& a, const RepParams& P, int s, const double* x)
{
    double r[DIM];
    load_vec<DIM>(a.XR, s, r);
    return displaced_too_far<DIM>(P, r, x);
}

// ------------------------------------------------------------ bonded forces

// Bond vector of the reference (inbounds_diff(pos[u], pos[v]) = x_u - x_v),
// its squared length, inverse length and unit vector (unit = raw when the
// length is zero, as forces.py:326-349 leaves it). BT is the arithmetic
// type: double for the FP64 path, float for the mixed path (the difference
// of the FP64 positions is taken in FP64, then rounded). inv64 is the inverse
// length in FP64 in both: the spring factor a/|t| - 1 cancels for bonds near
// their rest length, so the mixed path forms it in FP64 too.
template <int DIM, typename BT>
struct Bond {
    BT t[DIM];
    BT u[DIM];
    BT ds;
    double inv64;
    double ds64;
};

// 1/sqrt: the FP64 path keeps the correctly rounded form of the reference's
// arithmetic; the mixed path uses the hardware estimate (FP32: 1 ulp) and,
// for the FP64 spring factor, the FP64 estimate refined by one Newton step.
__device__ __forceinline__ double inv_sqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float inv_sqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double inv_sqrt_fast(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    return r * fma(-0.5 * x * r, r, 1.5);
}

template <int DIM, typename BT, bool SPRING = true>
__device__ __forceinline__ void make_bond(const double* xu, const double* xv, const RepParams& P,
                                          Bond<DIM, BT>& b)
{
    double t64[DIM];
    double ds64 = 0.0;
    b.ds = (BT)0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        t64[i] = wrap1(xu[i] - xv[i], P.L, P.halfL);
        b.t[i] = (BT)t64[i];
        b.ds += b.t[i] * b.t[i];
        if (SPRING && sizeof(BT) == 4) ds64 += t64[i] * t64[i];
    }
    b.ds64 = sizeof(BT) == 8 ? (double)b.ds : ds64;
    if (b.ds != (BT)0) {
        const BT inv = inv_sqrt(b.ds);
        b.inv64 = sizeof(BT) == 8 ? (double)inv : (SPRING ? inv_sqrt_fast(ds64) : 0.0);
#pragma unroll
        for (int i = 0; i < DIM; ++i) b.u[i] = b.t[i] * inv;
    } else {
        b.inv64 = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) b.u[i] = b.t[i];
    }
}

// Spring factor f of a bond (the force on its first bead is f * t, t the
// bond vector, forces.py:326-348), in FP64: Hookean k (a/|t| - 1), or FENE
// -k st / |st^2/m^2 - 1| / |t| with st = |t| - a and m = tol * a, with the
// reference's guard (forces.py:336-339), which compares st^2 with m itself
// and divides by zero when it fails. FENE = false compiles Hookean springs
// only (the lean kernels).
template <int DIM, typename BT, bool FENE = true>
__device__ __forceinline__ double spring_factor(const Bond<DIM, BT>& b, double a, const RepParams& P)
{
    if (FENE && (P.flags & GCR_F_FENE)) {
        const double m = P.fene_tol * a;
        const double st = sqrt(b.ds64) - a;
        const double den = (st * st != m && m != 0.0) ? fabs(st * st / (m * m) - 1.0) : 0.0;
        return -P.k * b.inv64 * st * 1.0 / den;
    }
    return P.k * (a * b.inv64 - 1.0);
}

template <typename BT>
__device__ __forceinline__ BT sgn(BT x)
{
    return (BT)((x > (BT)0) - (x < (BT)0));
}

// Bending prefactor phi of forces.py:364-373 for the angle between ti (bond
// leaving the vertex forward) and tj (bond arriving from behind):
// phi = kappa (theta + theta0 sign(.)) / sin(theta). Returns false when the
// term is skipped.
//   FP64: the reference's arithmetic, theta = acos(dot), sin = sqrt(1 - dot^2),
//   skipped when |dot| == 1.
//   FP32 (mixed): acos of a dot product near 1 loses most digits on nearly
//   straight chains, so theta = atan2(|ti x tj|, dot) (atan2_pos) and
//   sin = |ti x tj|, skipped when the cross product is exactly zero.
template <int DIM>
__device__ __forceinline__ bool bend_phi(const double* ti, const double* tj, double kappa, double theta0,
                                         double& dot, double& phi)
{
    dot = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) dot += ti[i] * tj[i];
    if (!(-1.0 < dot && dot < 1.0)) return false;
    phi = acos(dot);
    if (theta0 != 0.0) phi += theta0 * sgn<double>(-ti[0] * tj[1] + ti[1] * tj[0]);
    phi *= kappa / sqrt(1.0 - dot * dot);
    return true;
}

// atan2(y, x) in [0, pi] for y > 0, in FP32: a = min(|x|, y) / max(|x|, y)
// from the hardware reciprocal, atan(a) = a P(a^2) with a degree-7
// polynomial fitted to atan on [0, 1] (weighted least squares on Chebyshev
// nodes; largest error 3e-7, relative error 1.2e-7 near 0), then the octant
// and the sign of x. About 17 instructions against about 40 for the
// IEEE-division atan2f.
__device__ __forceinline__ float atan2_pos(float y, float x)
{
    const float ax = fabsf(x);
    const float lo = fminf(ax, y), hi = fmaxf(ax, y);
    const float a = lo * __builtin_amdgcn_rcpf(hi);
    const float u = a * a;
    float p = -0.004249828401952982f;
    p = fmaf(p, u, 0.022837471216917038f);
    p = fmaf(p, u, -0.05779505893588066f);
    p = fmaf(p, u, 0.09822430461645126f);
    p = fmaf(p, u, -0.13998103141784668f);
    p = fmaf(p, u, 0.19968131184577942f);
    p = fmaf(p, u, -0.3333188593387604f);
    p = fmaf(p, u, 0.9999998807907104f);
    float th = a * p;
    th = y > ax ? 1.5707963267948966f - th : th;
    return x < 0.0f ? 3.141592653589793f - th : th;
}

template <int DIM>
__device__ __forceinline__ bool bend_phi(const float* ti, const float* tj, float kappa, float theta0,
                                         float& dot, float& phi)
{
    dot = 0.0f;
#pragma unroll
    for (int i = 0; i < DIM; ++i) dot += ti[i] * tj[i];
    const float cz = ti[0] * tj[1] - ti[1] * tj[0];
    float cn;
    if (DIM == 2) {
        cn = fabsf(cz);
    } else {
        const float cx = ti[1] * tj[2] - ti[2] * tj[1];
        const float cy = ti[2] * tj[0] - ti[0] * tj[2];
        cn = sqrtf(cx * cx + cy * cy + cz * cz);
    }
    if (cn == 0.0f) return false;
    phi = atan2_pos(cn, dot);
    if (theta0 != 0.0f) phi += theta0 * sgn<float>(-cz);
    phi *= kappa * __builtin_amdgcn_rcpf(cn);
    return true;
}

// Positions of the chain neighbours of a bead (2 and 1 before, 1 and 2
// after); loaded before the pair loop so their latency overlaps it.
template <int DIM>
struct ChainNb {
    int4 lk;
    double p2[DIM], p1[DIM], n1[DIM], n2[DIM];
};

template <int DIM>
__device__ __forceinline__ void load_chain_nb(const double* X, const int32_t* LINK, int s, ChainNb<DIM>& c)
{
    c.lk = *reinterpret_cast<const int4*>(LINK + 4 * s);
    if (c.lk.x >= 0) load_vec<DIM>(X, c.lk.x, c.p2);
    if (c.lk.y >= 0) load_vec<DIM>(X, c.lk.y, c.p1);
    if (c.lk.z >= 0) load_vec<DIM>(X, c.lk.z, c.n1);
    if (c.lk.w >= 0) load_vec<DIM>(X, c.lk.w, c.n2);
}

// All bonded terms acting on the bead at xs, added to acc in the
// reference's order (forces.py:305-402 walks beads in index order: bond
// p1->s, angle at p1, bond s->n1, angle at s, angle at n1). Writes the
// tangent at s when tangents follow the bonds. FENE = false: Hookean only.
template <int DIM, typename BT, bool FENE = true>
__device__ __forceinline__ void bonded(const ChainNb<DIM>& c, const double* xs, double a,
                                       const RepParams& P, double* acc, double* tang, bool& tang_set)
{
    const bool bend = (P.flags & GCR_F_BEND) != 0;
    const bool orient = (P.flags & GCR_F_ORIENT) != 0;
    const BT kappa = (BT)P.kappa, theta0 = (BT)P.theta0;
    const int4 lk = c.lk;
    Bond<DIM, BT> A, B, Pb, Nb;  // (p1,s), (s,n1), (p2,p1), (n1,n2)
    if (lk.y >= 0) make_bond<DIM, BT>(c.p1, xs, P, A);
    if (lk.z >= 0) make_bond<DIM, BT>(xs, c.n1, P, B);
    // bond p1 -> s: forces[b1+1] -= f (forces.py:345-348)
    if (lk.y >= 0 && A.ds != (BT)0) {
        const BT mag = (BT)spring_factor<DIM, BT, FENE>(A, a, P);
#pragma unroll
        for (int i = 0; i < DIM; ++i) acc[i] -= (double)(A.t[i] * mag);
    }
    // angle at p1: forces[b1+1] -= f_bend_backward
    if (bend && lk.x >= 0) {
        make_bond<DIM, BT, false>(c.p2, c.p1, P, Pb);
        BT dot, phi;
        if (bend_phi<DIM>(A.u, Pb.u, kappa, theta0, dot, phi)) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) acc[i] -= (double)(phi * (Pb.u[i] - dot * A.u[i]));
        }
    }
    // bond s -> n1: forces[b1] += f
    if (lk.z >= 0 && B.ds != (BT)0) {
        const BT mag = (BT)spring_factor<DIM, BT, FENE>(B, a, P);
#pragma unroll
        for (int i = 0; i < DIM; ++i) acc[i] += (double)(B.t[i] * mag);
    }
    // angle at s (interior bead), and the tangent
    if (lk.y >= 0 && lk.z >= 0) {
        if (bend || orient) {
            BT dot = (BT)0;
#pragma unroll
            for (int i = 0; i < DIM; ++i) dot += B.u[i] * A.u[i];
            const bool inside = ((BT)-1 < dot && dot < (BT)1);
            BT phi, dot2;
            if (bend && bend_phi<DIM>(B.u, A.u, kappa, theta0, dot2, phi)) {
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    const BT ff = phi * (dot * A.u[i] - B.u[i]);
                    const BT fb = phi * (A.u[i] - dot * B.u[i]);
                    acc[i] += (double)(ff + fb);
                }
            }
            if (orient) {
                if (inside) {
                    const BT cc = inv_sqrt((BT)2 * ((BT)1 + dot));
#pragma unroll
                    for (int i = 0; i < DIM; ++i) tang[i] = (double)(cc * (B.u[i] + A.u[i]));
                } else {
#pragma unroll
                    for (int i = 0; i < DIM; ++i) tang[i] = (double)B.u[i];
                }
                tang_set = true;
            }
        }
    } else if (orient) {
        if (lk.y < 0 && lk.z >= 0) {  // head: first bond
#pragma unroll
            for (int i = 0; i < DIM; ++i) tang[i] = (double)B.u[i];
            tang_set = true;
        } else if (lk.y >= 0 && lk.z < 0) {  // tail: last bond
#pragma unroll
            for (int i = 0; i < DIM; ++i) tang[i] = (double)A.u[i];
            tang_set = true;
        }
    }
    // angle at n1: forces[b1-1] -= f_bend_forward
    if (bend && lk.w >= 0) {
        make_bond<DIM, BT, false>(c.n1, c.n2, P, Nb);
        BT dot, phi;
        if (bend_phi<DIM>(Nb.u, B.u, kappa, theta0, dot, phi)) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) acc[i] -= (double)(phi * (dot * B.u[i] - Nb.u[i]));
        }
    }
}

// ------------------------------------------------------------- pair forces

// Pair law forces.py:13-15 evaluated in FT; the separation and the cutoff
// test are FP64, the returned force component is added to an FP64 sum.
template <typename FT>
struct PairLaw;

template <>
struct PairLaw<double> {
    // 1/ds from v_rcp_f64 refined by two Newton steps (error ~1 ulp)
    static __device__ __forceinline__ double inv(double ds)
    {
        double r = __builtin_amdgcn_rcp(ds);
        double e = fma(-ds, r, 1.0);
        r = fma(r, e, r);
        e = fma(-ds, r, 1.0);
        return fma(r, e, r);
    }
    static __device__ __forceinline__ double coeff(double ds, double p0, double p1)
    {
        double iv = inv(ds);
        double i2 = iv * iv;
        double i3 = i2 * iv;
        return (i2 * i2) * fma(-p1, i3, p0);
    }
    static __device__ __forceinline__ double coeff(double ds, const RepParams& P)
    {
        return coeff(ds, P.wca[0], P.wca[1]);
    }
};

template <>
struct PairLaw<float> {
    static __device__ __forceinline__ float coeff(float ds, float p0, float p1)
    {
        float iv = __builtin_amdgcn_rcpf(ds);
        float i2 = iv * iv;
        float i3 = i2 * iv;
        return (i2 * i2) * fmaf(-p1, i3, p0);
    }
    static __device__ __forceinline__ float coeff(float ds, const RepParams& P)
    {
        return coeff(ds, (float)P.wca[0], (float)P.wca[1]);
    }
};

// Scalars of the pair law, held in registers for a whole pair loop.
template <typename FT>
struct PairConsts {
    FT h, rc2, p0, p1, shift;
    bool shifted;
};

template <typename FT>
__device__ __forceinline__ PairConsts<FT> pair_consts(const RepParams& P)
{
    PairConsts<FT> k;
    k.h = (FT)P.h;
    k.rc2 = (FT)P.rc2;
    k.p0 = (FT)P.wca[0];
    k.p1 = (FT)P.wca[1];
    k.shift = (FT)P.wca[2];
    k.shifted = P.wca[2] != 0.0;
    return k;
}

// Separation along one axis from positions relative to their cell centres
// (rounded to FT) and the cell offset dc in {-1, 0, 1}: dc * h is exact, so
// the result carries one rounding of a cell-sized difference. Every pair path
// (LDS lists, global cell scan) goes through this, in the same operations.
template <typename FT>
__device__ __forceinline__ FT cell_sep(FT dc, FT h, FT rel_j, FT rel_i)
{
    return fma(dc, h, rel_j - rel_i);
}

// 0 < ds <= rc2 (forces.py:122). In FP32 one unsigned compare: for the
// non-negative ds of a sum of squares the bit patterns order as the values,
// and ds = +0 wraps to the largest pattern (NaN is excluded in both forms).
__device__ __forceinline__ bool in_cutoff(double ds, double rc2) { return (ds <= rc2) & (ds > 0.0); }
__device__ __forceinline__ bool in_cutoff(float ds, float rc2)
{
    return __float_as_uint(ds) - 1u < __float_as_uint(rc2);
}

// Force of one candidate at separation sep (pointing to the partner), added
// to f. Branch-free: the cutoff and zero-distance mask keep every lane on one
// path (forces.py:96-118); the law is evaluated for every candidate and its
// value selected away outside the cutoff (at ds = 0 it is not finite, and
// the select drops it).
template <int DIM, typename FT>
__device__ __forceinline__ void pair_term(const FT* sep, const PairConsts<FT>& k, FT* f)
{
    FT ds = sep[0] * sep[0];
#pragma unroll
    for (int i = 1; i < DIM; ++i) ds = fma(sep[i], sep[i], ds);
    const bool use = in_cutoff(ds, k.rc2);
    const FT cf = PairLaw<FT>::coeff(ds, k.p0, k.p1);
    const FT w = use ? cf : (FT)0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) f[i] = fma(sep[i], w, f[i]);
}

// WCA_shift (forces.py:123-131): the cutoff applies to the unshifted distance.
template <int DIM, typename FT>
__device__ __forceinline__ void pair_term_shifted(const FT* sep_in, const PairConsts<FT>& k, FT* f)
{
    FT sep[DIM];
    FT ds = sep_in[0] * sep_in[0];
#pragma unroll
    for (int i = 1; i < DIM; ++i) ds = fma(sep_in[i], sep_in[i], ds);
    if (!(ds <= k.rc2 && ds > (FT)0)) return;
    const FT mag = sqrt(ds);
    FT nm = mag + k.shift;
    nm = nm < (FT)0 ? (FT)0 : nm;
    const FT rs = nm / mag;
#pragma unroll
    for (int i = 0; i < DIM; ++i) sep[i] = sep_in[i] * rs;
    ds = nm * nm;
    const FT cf = PairLaw<FT>::coeff(ds, k.p0, k.p1);
#pragma unroll
    for (int i = 0; i < DIM; ++i) f[i] = fma(sep[i], cf, f[i]);
}

// Position relative to the centre of cell coordinate c along one axis. The
// fma is explicit: left to contraction, a centre hoisted out of a loop would
// be rounded on its own in one path and fused in another.
__device__ __forceinline__ double cell_rel(double x, int c, const RepParams& P)
{
    return wrap1(fma(-((double)c + 0.5), P.h, x), P.L, P.halfL);
}

// Pair forces of the bead at xs (binned in `cell`) over the surrounding
// cells, from global memory, summed in FT into f in the order cells (z, y, x),
// then slots -- the order of the neighbour lists. With 3 or more cells per
// side the separations are formed as in the LDS tiles (cell_sep), so a bead
// gets the same sum from either path; smaller grids scan every cell with the
// plain minimum-image difference.
template <int DIM, typename FT>
__device__ __forceinline__ void pair_sum(const double* X, const int32_t* META, const int32_t* CSTART,
                                         const double* xs, int orig, int chain, int cell,
                                         const RepParams& P, FT* f)
{
    const int nc = P.nc;
    int cc[3] = {cell % nc, (cell / nc) % nc, DIM == 3 ? cell / (nc * nc) : 0};
    const bool rel = nc >= 3;
    const int lo = rel ? -1 : 0;
    const int hi = rel ? 1 : nc - 1;
    const int zlo = DIM == 3 ? lo : 0, zhi = DIM == 3 ? hi : 0;
    const PairConsts<FT> k = pair_consts<FT>(P);
    FT ri[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) ri[i] = (FT)cell_rel(xs[i], cc[i], P);
    for (int dz = zlo; dz <= zhi; ++dz) {
        const int z = DIM == 3 ? (rel ? (cc[2] + dz + nc) % nc : dz) : 0;
        for (int dy = lo; dy <= hi; ++dy) {
            const int y = rel ? (cc[1] + dy + nc) % nc : dy;
            for (int dx = lo; dx <= hi; ++dx) {
                const int x = rel ? (cc[0] + dx + nc) % nc : dx;
                const int cj[3] = {x, y, z};
                const FT dc[3] = {(FT)dx, (FT)dy, (FT)dz};
                const int c2 = P.cell_off + (z * nc + y) * nc + x;
                const int t0 = CSTART[c2], t1 = CSTART[c2 + 1];
                for (int t = t0; t < t1; ++t) {
                    const int2 mj = *reinterpret_cast<const int2*>(META + 2 * t);
                    if (mj.y == chain) {
                        if (P.n_nearby < 0) continue;
                        const int dd = mj.x - orig;
                        if ((dd < 0 ? -dd : dd) <= P.n_nearby) continue;
                    }
                    double xj[DIM];
                    load_vec<DIM>(X, t, xj);
                    FT sep[DIM];
#pragma unroll
                    for (int i = 0; i < DIM; ++i)
                        sep[i] = rel ? cell_sep<FT>(dc[i], k.h, (FT)cell_rel(xj[i], cj[i], P), ri[i])
                                     : (FT)wrap1(xj[i] - xs[i], P.L, P.halfL);
                    if (k.shifted)
                        pair_term_shifted<DIM, FT>(sep, k, f);
                    else
                        pair_term<DIM, FT>(sep, k, f);
                }
            }
        }
    }
}

// Pair forces added to an FP64 accumulator.
template <int DIM, typename FT>
__device__ __forceinline__ void add_pair_sum(const double* X, const int32_t* META, const int32_t* CSTART,
                                             const double* xs, int orig, int chain, int cell,
                                             const RepParams& P, double* acc)
{
    FT f[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) f[i] = (FT)0;
    pair_sum<DIM, FT>(X, META, CSTART, xs, orig, chain, cell, P, f);
#pragma unroll
    for (int i = 0; i < DIM; ++i) acc[i] += (double)f[i];
}

// Pair forces of a bead from the cell scan, added to its FP64 force (the
// force evaluations outside the tile kernels).
template <int DIM, typename FT>
__device__ __forceinline__ void add_pairs(const PoolArgs& a, const RepParams& P, const double* X, int s,
                                          const double* xs, int2 me, double* acc)
{
    add_pair_sum<DIM, FT>(X, a.META, a.CSTART, xs, me.x, me.y, a.CELL[s], P, acc);
}

// ------------------------------------------------------------------ stages

// Max over the block (wave64 shuffles, then the block's waves through LDS);
// the result is valid in thread 0.
template <int NT = BS>
__device__ __forceinline__ double block_max(double v)
{
    __shared__ double red[NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int i = 1; i < NT / 64; ++i) m = red[i] > m ? red[i] : m;
    __syncthreads();
    return m;
}

// Self-propulsion (forces.py:206-223), per-chain (v0, cos a, sin a).
template <int DIM>
__device__ __forceinline__ void propel(const PoolArgs& a, int chain, const double* tang, double* acc)
{
    const double* cs = a.CHAIN_CS + 3 * chain;
    if (DIM == 2) {
        acc[0] += cs[0] * (cs[1] * tang[0] - cs[2] * tang[1]);
        acc[1] += cs[0] * (cs[2] * tang[0] + cs[1] * tang[1]);
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i) acc[i] += cs[0] * tang[i];
    }
}

// A virtual update x1 = (x + dt resc / r F + sqrt(2 D dt) xi) mod L with the
// previous step's dt and this step's noise xi, which every stage shares
// (SRK.py:355-388). A rod replica's x1 is replaced (and checked) by k_rod.
template <int DIM>
__device__ __forceinline__ void virtual_update_in(const PoolArgs& a, const RepParams& P, RepState* S,
                                                  int s, const double* x, const double* f,
                                                  double resc, int sub, const double* xi,
                                                  const double* xr, double dt)
{
    double dt_eff = dt * resc * 1.0;
    const double s2 = sqrt(2.0 * dt * P.D);
    dt_eff /= P.radius;
    double x1[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double tmp = x[i] + dt_eff * f[i];
        if (s2 != 0.0) tmp += s2 * xi[i];
        x1[i] = pymod(tmp, P.L);
    }
    store_vec<DIM>(a.X1, s, x1);
    if (!(P.flags & GCR_F_ROD) && displaced_too_far<DIM>(P, xr, x1)) flag_predictor(a, S, sub);
}

// The same, loading the bead's noise and re-sort position (xi: this step's
// noise, which every stage shares; xr: the position at the last re-sort).
template <int DIM>
__device__ __forceinline__ void virtual_update(const PoolArgs& a, const RepParams& P, RepState* S,
                                               int s, const double* x, const double* f,
                                               double resc, int sub)
{
    double xi[DIM], xr[DIM];
    load_vec<DIM>(a.XI, s, xi);
    load_vec<DIM>(a.XR, s, xr);
    virtual_update_in<DIM>(a, P, S, s, x, f, resc, sub, xi, xr, S->dt);
}

// WCA walls normal to the axes in P.walls (forces.py:405-435), `wall_zone`
// inside the faces, evaluated at x = the bead's position at the start of the
// step in every stage (SRK.py:270 passes pos): sep from the nearer wall,
// reduced by WCA_shift (forces.py:425-429, the opposite sign convention to
// the pair law's), and the force -sep coeff(sep^2) while 0 < sep^2 <= the
// walls' cutoff -- the squared distance against the cutoff itself, as
// forces.py:432 has it.
// The reference visits only beads whose last grid cell lies within
// 2 grid_window cells of a face; all beads in the walls' range do while beads
// move less than grid_window cells between grid updates.
template <int DIM>
__device__ __forceinline__ void add_walls(const RepParams& P, const double* x, double* acc)
{
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        if (!((P.walls >> i) & 1)) continue;
        double sep = x[i];
        if (sep >= P.halfL)
            sep -= (P.L - P.wall_zone);
        else
            sep -= P.wall_zone;
        double ds;
        if (P.wca[2] != 0.0) {
            double m = fabs(sep) - P.wca[2];
            m = m > 0.0 ? m : 0.0;
            sep = sep > 0.0 ? m : (sep < 0.0 ? -m : 0.0);
            ds = m * m;
        } else {
            ds = sep * sep;
        }
        if (ds > 0.0 && ds <= P.wall_rc) acc[i] += -sep * PairLaw<double>::coeff(ds, P);
    }
}

// What a force stage does with the total force on slot s:
// STAGE 0 (forces at x): store F0; SRK4 also starts the weighted total
//          FT = F0/6; the virtual update from x (SRK2: the predictor);
// STAGE 1 (forces at x1), SRK2: store F = F0 + F1 and return |F|^2 for the
//          max-force reduction; SRK4 stage 1, 2: FT += F/3, F0 = F (the
//          virtual update runs in k_virtual once every bead's forces are
//          known); SRK4 stage 3: FT += F/6 and return |FT|^2;
// STAGE 2: (force evaluation only) store the force.
// GEN = false compiles the SRK2 path only (the kernels' lean variant).
template <int DIM, int STAGE, bool GEN = true>
__device__ __forceinline__ double stage_store(const PoolArgs& a, const RepParams& P, RepState* S,
                                              int s, int orig_global, const double* xs,
                                              const double* acc)
{
    const bool srk4 = GEN && (P.flags & GCR_F_SRK4) != 0;
    if (GEN && STAGE != 2 && (P.flags & (GCR_F_GROUP | GCR_F_SLIDE))) {
        // a group member's or a sliding replica's own forces; the sliding
        // pushes, the other populations' forces, the virtual update and the
        // max follow in k_group_stage, once every bead's tangent is known
        store_vec<DIM>(STAGE == 0 ? a.F0 : a.FT, s, acc);
        if (!(P.flags & GCR_F_MOTOR)) {
            if (STAGE == 0) a.WA[s] = -1;  // SRK.py:132
            a.MC[s] = 0;
        }
        return 0.0;
    }
    if (STAGE == 0) {
        store_vec<DIM>(a.F0, s, acc);
        if (srk4) {
            double ft[DIM];
#pragma unroll
            for (int i = 0; i < DIM; ++i) ft[i] = acc[i] * (1. / 6.);
            store_vec<DIM>(a.FT, s, ft);
        }
        virtual_update<DIM>(a, P, S, s, xs, acc, GEN ? virtual_rescale(P, 0) : 1.0, 0);
        return 0.0;
    } else if (!srk4) {
        store_vec<DIM>(a.FT, s, acc);
        double q = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) q += acc[i] * acc[i];
        return q;
    } else {
        const double w = a.sub == 3 ? 1. / 6. : 1. / 3.;
        double ft[DIM];
        load_vec<DIM>(a.FT, s, ft);
#pragma unroll
        for (int i = 0; i < DIM; ++i) ft[i] = ft[i] + acc[i] * w;
        store_vec<DIM>(a.FT, s, ft);
        if (a.sub < 3) {
            store_vec<DIM>(a.F0, s, acc);
            return 0.0;
        }
        double q = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) q += ft[i] * ft[i];
        return q;
    }
}

// SRK4 virtual update of stages 1 and 2 (forces of the stage in F0), from x.
template <int DIM>
__global__ void __launch_bounds__(BS) k_virtual(PoolArgs a, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (!(P.flags & GCR_F_SRK4) || (P.flags & GCR_F_ROD) || !step_runs(S, j)) return;
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (s >= P.bead_off + P.n) return;
    double x[DIM], f[DIM];
    load_vec<DIM>(a.X, s, x);
    load_vec<DIM>(a.F0, s, f);
    virtual_update<DIM>(a, P, S, s, x, f, virtual_rescale(P, a.sub), a.sub);
}

// Step bookkeeping of a replica from the largest |F|^2 over its blocks
// (SRK.py:323-353): the adaptive dt, dt_used clipped to max_dt_usable, the
// clock and the step counters. A group member only reports its part
// (k_group_dt takes the group's max). One thread.
__device__ __forceinline__ void finish_step(const RepParams& P, RepState* S, double m2)
{
    S->sort_due = 0;  // this step's re-sort, if due, has run
    double m = sqrt(2.0 * P.D);
    if (P.n > 0) m += force_prefactor(P) * sqrt(m2) / P.radius;
    if (P.flags & GCR_F_GROUP) {
        S->mpart = m;
        return;
    }
    double dt = (m == 0.0) ? P.dtmax : fmin(P.dtmax, P.mdps / m);
    double dt_used = P.cheat * dt;
    if (S->mdu > 0) dt_used = fmin(dt_used, S->mdu - S->elapsed);
    S->dt = dt;
    S->dt_used = dt_used;
    S->elapsed += dt_used;
    S->t_total += dt_used;
    S->maxm = m;
    S->steps += 1;
    S->step_global += 1;
}

// Max force and adaptive dt of each replica, one block per replica, after
// the stage-1 launch has written every block's maximum.
__global__ void __launch_bounds__(BS) k_finalize(PoolArgs a, int j)
{
    const int r = blockIdx.x;
    RepState* S = a.S + r;
    const RepParams& P = a.P[r];
    // the reduction's extent is read with the run test's reads
    const int red_n = P.red_n, red_off = P.red_off;
    asm volatile("" ::"s"(red_n), "s"(red_off));
    if (!step_runs(S, j)) return;
    double m2 = 0.0;
    for (int i = threadIdx.x; i < red_n; i += BS) {
        double v = a.BMAX[red_off + i];
        m2 = v > m2 ? v : m2;
    }
    m2 = block_max(m2);
    if (threadIdx.x == 0) finish_step(P, S, m2);
}

// ------------------------------------------------------------ populations

// Calls fn(t) for every slot t of the cells within w cells (per axis) of the
// cell holding x in a grid of nc cells per side starting at cell_off (every
// stride-th slot of each range from its first + `first`, for lanes that split
// a scan). Slots are sorted by cell, so the cells of one row of the window
// are one slot range, or two where the window wraps round the box; the whole
// grid when 2w + 1 >= nc.
template <int DIM, typename Fn>
__device__ __forceinline__ void for_window(const PoolArgs& a, int cell_off, int nc, double inv_h,
                                           const double* x, int w, Fn&& fn, int first = 0,
                                           int stride = 1)
{
    int c[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < DIM; ++i) c[i] = cell_coord(x[i], inv_h, nc);
    const bool all = 2 * w + 1 >= nc;
    const int lo = all ? 0 : -w, hi = all ? nc - 1 : w;
    for (int dz = (DIM == 3 ? lo : 0); dz <= (DIM == 3 ? hi : 0); ++dz) {
        const int z = DIM == 3 ? (all ? dz : (c[2] + dz + nc) % nc) : 0;
        for (int dy = lo; dy <= hi; ++dy) {
            const int y = all ? dy : (c[1] + dy + nc) % nc;
            const int row = cell_off + (z * nc + y) * nc;
            int x0 = all ? 0 : c[0] - w, x1 = all ? nc - 1 : c[0] + w;  // inclusive
            if (x0 < 0) {  // the part beyond the lower face
                const int t1 = a.CSTART[row + nc];
                for (int t = a.CSTART[row + nc + x0] + first; t < t1; t += stride) fn(t);
                x0 = 0;
            } else if (x1 >= nc) {  // the part beyond the upper face
                const int t1 = a.CSTART[row + x1 - nc + 1];
                for (int t = a.CSTART[row] + first; t < t1; t += stride) fn(t);
                x1 = nc - 1;
            }
            const int t1 = a.CSTART[row + x1 + 1];
            for (int t = a.CSTART[row + x0] + first; t < t1; t += stride) fn(t);
        }
    }
}

// Forces on a bead at x from the beads of another population Q
// (forces.py:57-144 with same_pop False: every Q bead within the law's
// cutoff, no exclusions, the force coeff * (x_q - x), WCA_shift as in the
// pair law), in FP64, found by scanning Q's cells within e.window of the
// cell holding x in Q's grid (the window covers the cutoff plus the half skin
// by which Q's beads may have left their cells).
template <int DIM>
__device__ __forceinline__ void cross_sum(const PoolArgs& a, const RepParams& Q, const double* XQ,
                                          const double* x, const XPair& e, double* f, int first = 0,
                                          int stride = 1)
{
    const double rc2 = e.rc * e.rc;
    for_window<DIM>(
        a, Q.cell_off, Q.nc, Q.inv_h, x, e.window,
        [&](int t) {
        double q[DIM], sep[DIM], ds = 0.0;
        load_vec<DIM>(XQ, t, q);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            double d = q[i] - x[i];
            if (d < -Q.halfL)
                d += Q.L;
            else if (d >= Q.halfL)
                d -= Q.L;
            sep[i] = d;
            ds += d * d;
        }
        if (!(ds > 0.0 && ds <= rc2)) return;
        if (e.shift != 0.0) {
            const double mag = sqrt(ds);
            double nm = mag + e.shift;
            nm = nm > 0.0 ? nm : 0.0;
            const double sc = nm / mag;
#pragma unroll
            for (int i = 0; i < DIM; ++i) sep[i] *= sc;
            ds = nm * nm;
        }
        const double cf = PairLaw<double>::coeff(ds, e.p0, e.p1);
#pragma unroll
        for (int i = 0; i < DIM; ++i) f[i] += sep[i] * cf;
        },
        first, stride);
}

// Active sliding partners of the bead in slot s (forces.py:79-83, 114-120):
// beads of its own replica closer than the sliding range (strictly) whose
// tangent points against the bead's, the same-chain rule applying, and a
// partner in the bead's own thread chunk only when its index is lower (the
// reference skips b2 >= b1 within a chunk, a pair whose pair force it
// applies to both beads but whose push it gives only to the later one).
// Found by scanning the replica's cells within the range plus the half skin.
template <int DIM>
__device__ __forceinline__ int slide_partners(const PoolArgs& a, const RepParams& P, const double* XA,
                                              int s, const double* x, const double* t1, int first = 0,
                                              int stride = 1)
{
    const int nc = P.nc;
    const int2 me = *reinterpret_cast<const int2*>(a.META + 2 * s);
    const int ch1 = a.CHAIN_CHUNK[me.y];
    const int nn = P.n_nearby;
    const double wd = ceil((sqrt(P.slide_d2) + fmin(P.skin_half, P.L)) * P.inv_h);
    const int w = wd >= nc ? nc : max(1, (int)wd);
    int cnt = 0;
    for_window<DIM>(
        a, P.cell_off, nc, P.inv_h, x, w,
        [&](int t) {
        // range first: the rules below load more, for few beads
        double q[DIM], ds = 0.0;
        load_vec<DIM>(XA, t, q);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            double sep = q[i] - x[i];
            if (sep < -P.halfL)
                sep += P.L;
            else if (sep >= P.halfL)
                sep -= P.L;
            ds += sep * sep;
        }
        if (!(ds < P.slide_d2)) return;
        const int2 m = *reinterpret_cast<const int2*>(a.META + 2 * t);
        const int d = m.x - me.x;
        if (m.y == me.y && (nn < 0 || (d < 0 ? -d : d) <= nn)) return;
        if (m.x >= me.x && a.CHAIN_CHUNK[m.y] == ch1) return;
        double t2[DIM], dot = 0.0;
        load_vec<DIM>(a.Tn, t, t2);
#pragma unroll
        for (int i = 0; i < DIM; ++i) dot += t1[i] * t2[i];
        if (dot < 0.0) ++cnt;
        },
        first, stride);
    return cnt;
}

// Active sliding partners of every bead of a replica with GCR_F_SLIDE at
// this force stage (x in stage 0, x1 in stage 1; tangents of the stage from
// Tn), into SCNT for k_group_stage. SLIDE_LANES lanes per bead split the
// window's slots and add their counts.
template <int DIM, int STAGE>
__global__ void __launch_bounds__(BS) k_slide_count(PoolArgs a, int j, int blk0)
{
    const int e = blk0 + (int)blockIdx.x / SLIDE_LANES;
    const int r = a.blk_rep[e];
    const RepParams& P = a.P[r];
    if (!(P.flags & GCR_F_SLIDE) || !step_runs(a.S + r, j)) return;
    const int g = threadIdx.x % SLIDE_LANES;
    const int s = a.blk_slot0[e] + ((int)blockIdx.x % SLIDE_LANES) * (BS / SLIDE_LANES) +
                  (int)threadIdx.x / SLIDE_LANES;
    if (s >= P.bead_off + P.n) return;  // the whole group
    const double* XA = STAGE == 0 ? a.X : a.X1;
    double x[DIM], t1[DIM];
    load_vec<DIM>(XA, s, x);
    load_vec<DIM>(a.Tn, s, t1);
    int cnt = slide_partners<DIM>(a, P, XA, s, x, t1, g, SLIDE_LANES);
#pragma unroll
    for (int k = 1; k < SLIDE_LANES; k <<= 1) cnt += __shfl_xor(cnt, k, 64);
    if (g == 0) a.SCNT[s] = cnt;
}

// Second half of a force stage for the moving populations of a group and for
// replicas with active sliding: the sliding pushes (partners from
// k_slide_count), then the forces of the other populations, in the order
// (Q - P) mod n (SRK.py:287-293), at their positions of this stage (x1 in
// stage 1 unless fixed), added to the bead's own forces; then STAGE 0 the
// predictor, STAGE 1 the block's max |F|^2 for k_finalize. GROUP_LANES lanes
// per bead split the cell windows of the other populations and add their
// partial sums; the first lane writes.
template <int DIM, int STAGE>
__global__ void __launch_bounds__(BS) k_group_stage(PoolArgs a, int j)
{
    const int eb = (int)blockIdx.x / GROUP_LANES;
    const int r = a.blk_rep[eb];
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (!(P.flags & (GCR_F_GROUP | GCR_F_SLIDE)) || !step_runs(S, j)) return;
    const int g = threadIdx.x % GROUP_LANES;
    const int s = a.blk_slot0[eb] + ((int)blockIdx.x % GROUP_LANES) * (BS / GROUP_LANES) +
                  (int)threadIdx.x / GROUP_LANES;
    double fsq = 0.0;
    if (s < P.bead_off + P.n) {  // the whole group of lanes
        const double* XA = STAGE == 0 ? a.X : a.X1;
        double x[DIM], f[DIM];
        load_vec<DIM>(XA, s, x);
        load_vec<DIM>(STAGE == 0 ? a.F0 : a.FT, s, f);
        if (P.flags & GCR_F_SLIDE) {
            double t1[DIM];
            load_vec<DIM>(a.Tn, s, t1);
            const int cnt = a.SCNT[s];  // k_slide_count
            for (int k = 0; k < cnt; ++k)
#pragma unroll
                for (int i = 0; i < DIM; ++i) f[i] += P.slide_v0 * t1[i];
            if (g == 0) a.WA[s] = (STAGE == 0 ? -1 : a.WA[s]) + cnt;  // SRK.py:132, forces.py:119
        }
        for (int k = P.xin_off; k < P.xin_off + P.xin_n; ++k) {
            const XPair e = a.XPAIRS[k];
            if (e.kind == 2) {  // the bindings k_motor counted, imparts * t each
                double t1[DIM];
                load_vec<DIM>(a.Tn, s, t1);
                const double imp = a.P[e.src].motor_imparts;
                for (int c = a.MC[s]; c > 0; --c)
#pragma unroll
                    for (int i = 0; i < DIM; ++i) f[i] += imp * t1[i];
                continue;
            }
            const double* XQ = (STAGE == 0 || a.S[e.src].fixed) ? a.X : a.X1;
            double fc[DIM];
#pragma unroll
            for (int i = 0; i < DIM; ++i) fc[i] = 0.0;
            cross_sum<DIM>(a, a.P[e.src], XQ, x, e, fc, g, GROUP_LANES);
            // butterfly: every lane ends with the same sum (each add commutes)
#pragma unroll
            for (int m = 1; m < GROUP_LANES; m <<= 1)
#pragma unroll
                for (int i = 0; i < DIM; ++i) fc[i] += __shfl_xor(fc[i], m, 64);
#pragma unroll
            for (int i = 0; i < DIM; ++i) f[i] += fc[i];
        }
        if (STAGE == 0) {
            if (g == 0) {
                store_vec<DIM>(a.F0, s, f);
                // a bound motor that does not diffuse keeps its last x1
                // (SRK.py:389-396 writes only the movers)
                if (!((P.flags & GCR_F_MSTILL) && a.WA[s] >= 0)) virtual_update<DIM>(a, P, S, s, x, f, 1.0, 0);
            }
        } else {
            if (g == 0) store_vec<DIM>(a.FT, s, f);
#pragma unroll
            for (int i = 0; i < DIM; ++i) fsq += f[i] * f[i];
        }
    }
    if (STAGE == 1) {
        const double m = block_max(fsq);
        if (threadIdx.x == 0) a.BMAX[P.red_off + ((int)blockIdx.x - GROUP_LANES * P.blk_off)] = m;
    }
}

// The motor rule's uniform draw: Philox4x32-10 keyed by the motor
// population's seed, counter (motor, bead, step lo, step hi << 8 | 0x80 |
// stage << 1 | kind), 53 bits in [0, 1). Same definition as the oracle's
// orc_motor_uniform.
__device__ __forceinline__ double motor_uniform(uint64_t seed, uint32_t motor, uint32_t bead,
                                                int64_t step, int stage, int kind)
{
    uint32_t c[4] = {motor, bead, (uint32_t)(uint64_t)step,
                     ((uint32_t)((uint64_t)step >> 32) << 8) | 0x80u | ((uint32_t)stage << 1) |
                         (uint32_t)kind};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * (1.0 / 9007199254740992.0);
}

// Motors (forces.py:248-283), one thread per motor of a motor population of a
// group: the beads of the pushed population within its reach, in index
// order (each motor's history depends on the beads in order and on nothing
// else, so the reference's bead-major loop is the same as this motor-major
// one): a bound motor comes free when its draw is <= k_off (the visiting
// bead's where_active drops by one), and a free motor binds to the visiting
// bead when its draw is < k_on and the bead is within the motor range
// (where_active and the bead's binding count rise by one; k_group_stage adds
// imparts * t per binding). A moving motor population starts each step free
// (SRK.py:132); a fixed one keeps its bindings.
template <int DIM, int STAGE>
__global__ void __launch_bounds__(BS) k_motor(PoolArgs a, int j, int blk0)
{
    // MOTOR_LANES consecutive lanes per motor: they split the window's slots
    // and agree on the next bead through a minimum over the group, so every
    // decision below is the same on the group's lanes
    const int e = blk0 + (int)blockIdx.x / MOTOR_LANES;
    const int q = a.blk_rep[e];
    const RepParams& Q = a.P[q];
    if (!(Q.flags & GCR_F_MOTOR) || Q.motor_target < 0) return;
    const int tr = Q.motor_target;
    const RepParams& T = a.P[tr];
    const RepState* ST = a.S + tr;
    if (!step_runs(ST, j)) return;
    const int g = threadIdx.x % MOTOR_LANES;
    const int m = a.blk_slot0[e] + ((int)blockIdx.x % MOTOR_LANES) * (BS / MOTOR_LANES) +
                  (int)threadIdx.x / MOTOR_LANES;
    if (m >= Q.bead_off + Q.n) return;  // the whole group
    const bool qfixed = a.S[q].fixed != 0;
    double xm[DIM];
    load_vec<DIM>((STAGE == 0 || qfixed) ? a.X : a.X1, m, xm);
    const double* XT = STAGE == 0 ? a.X : a.X1;
    int state = a.WA[m];
    if (!qfixed && STAGE == 0) state = -1;
    const uint32_t mloc = (uint32_t)(a.META[2 * m] - Q.bead_off);
    const int64_t step = ST->step_global;
    const int nc = T.nc;
    const double reach2 = Q.motor_reach * Q.motor_reach;
    const double wd = ceil((Q.motor_reach + fmin(T.skin_half, T.L)) * T.inv_h);
    const int w = wd >= nc ? nc : max(1, (int)wd);
    int last = -1;  // the beads within reach, visited in increasing index
    for (;;) {
        int best = 0x7fffffff, bslot = -1;
        double bd2 = 0.0;
        for_window<DIM>(
            a, T.cell_off, nc, T.inv_h, xm, w,
            [&](int t) {
                // range first: most candidates are out of reach
                double xb[DIM], d2 = 0.0;
                load_vec<DIM>(XT, t, xb);
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    double sep = xm[i] - xb[i];
                    if (sep < -T.halfL)
                        sep += T.L;
                    else if (sep >= T.halfL)
                        sep -= T.L;
                    d2 += sep * sep;
                }
                if (d2 > reach2) return;
                const int o = a.META[2 * t];
                if (o <= last || o >= best) return;
                best = o;
                bslot = t;
                bd2 = d2;
            },
            g, MOTOR_LANES);
#pragma unroll
        for (int k = 1; k < MOTOR_LANES; k <<= 1) {
            const int ob = __shfl_xor(best, k, 64);
            const int os = __shfl_xor(bslot, k, 64);
            const double od = __shfl_xor(bd2, k, 64);
            if (ob < best) {
                best = ob;
                bslot = os;
                bd2 = od;
            }
        }
        if (bslot < 0) break;
        last = best;
        const uint32_t bloc = (uint32_t)(best - T.bead_off);
        if (state >= 0) {
            if (motor_uniform(Q.seed, mloc, bloc, step, STAGE, 0) > Q.k_off) continue;
            state = -1;
            if (g == 0) atomicAdd(&a.WA[bslot], -1);
        }
        if (motor_uniform(Q.seed, mloc, bloc, step, STAGE, 1) < Q.k_on && bd2 <= Q.motor_d2) {
            if (g == 0) {
                atomicAdd(&a.WA[bslot], 1);
                atomicAdd(&a.MC[bslot], 1);
            }
            state = (int)bloc;
        }
    }
    if (g == 0) a.WA[m] = state;
}

// One dt for a group: the largest member part (SRK.py:336-337 takes the max
// over every thread of every population), then the step bookkeeping of
// k_finalize for each moving member.
__global__ void k_group_dt(PoolArgs a, int j)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.nrep) return;
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (!(P.flags & GCR_F_GROUP) || !step_runs(S, j)) return;
    double m = 0.0;
    for (int k = P.grp_first; k < P.grp_first + P.grp_n; ++k) {
        const double v = a.S[a.GRP[k]].mpart;
        m = v > m ? v : m;
    }
    double dt = (m == 0.0) ? P.dtmax : fmin(P.dtmax, P.mdps / m);
    double dt_used = P.cheat * dt;
    if (S->mdu > 0) dt_used = fmin(dt_used, S->mdu - S->elapsed);
    S->dt = dt;
    S->dt_used = dt_used;
    S->elapsed += dt_used;
    S->t_total += dt_used;
    S->maxm = m;
    S->steps += 1;
    S->step_global += 1;
}

// Corrector update (SRK.py:355-388) and rotational diffusion (SRK.py:398-423).
// The replica's parameters and step state are read before the first store
// (after one the compiler could not keep them in scalar registers: every use
// would be a vector load waiting on memory), and every input of the bead is
// in flight at once, (orig, chain) first, so that the next step's noise is
// drawn while the rest arrives. Before that, the block's table entries, then
// the replica's state and extent, are each read in one go (one wait on
// memory each, rather than one per test).
template <int DIM>
__global__ void __launch_bounds__(BS) k_update(PoolArgs a, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const int slot0 = a.blk_slot0[blockIdx.x];
    const RepParams P = a.P[r];
    const RepState* S = a.S + r;
    const int32_t frozen = S->frozen;
    const int64_t steps = S->steps, base = S->batch_base, step_global = S->step_global;
    const double dt_used = S->dt_used;
    const int s = slot0 + threadIdx.x;
    if ((frozen != 0) | (steps != base + j + 1) | (s >= P.bead_off + P.n)) return;
    const bool rot = (P.flags & GCR_F_ROTDIFF) != 0;
    // rod replicas: the corrector ran in k_rod; a bound motor stays put
    const bool moves = !(P.flags & GCR_F_ROD) && !((P.flags & GCR_F_MSTILL) && a.WA[s] >= 0);
    // loaded whether or not the bead moves (in bounds either way), so that
    // no branch joins behind them and the noise waits for (orig, chain) only
    const int orig = a.META[2 * s];
    double x[DIM], f[DIM], xi[2 * DIM], xr[DIM], t[DIM];
    load_vec<DIM>(a.X, s, x);
    load_vec<DIM>(a.FT, s, f);
    load_vec<DIM>(a.XI, s, xi);
    load_vec<DIM>(a.XR, s, xr);
    // the next step's noise (finalize already advanced the step counters)
    double xin[2 * DIM];
    draw_noise<DIM>(P, orig, steps, step_global, xin);
    if (rot) {
        load_vec<DIM>(a.T, s, t);
        load_vec<DIM>(a.XIR, s, xi + DIM);
    }
    if (moves) {
        double dt_eff = dt_used * force_prefactor(P);
        const double s2 = sqrt(2.0 * dt_used * P.D);
        dt_eff /= P.radius;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            double tmp = x[i] + dt_eff * f[i];
            if (s2 != 0.0) tmp += s2 * xi[i];
            x[i] = pymod(tmp, P.L);
        }
        store_vec<DIM>(a.X, s, x);
        // outside the safe batch a replay, inside it a re-sort before the next step
        if (displaced_too_far<DIM>(P, xr, x)) atomicOr(a.safe ? &a.S[r].sort_due : &a.S[r].violation, 1);
    }
    // the step is over: the next mid-step re-sort waits for a new predictor
    if (threadIdx.x == 0 && blockIdx.x == P.blk_off) a.S[r].pred_far = 0;
    if (rot) {
        const double sr = sqrt(2.0 * dt_used * P.Dr) * 3.141592653589793;
        double nr[DIM], proj = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) nr[i] = xi[DIM + i];
#pragma unroll
        for (int i = 0; i < DIM; ++i) proj += nr[i] * t[i];
#pragma unroll
        for (int i = 0; i < DIM; ++i) nr[i] -= proj * t[i];
        double mag = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            t[i] = t[i] + sr * nr[i];
            mag += t[i] * t[i];
        }
        mag = sqrt(mag);
#pragma unroll
        for (int i = 0; i < DIM; ++i) t[i] /= mag;
        store_vec<DIM>(a.T, s, t);
    }
    store_vec<DIM>(a.XI, s, xin);
    if (rot) store_vec<DIM>(a.XIR, s, xin + DIM);
}

// One force evaluation of the bead in slot s without an update: forces at X
// (which == 0) or X1 (which == 1) into FT, tangents into Tn.
template <int DIM, typename FT>
__device__ __forceinline__ void eval_bead(const PoolArgs& a, const RepParams& P, int s, int which)
{
    const double* Xa = which == 0 ? a.X : a.X1;
    double xs[DIM], acc[DIM], tang[DIM];
    load_vec<DIM>(Xa, s, xs);
#pragma unroll
    for (int i = 0; i < DIM; ++i) acc[i] = 0.0;
    int2 me = *reinterpret_cast<const int2*>(a.META + 2 * s);
    bool tang_set = false;
    if (P.flags & GCR_F_SPRINGS) {
        ChainNb<DIM> cn;
        load_chain_nb<DIM>(Xa, a.LINK, s, cn);
        bonded<DIM, FT>(cn, xs, a.CHAIN_A[me.y], P, acc, tang, tang_set);
    }
    if (P.flags & GCR_F_WCA) add_pairs<DIM, FT>(a, P, Xa, s, xs, me, acc);
    if (!tang_set) load_vec<DIM>(a.T, s, tang);
    if (P.flags & GCR_F_PROPEL) propel<DIM>(a, me.y, tang, acc);
    if (P.flags & GCR_F_WALLS) {
        double x0[DIM];
        load_vec<DIM>(a.X, s, x0);
        add_walls<DIM>(P, x0, acc);
    }
    store_vec<DIM>(a.FT, s, acc);
    store_vec<DIM>(a.Tn, s, tang);
}

// ----------------------------------------------------------------- rebuild

// Every chain's speed (the first of its three CHAIN_CS entries) set to v0.
__global__ void __launch_bounds__(BS) k_set_speed(double* cs, int n_chains, double v0)
{
    const int c = blockIdx.x * BS + threadIdx.x;
    if (c < n_chains) cs[3 * c] = v0;
}

__global__ void __launch_bounds__(BS) k_fill_i32(int32_t* p, int n, int32_t v)
{
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) p[i] = v;
}

// Beads are stored sorted by cell, so neighbouring lanes mostly fall in the
// same cell: each run of equal cells in a wave takes one atomic for the whole
// run (its first lane adds the run length) and the lanes of the run number
// themselves from the returned base. Offsets only need to be distinct within
// a cell; the gather orders each cell by original index.
template <int DIM>
__global__ void __launch_bounds__(BS) k_bin(PoolArgs a, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    const bool live = s < P.bead_off + P.n;
    int c = -1;
    if (live) {
        double x[DIM];
        load_vec<DIM>(a.on_x1 ? a.X1 : a.X, s, x);
        c = cell_of<DIM>(x, P);
    }
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(c, 1, 64);
    const bool head = lane == 0 || prev != c;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int hl = 63 - __clzll((long long)(heads & upto));  // head of my run (lane 0 always is one)
    const unsigned long long after = heads & ~upto;         // heads past my lane
    const int end = after ? __ffsll(after) - 1 : 64;        // one past my run
    int base = 0;
    if (head && live) base = atomicAdd(&a.CCOUNT[P.cell_off + c], end - lane);
    base = __shfl(base, hl, 64);
    if (live) *reinterpret_cast<int2*>(a.KEYOFF + 2 * s) = make_int2(c, base + lane - hl);
}

// Inclusive wave64 scan with shuffles, exclusive block scan through LDS.
template <int NT = BS>
__device__ __forceinline__ int block_exclusive_scan(int v, int* total)
{
    __shared__ int wsum[NT / 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int o = __shfl_up(inc, off, 64);
        if (l >= off) inc += o;
    }
    if (l == 63) wsum[w] = inc;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        if (i < w) base += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

#define SCAN_PER_THREAD 8
#define SCAN_CHUNK (BS * SCAN_PER_THREAD)

// scan phase A: one chunk of SCAN_CHUNK cells of one replica -> its sum
__global__ void __launch_bounds__(BS) k_scan_chunks(PoolArgs a, int j)
{
    const int r = a.chk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int c0 = a.chk_cell0[blockIdx.x];
    const int cend = P.cell_off + P.ncell;
    int sum = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        int c = c0 + threadIdx.x * SCAN_PER_THREAD + k;
        if (c < cend) sum += a.CCOUNT[c];
    }
    int tot;
    block_exclusive_scan(sum, &tot);
    if (threadIdx.x == 0) a.CHKSUM[blockIdx.x] = tot;
}

// scan phase B: one block per replica scans its chunk sums in place
__global__ void __launch_bounds__(BS) k_scan_reps(PoolArgs a, int j)
{
    const int r = blockIdx.x;
    if (!rebuild_runs(a, r, j)) return;
    const int k0 = a.rep_chk_off[r], k1 = a.rep_chk_off[r + 1];
    int carry = 0;
    for (int base = k0; base < k1; base += BS) {
        int k = base + threadIdx.x;
        int v = k < k1 ? a.CHKSUM[k] : 0;
        int tot;
        int ex = block_exclusive_scan(v, &tot);
        if (k < k1) a.CHKSUM[k] = carry + ex;
        carry += tot;
    }
}

// scan phase C: exclusive scan inside each chunk, offset by the chunk prefix
// and the replica's first slot -> CSTART; the last chunk also writes the end
__global__ void __launch_bounds__(BS) k_scan_final(PoolArgs a, int j)
{
    const int r = a.chk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int c0 = a.chk_cell0[blockIdx.x];
    const int cend = P.cell_off + P.ncell;
    int v[SCAN_PER_THREAD];
    int sum = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        int c = c0 + threadIdx.x * SCAN_PER_THREAD + k;
        v[k] = c < cend ? a.CCOUNT[c] : 0;
        if (c < cend) a.CCOUNT[c] = 0;  // counts start from zero at the next rebuild
        sum += v[k];
    }
    int tot;
    int ex = block_exclusive_scan(sum, &tot);
    int run = P.bead_off + a.CHKSUM[blockIdx.x] + ex;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        int c = c0 + threadIdx.x * SCAN_PER_THREAD + k;
        if (c < cend) a.CSTART[c] = run;
        run += v[k];
        if (c == cend - 1) a.CSTART[cend] = run;
    }
}

__global__ void __launch_bounds__(BS) k_scatter(PoolArgs a, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (s >= P.bead_off + P.n) return;
    int2 ko = *reinterpret_cast<const int2*>(a.KEYOFF + 2 * s);
    int ns = a.CSTART[P.cell_off + ko.x] + ko.y;
    *reinterpret_cast<int2*>(a.PAIRS + 2 * ns) = make_int2(a.META[2 * s], s);
}

// Move each bead to its new slot. The scatter left a cell's beads in atomic
// (arbitrary) order; each bead's final place inside its cell is its rank by
// original index, counted over the cell, so the layout (and with it the pair
// summation order) depends on the positions alone.
template <int DIM>
__global__ void __launch_bounds__(BS) k_gather(PoolArgs a, int with_t, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int ns = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (ns >= P.bead_off + P.n) return;
    // everything this kernel reads, loaded before its first store (a store
    // could alias the replica's parameters, and every later read of them
    // would wait on memory)
    const uint32_t flags = P.flags;
    const bool rot = (flags & GCR_F_ROTDIFF) != 0, srk4 = (flags & GCR_F_SRK4) != 0;
    const bool wa = a.on_x1 && (flags & (GCR_F_SLIDE | GCR_F_GROUP | GCR_F_MOTOR));
    const int2 pr = *reinterpret_cast<const int2*>(a.PAIRS + 2 * ns);
    const int so = pr.y;
    double x[DIM], t[DIM], xi[DIM], xir[DIM], x1[DIM], fl[DIM];
    load_vec<DIM>(a.X, so, x);
    if (with_t) load_vec<DIM>(a.T, so, t);
    const int2 meta = *reinterpret_cast<const int2*>(a.META + 2 * so);
    // the pending noise and the last predictor travel with the bead
    load_vec<DIM>(a.XI, so, xi);
    if (rot) load_vec<DIM>(a.XIR, so, xir);
    load_vec<DIM>(a.X1, so, x1);
    // mid-step: one force array is live and the other free -- SRK2's stage-0
    // forces (FT is written by the last stage), SRK4's running total (F0
    // holds the forces of a stage already applied); the live one travels
    // through the free one
    if (a.on_x1) load_vec<DIM>(srk4 ? a.FT : a.F0, so, fl);
    const int wav = wa ? a.WA[so] : 0;  // where_active so far, or a motor's bound bead
    const int key = a.KEYOFF[2 * so];
    const int t0 = a.CSTART[P.cell_off + key], t1 = a.CSTART[P.cell_off + key + 1];
    int rank = 0;
    for (int q = t0; q < t1; ++q) rank += a.PAIRS[2 * q] < pr.x;
    const int d = t0 + rank;
    store_vec<DIM>(a.Xn, d, x);
    if (with_t) store_vec<DIM>(a.Tn, d, t);
    *reinterpret_cast<int2*>(a.METAn + 2 * d) = meta;
    a.CELL[d] = key;
    a.SLOT[pr.x] = d;
    store_vec<DIM>(a.XIn, d, xi);
    if (rot) store_vec<DIM>(a.XIRn, d, xir);
    store_vec<DIM>(a.X1n, d, x1);
    if (a.on_x1) store_vec<DIM>(srk4 ? a.F0 : a.FT, d, fl);
    if (wa) a.WAn[d] = wav;
}

template <int DIM>
__global__ void __launch_bounds__(BS) k_links(PoolArgs a, int with_t, int j)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    if (!rebuild_runs(a, r, j)) return;
    const int ns = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (ns >= P.bead_off + P.n) return;
    // every read before the first store (as in k_gather)
    const uint32_t flags = P.flags;
    const bool rot = (flags & GCR_F_ROTDIFF) != 0, srk4 = (flags & GCR_F_SRK4) != 0;
    const bool wa = a.on_x1 && (flags & (GCR_F_SLIDE | GCR_F_GROUP | GCR_F_MOTOR));
    const int2 me = *reinterpret_cast<const int2*>(a.METAn + 2 * ns);
    double x[DIM], t[DIM], xi[DIM], xir[DIM], x1[DIM], fl[DIM];
    load_vec<DIM>(a.Xn, ns, x);
    if (with_t) load_vec<DIM>(a.Tn, ns, t);
    load_vec<DIM>(a.XIn, ns, xi);
    if (rot) load_vec<DIM>(a.XIRn, ns, xir);
    load_vec<DIM>(a.X1n, ns, x1);
    if (a.on_x1) load_vec<DIM>(srk4 ? a.F0 : a.FT, ns, fl);
    const int wav = wa ? a.WAn[ns] : 0;
    const int first = a.CHAIN_FIRST[me.y], len = a.CHAIN_LEN[me.y];
    const int kk = me.x - first;
    int4 lk;
    lk.x = kk >= 2 ? a.SLOT[me.x - 2] : -1;
    lk.y = kk >= 1 ? a.SLOT[me.x - 1] : -1;
    lk.z = kk + 1 < len ? a.SLOT[me.x + 1] : -1;
    lk.w = kk + 2 < len ? a.SLOT[me.x + 2] : -1;
    *reinterpret_cast<int4*>(a.LINK + 4 * ns) = lk;
    store_vec<DIM>(a.X, ns, x);
    if (!a.on_x1) store_vec<DIM>(a.XR, ns, x);
    if (with_t) store_vec<DIM>(a.T, ns, t);
    *reinterpret_cast<int2*>(a.META + 2 * ns) = me;
    store_vec<DIM>(a.XI, ns, xi);
    if (rot) store_vec<DIM>(a.XIR, ns, xir);
    store_vec<DIM>(a.X1, ns, x1);
    if (a.on_x1) {
        store_vec<DIM>(a.XR, ns, x1);
        store_vec<DIM>(srk4 ? a.FT : a.F0, ns, fl);
        if (wa) a.WA[ns] = wav;
    }
}

// ------------------------------------------------------------ batch control

// Start of a batch: remember where each replica's step count stands, clear
// the violation flag and snapshot the state for a possible replay.
template <int DIM>
__global__ void __launch_bounds__(BS) k_batch_begin(PoolArgs a, int with_t)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (S->frozen) return;
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (threadIdx.x == 0 && blockIdx.x == P.blk_off) {
        S->batch_base = S->steps;
        S->violation = 0;
        S->pred_far = 0;
        a.Ss[r] = *S;
    }
    if (s >= P.bead_off + P.n) return;
    // the first step of the batch needs its noise (after a begin or a replay
    // nothing else has drawn it)
    store_noise<DIM>(a, P, s, S->steps, S->step_global);
    double x[DIM];
    load_vec<DIM>(a.X, s, x);
    store_vec<DIM>(a.Xs, s, x);
    if (with_t) {
        double t[DIM];
        load_vec<DIM>(a.T, s, t);
        store_vec<DIM>(a.Ts, s, t);
    }
    *reinterpret_cast<int2*>(a.METAs + 2 * s) = *reinterpret_cast<const int2*>(a.META + 2 * s);
}

// Undo a batch for replicas whose flag `frozen` is 2 (host-marked for replay).
template <int DIM>
__global__ void __launch_bounds__(BS) k_restore(PoolArgs a, int with_t)
{
    const int r = a.blk_rep[blockIdx.x];
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (S->frozen != 2) return;
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (s < P.bead_off + P.n) {
        double x[DIM];
        load_vec<DIM>(a.Xs, s, x);
        store_vec<DIM>(a.X, s, x);
        if (with_t) {
            double t[DIM];
            load_vec<DIM>(a.Ts, s, t);
            store_vec<DIM>(a.T, s, t);
        }
        *reinterpret_cast<int2*>(a.META + 2 * s) = *reinterpret_cast<const int2*>(a.METAs + 2 * s);
    }
}

__global__ void k_restore_state(PoolArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.nrep) return;
    if (a.S[r].frozen == 2) a.S[r] = a.Ss[r];  // snapshot has frozen == 0
}

// ---------------------------------------------------------------- download

// Gather one replica's per-slot int32 array into original bead order.
__global__ void __launch_bounds__(BS) k_unpermute_i32(const int32_t* src, int32_t* dst, const int32_t* SLOT,
                                                      int orig0, int n)
{
    const int b = blockIdx.x * BS + threadIdx.x;
    if (b < n) dst[b] = src[SLOT[orig0 + b]];
}

// Gather one replica's per-slot array into original bead order.
__global__ void __launch_bounds__(BS) k_unpermute(const double* src, double* dst, const int32_t* SLOT,
                                                  int orig0, int n, int width)
{
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    int s = SLOT[orig0 + i];
    for (int c = 0; c < width; ++c) dst[(size_t)i * width + c] = src[(size_t)s * width + c];
}

// Scatter an original-order array of one replica into slots.
__global__ void __launch_bounds__(BS) k_permute_in(const double* src, double* dst, const int32_t* SLOT,
                                                   int orig0, int n, int width)
{
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    int s = SLOT[orig0 + i];
    for (int c = 0; c < width; ++c) dst[(size_t)s * width + c] = src[(size_t)i * width + c];
}

// Tangents from the bonds at X1 (tangent_list after a step holds the
// tangents computed during the corrector's force evaluation, SRK.py:228-233).
template <int DIM>
__global__ void __launch_bounds__(BS) k_tangents(PoolArgs a, int rsel)
{
    const int r = a.blk_rep[blockIdx.x];
    if (r != rsel) return;
    const RepParams& P = a.P[r];
    const int s = a.blk_slot0[blockIdx.x] + threadIdx.x;
    if (s >= P.bead_off + P.n) return;
    double xs[DIM], acc[DIM] = {}, tang[DIM];
    load_vec<DIM>(a.X1, s, xs);
    int2 me = *reinterpret_cast<const int2*>(a.META + 2 * s);
    ChainNb<DIM> cn;
    load_chain_nb<DIM>(a.X1, a.LINK, s, cn);
    bool set = false;
    bonded<DIM, double>(cn, xs, a.CHAIN_A[me.y], P, acc, tang, set);
    if (!set) load_vec<DIM>(a.T, s, tang);
    store_vec<DIM>(a.Tn, s, tang);
}

// ------------------------------------------------------- rigid-rod updates

// Makes this wavefront's LDS writes visible to its own later LDS reads of
// other lanes' entries (the LDS serves a wavefront's operations in order; the
// fence keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Rigid-rod projection of one stage's update (updates.py:347-499) for
// replicas with GCR_F_ROD. P.rod_lanes lanes per chain, as many chains per
// wavefront as fit (lane k of a chain's group holds bead k; chains longer
// than 64 beads in segments of 64). STAGE 0: a virtual update, after the
// force stage of SRK stage a.sub (x1 from x in stage 0 and from x1 in the
// inner SRK4 stages -- the position the forces were evaluated at, as
// SRK.py:364-371 passes it -- with that stage's forces in F0 and dt_prev
// times the stage's rescale factor). STAGE 1: the real update after
// k_finalize (x from x1, with the total force and dt_used times the force
// prefactor, 1/2 in SRK2). The update of a bead is dt/r F + sqrt(2 D dt) xi
// with that dt. The chain sums run over the beads in order, as the
// reference's loops do: the lanes put their terms in LDS and the chain's
// first lane adds them one by one, so each sum rounds as the sequential loop
// does. Within a stage a chain's lanes read and write that chain only, so x1
// may be both the input and the output. Each entry of the chain-block tables
// (BS chains) is covered by the first of its 64 blocks of four wavefronts.
template <int DIM, int STAGE>
__global__ void __launch_bounds__(256) k_rod(PoolArgs a, int j)
{
    __shared__ double sh[4][64][2 * DIM];  // per lane: (u, sep), then the torque term
    __shared__ double res[4][ROD_MAX_CPW][3 * DIM + 1];  // per chain: mean, com, r_hat, |r|
    const int e = blockIdx.x >> 6;
    const int r = a.chb_rep[e];
    const RepParams& P = a.P[r];
    const RepState* S = a.S + r;
    if (STAGE == 0) {
        if (!step_runs(S, j)) return;
        if (a.sub > 0 && !(P.flags & GCR_F_SRK4)) return;
    } else if (S->frozen || S->steps != S->batch_base + j + 1) {
        return;
    }
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int G = P.rod_lanes, cpw = 64 / G;
    const int grp = lane / G, k = lane - grp * G, lead = grp * G;
    const int c0 = a.chb_chain0[e] + ((blockIdx.x & 63) * 4 + wv) * cpw;
    const int cend = min(a.chb_chain0[e] + BS, P.chain_off + P.n_chains);
    if (c0 >= cend) return;  // the whole wavefront
    const int c = c0 + grp;
    const bool on = grp < cpw && c < cend;  // the lane belongs to a chain
    // the projection starts from the positions the stage's forces were
    // evaluated at (SRK.py:364-371): x in SRK stage 0, x1 after it
    const double* prev = (STAGE == 0 && a.sub == 0) ? a.X : a.X1;
    const double* F = STAGE == 0 ? a.F0 : a.FT;
    double* out = STAGE == 0 ? a.X1 : a.X;
    const double dt = STAGE == 0 ? S->dt * virtual_rescale(P, a.sub) * 1.0
                                 : S->dt_used * force_prefactor(P);
    const double dt_over_r = dt / P.radius;
    const double sq = sqrt(2.0 * P.D * dt);
    const double L = P.L, half = 0.5 * L, fac = P.rod;
    const int first = on ? a.CHAIN_FIRST[c] : 0, cl = on ? a.CHAIN_LEN[c] : 0;
    const double ca = on ? a.CHAIN_A[c] : 0.0;
    const double half_clm1 = (cl - 1) / 2.0;
    // bead kk's update u and position x (kk >= cl: zeros, slot -1)
    auto bead = [&](int kk, double* u, double* x, int& sl) {
        sl = -1;
#pragma unroll
        for (int i = 0; i < DIM; ++i) u[i] = x[i] = 0.0;
        if (kk >= cl) return;
        sl = a.SLOT[first + kk];
        double f[DIM], xi[DIM];
        load_vec<DIM>(F, sl, f);
        load_vec<DIM>(a.XI, sl, xi);
        load_vec<DIM>(prev, sl, x);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            u[i] = dt_over_r * f[i];
            if (sq != 0.0) u[i] += sq * xi[i];
        }
    };
    // the lane's bead of the first segment stays in registers
    double u0[DIM], x0[DIM];
    int sl0;
    bead(k, u0, x0, sl0);
    auto seg_bead = [&](int base, double* u, double* x, int& sl) {
        if (base == 0) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                u[i] = u0[i];
                x[i] = x0[i];
            }
            sl = sl0;
        } else {
            bead(base + k, u, x, sl);
        }
    };
    double p0[DIM];
    if (on) load_vec<DIM>(prev, a.SLOT[first], p0);
    // pass 1: mean update, centre of mass and head-to-tail direction, summed
    // by the chain's first lane
    double mean[DIM], com[DIM], rh[DIM], rmag = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) mean[i] = com[i] = rh[i] = 0.0;
    for (int base = 0; base < cl; base += G) {
        double u[DIM], x[DIM];
        int sl;
        seg_bead(base, u, x, sl);
        if (sl >= 0) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                double d = x[i] - p0[i];
                if (d < -half)
                    d += L;
                else if (d > half)
                    d -= L;
                sh[wv][lane][i] = u[i];
                sh[wv][lane][DIM + i] = d;
            }
        }
        wave_lds_sync();
        if (k == 0) {
            const int nk = min(G, cl - base);
#pragma unroll 2
            for (int q = 0; q < nk; ++q)
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    mean[i] += sh[wv][lead + q][i];
                    com[i] += sh[wv][lead + q][DIM + i];
                }
            if (base + G >= cl)
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    rh[i] = sh[wv][lead + cl - 1 - base][DIM + i];
                    rmag += rh[i] * rh[i];
                }
        }
        wave_lds_sync();
    }
    if (k == 0 && on) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            mean[i] /= cl;
            com[i] = pymod(p0[i] + com[i] / cl, L);
        }
        if (rmag > 0) {
            rmag = sqrt(rmag);
#pragma unroll
            for (int i = 0; i < DIM; ++i) rh[i] /= rmag;
        }
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            res[wv][grp][i] = mean[i];
            res[wv][grp][DIM + i] = com[i];
            res[wv][grp][2 * DIM + i] = rh[i];
        }
        res[wv][grp][3 * DIM] = rmag;
    }
    wave_lds_sync();
    if (on) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            mean[i] = res[wv][grp][i];
            com[i] = res[wv][grp][DIM + i];
            rh[i] = res[wv][grp][2 * DIM + i];
        }
        rmag = res[wv][grp][3 * DIM];
    }
    // pass 2: the torque term, summed by the chain's first lane
    double tq[3] = {0.0, 0.0, 0.0};
    for (int base = 0; base < cl; base += G) {
        double u[DIM], x[DIM];
        int sl;
        seg_bead(base, u, x, sl);
        if (sl >= 0) {
            if (DIM == 2) {
                const double d = ca * (base + k - half_clm1);
#pragma unroll
                for (int i = 0; i < 2; ++i) sh[wv][lane][i] = (2 * i - 1) * rh[1 - i] * u[i] * d;
            } else {
                sh[wv][lane][0] = rh[1] * u[2] - rh[2] * u[1];
                sh[wv][lane][1] = rh[2] * u[0] - rh[0] * u[2];
                sh[wv][lane][2] = rh[0] * u[1] - rh[1] * u[0];
            }
        }
        wave_lds_sync();
        if (k == 0) {
            const int nk = min(G, cl - base);
#pragma unroll 2
            for (int q = 0; q < nk; ++q)
#pragma unroll
                for (int i = 0; i < (DIM == 2 ? 2 : 3); ++i) tq[i] += sh[wv][lead + q][i];
        }
        wave_lds_sync();
    }
    double rot[3] = {0.0, 0.0, 0.0};  // 3D: (|r| / I) (torque x r_hat), the same for every bead
    if (k == 0 && on) {
        double moi = 0.0;
#pragma unroll 1
        for (int q = 0; q < cl; ++q) {
            const double d = ca * (q - half_clm1);
#pragma unroll
            for (int i = 0; i < DIM; ++i) moi += d * d;
        }
        if (DIM == 2) {
            if (moi != 0)
                for (int i = 0; i < 2; ++i) tq[i] /= moi;
        } else {
            for (int i = 0; i < 3; ++i) tq[i] *= rmag;
            const double w = rmag / moi;
            rot[0] = w * (tq[1] * rh[2] - tq[2] * rh[1]);
            rot[1] = w * (tq[2] * rh[0] - tq[0] * rh[2]);
            rot[2] = w * (tq[0] * rh[1] - tq[1] * rh[0]);
        }
        // 2D: the scaled torque; 3D: the rotation term
#pragma unroll
        for (int i = 0; i < DIM; ++i) sh[wv][lane][i] = DIM == 2 ? tq[i] : rot[i];
    }
    wave_lds_sync();
    if (on)
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            if (DIM == 2)
                tq[i] = sh[wv][lead][i];
            else
                rot[i] = sh[wv][lead][i];
        }
    // pass 3: blend with the rigid-rod update and write, one bead per lane
    bool viol = false;
    for (int base = 0; base < cl; base += G) {
        double u[DIM], x[DIM], y[DIM];
        int sl;
        seg_bead(base, u, x, sl);
        if (sl < 0) continue;
        const double d = ca * (base + k - half_clm1);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            double ool = x[i] - (com[i] + rh[i] * d);
            if (ool < -half)
                ool += L;
            else if (ool >= half)
                ool -= L;
            double v = u[i] * (1.0 - fac);
            if (DIM == 2)
                v += fac * (-ool + mean[i] + (tq[i] * d * (2 * i - 1) * rh[1 - i]));
            else
                v += fac * (-ool + mean[i] + rot[i]);
            y[i] = pymod(x[i] + v, L);
        }
        store_vec<DIM>(out, sl, y);
        viol |= moved_too_far<DIM>(a, P, sl, y);
    }
    if (__any(viol) && lane == 0) {
        if (STAGE == 0)
            flag_predictor(a, a.S + r, a.sub);
        else
            atomicOr(a.safe ? &a.S[r].sort_due : &a.S[r].violation, 1);
    }
}

// ------------------------------------------------------------ observables

// Order parameters of one replica from its chains' mean directions
// (output_methods.calc_avg_phi_chain, output_methods.py:492-510, n-fold
// symmetry 1): phi_c = atan2(sum_b t_y, sum_b t_x) over the chain's beads in
// bead order. Each block sums, over its chains, cos phi, sin phi,
// cos 2phi = 2cos^2 phi - 1 and sin 2phi = 2 cos phi sin phi (the Q-tensor
// terms of analysis.py:294-313) into OBS[4 * block + k]; k_order_final adds
// the blocks in order, so the result does not depend on scheduling.
// `tang` holds the tangents by slot (T, or the bond tangents at x1).
__global__ void __launch_bounds__(BS) k_chain_order(PoolArgs a, int r, const double* tang, int dim, double* OBS)
{
    const RepParams& P = a.P[r];
    const int c = blockIdx.x * BS + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < P.n_chains) {
        const int chain = P.chain_off + c;
        const int first = a.CHAIN_FIRST[chain], len = a.CHAIN_LEN[chain];
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < len; ++k) {
            const int s = a.SLOT[first + k];
            sx += tang[(size_t)s * dim];
            sy += tang[(size_t)s * dim + 1];
        }
        const double phi = atan2(sy, sx);
        double sn, cs;
        sincos(phi, &sn, &cs);
        v[0] = cs;
        v[1] = sn;
        v[2] = 2.0 * cs * cs - 1.0;
        v[3] = 2.0 * cs * sn;
    }
    __shared__ double red[4][BS];
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) OBS[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One sample of the collective observables of replica r, per chain: the
// order terms of k_chain_order, and the chain centre of mass -- the mean of
// the chain's beads unwrapped bond by bond from its first bead (minimum
// image), folded into the box -- whose displacement since the last sample
// (minimum image: a centre moves less than half a box between samples) is
// added to its unwrapped displacement U since the origin; reset makes the
// current centres the origin. Per block, the sums (cos phi, sin phi, Q11,
// Q12, |U|^2) go to PART[5 * block + k] in a fixed order.
template <int DIM>
__global__ void __launch_bounds__(BS) k_chain_sample(PoolArgs a, int r, const double* tang, int reset,
                                                     double* COMW, double* COMU, double* PART)
{
    const RepParams& P = a.P[r];
    const int c = blockIdx.x * BS + threadIdx.x;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < P.n_chains) {
        const int chain = P.chain_off + c;
        const int first = a.CHAIN_FIRST[chain], len = a.CHAIN_LEN[chain];
        double sx = 0.0, sy = 0.0;
        double prev[DIM], acc[DIM];
        const int s0 = a.SLOT[first];
        load_vec<DIM>(a.X, s0, prev);
#pragma unroll
        for (int i = 0; i < DIM; ++i) acc[i] = prev[i];
        for (int k = 0; k < len; ++k) {
            const int s = a.SLOT[first + k];
            sx += tang[(size_t)s * DIM];
            sy += tang[(size_t)s * DIM + 1];
            if (k > 0) {
                double x[DIM];
                load_vec<DIM>(a.X, s, x);
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    prev[i] += wrap1(x[i] - prev[i], P.L, P.halfL);
                    acc[i] += prev[i];
                }
            }
        }
        const double phi = atan2(sy, sx);
        double sn, cs;
        sincos(phi, &sn, &cs);
        v[0] = cs;
        v[1] = sn;
        v[2] = 2.0 * cs * cs - 1.0;
        v[3] = 2.0 * cs * sn;
        double* W = COMW + (size_t)chain * 3;
        double* U = COMU + (size_t)chain * 3;
        double u2 = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const double com = pymod(acc[i] / (double)len, P.L);
            double u = 0.0;
            if (!reset) u = U[i] + wrap1(com - W[i], P.L, P.halfL);
            W[i] = com;
            U[i] = u;
            u2 += u * u;
        }
        v[4] = u2;
    }
    __shared__ double red[5][BS];
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) PART[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The sample row of replica r from the block sums (added in block order):
// [time since the replica was added, polar order |sum e^{i phi}| / C,
// nematic order |(Q11, Q12)| / C, director arctan((|Q| - Q11) / Q12) (0 when
// Q12 = 0), mean |U|^2 over the chains, steps since the replica was added]
// (analysis.py:294-328). One thread.
__global__ void k_sample_final(PoolArgs a, int r, const double* PART, int nblk, double* row)
{
    if (threadIdx.x != 0) return;
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nblk; ++b)
        for (int k = 0; k < 5; ++k) t[k] += PART[5 * b + k];
    const double C = (double)a.P[r].n_chains;
    const double qf = hypot(t[2], t[3]);
    row[0] = a.S[r].t_total;
    row[1] = hypot(t[0], t[1]) / C;
    row[2] = qf / C;
    row[3] = t[3] != 0.0 ? atan((qf - t[2]) / t[3]) : 0.0;
    row[4] = t[4] / C;
    row[5] = (double)a.S[r].step_global;
}

// Sums of the per-block terms, in block order, into OBS[4 * nblk + k].
__global__ void k_order_final(double* OBS, int nblk)
{
    const int k = threadIdx.x;
    if (k >= 4) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += OBS[4 * b + k];
    OBS[4 * nblk + k] = acc;
}

}  // namespace gcr
