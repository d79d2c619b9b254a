/*
 * gcr_oracle.h -- CPU FP64 restatement of the reference Brownian-dynamics hot
 * path (WCA/LJ pair forces, Hookean bond + bending forces, chiral
 * self-propulsion, SRK2 step with adaptive dt, rotational diffusion).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the
 * smoke() entry point and bench.py's CPU-baseline leg may load it. The
 * product path (libgcr.so) never links or calls it.
 *
 * Each function cites the reference file:line it restates; tests/test_oracle.py
 * pins each one against closed forms, energy gradients and invariants.
 */
#ifndef GCR_ORACLE_H
#define GCR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics of one population of chains (or ABPs when every chain has one
 * bead), as SRK_jit sees it (SRK.py:46-97, updates.py:144-162). */
typedef struct orc_sys {
    int32_t dim;              /* 2 or 3 */
    int32_t n;                /* beads */
    int32_t n_nearby;         /* n_nearby_beads_dont_repel (pop_consts[4]) */
    int32_t has_springs;      /* "springs" in the force list, longest chain > 1 */
    int32_t has_wca;          /* "WCA" in the force list */
    int32_t has_propel;       /* "self-propulsion" and b_calc_active_forces */
    double  L;                /* Lx */
    double  D, Dr, radius;    /* pop_consts[0..2] */
    double  wca[4];           /* make_WCA_params output (populations.py:48-58) */
    double  spring[6];        /* spring_params (populations.py:345-353) */
    double  v0;               /* active_params[0] (populations.py:355-358) */
    double  rod;              /* rigid_rod_projection_factor (pop_consts[3]) */
    int32_t order;            /* SRK_order: 4, else 2 (config.SRK_order) */
    int32_t walls;            /* config.walls as bits: x 1, y 2, z 4 */
    double  wall_zone;        /* forbidden_zone_width (forces.py:417) */
    double  wall_cutoff;      /* wall_params[3]; <= 0: wca[3] */
    double  sliding_maxdist;  /* active sliding (active_sliding_and_WCA): partner range, 0: off */
    double  sliding_v0;       /* its push per antiparallel partner */
    const int32_t* chunk;     /* thread chunk of each bead (NULL: one chunk) */
    int32_t* where_active;    /* per bead: -1 at each step start, +1 per partner and stage */
} orc_sys;

/* Python float modulo (CPython float_rem semantics, used by `tmp %= Lx`
 * at SRK.py:387): result has the sign of b, can equal b for tiny negative a. */
double orc_pymod(double a, double b);

/* populations.py:48-58 */
void orc_wca_params(double repel_d_max, double eps, double shift, double cutoff_factor,
                    double out[4]);

/* forces.py:13-15: force coefficient; F_on_b1 = coeff * (x_b2 - x_b1). */
double orc_wca_coeff(double dist_sq, const double p[4]);

/* forces.py:44-54: out = a - b, each component corrected once into [-L/2, L/2). */
double orc_inbounds_diff(int dim, const double* a, const double* b, double half_L, double L,
                         double* out);

/* forces.py:285-402 (Hookean, or FENE when spring[3] != 0). Accumulates
 * into forces[n*dim]; writes tangent[n*dim] when spring[4] != 0. chain_id is
 * int64 as in populations.py:606. */
void orc_chain_forces(int dim, int n, const double* pos, const int64_t* chain_id, double L,
                      const double spring[6], const double* rest_lengths,
                      double* tangent, double* forces);

/* forces.py:57-144 with the neighbour set of updates.py:586-669: every pair
 * (b1, b2) with 0 < |d|^2 <= rc^2, skipping same-chain pairs with
 * |b1-b2| <= n_nearby (all same-chain pairs when n_nearby < 0). Brute force
 * O(n^2), full list (each pair evaluated from both sides, the cross-chunk
 * branch forces.py:141-143). Accumulates into forces. */
void orc_pair_forces_brute(int dim, int n, const double* pos, const int64_t* chain_id, double L,
                           const double wca[4], int n_nearby, double* forces);

/* forces.py:206-223 with a per-chain chirality angle (reduces to the
 * reference's scalar v_tan_angle when every entry is equal). */
void orc_self_propel(int dim, int n, const int64_t* chain_id, double v0,
                     const double* v_tan_angle_per_chain, const double* tangent, double* forces);

/* forces.py:57-144 with b_calc_active_sliding (forces.py:79-83, 114-120): for
 * every listed partner b2 of b1 (the same-chain rule applies) closer than
 * maxdist whose tangent points against b1's (t1 . t2 < 0), b1 gets v0 t1 and
 * where_active[b1] one more. A partner in b1's own thread chunk counts only
 * when its index is lower (the reference skips b2 >= b1 there, a pair it
 * handles once with Newton's third law for the pair force but not for the
 * sliding push). Accumulates into forces. */
void orc_sliding(int dim, int n, const double* pos, const int64_t* chain_id, double L,
                 const double* tangent, int n_nearby, double maxdist, double v0,
                 const int32_t* chunk, double* forces, int32_t* where_active);

/* forces.py:405-435: WCA walls normal to the axes in `walls`, at `zone`
 * inside the faces, from the positions pos; accumulates into forces. Every
 * bead is visited (the reference skips beads whose last grid cell is more
 * than 2 grid_window cells from a face, none of which is within range). */
void orc_walls(int dim, int n, const double* pos, double L, const double wca[4], int walls,
               double zone, double* forces);

/* One force evaluation in the reference order (SRK.py:228-275):
 * springs (+tangents) -> pair -> self-propulsion -> walls, accumulated into
 * forces; the walls take wall_pos (SRK.py:270 passes pos in every stage),
 * NULL meaning pos. */
void orc_force_eval(const orc_sys* s, const double* pos, const int64_t* chain_id,
                    const double* rest_lengths, const double* v_tan_angle_per_chain,
                    double* tangent, double* forces, int pair_mode, void* verlet,
                    const double* wall_pos);

/* Noise: Philox4x32-10 (Salmon et al., SC'11) and the build's bead-keyed
 * normal generator shared with the GPU path. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* xi[2*dim] for bead `orig` at global step `step`: translational normals in
 * xi[0:dim], rotational in xi[dim:2dim]. */
void orc_bead_noise(int dim, uint64_t seed, uint32_t orig, int64_t step, double* xi);

/* The SRK loop of SRK_jit (SRK.py:128-423) for one population with a single
 * thread, SRK_order = s->order (2, or 4: SRK.py:186-215, 297-321, 355-358).
 * time_vals as built at updates.py:144-146:
 * [dt, dtmax, max_dist_per_step, RK_cheat_factor, max_dt_usable,
 *  steps_until_update, n_steps_update, max_steps].
 * status_out = [dt, elapsed (dt_used sum), steps, max_forces_strength].
 * xi: NULL -> generated by orc_bead_noise(seed, b, step0 + k); otherwise
 * xi[k*n*2*dim + b*2*dim + c] for local step k (noise injection).
 * pair_mode: 0 = brute force, 1 = Verlet list (cell grid, rebuilt every
 * n_steps_update steps, half list with Newton-3 inside each OpenMP chunk, as
 * the reference's threads do).
 * Returns steps done. */
/* updates.py:347-499 (project_updates_for_rigid_rod): the update of every
 * chain blended with that of a rigid rod, rtn = (prev + update) % L, for the
 * chains of [0, n). dt is the update's time step (dt_eff before the division
 * by the radius), the noise amplitude sqrt(2 D dt). */
void orc_rod_project(int dim, int n, double* rtn, const double* prev, const double* F,
                     const double* xi, int nd, const int64_t* chain_id, const double* rest_lengths,
                     double factor, double L, double dt, double D, double radius);

int64_t orc_srk2_run(const orc_sys* s, double* pos, double* pos1, double* tangent,
                     double* forces, const int64_t* chain_id, const double* rest_lengths,
                     const double* v_tan_angle_per_chain, const double time_vals[8],
                     double status_out[4], const double* xi, uint64_t seed, int64_t step0,
                     int pair_mode, int n_threads, double grid_factor);

/* forces.py:57-144 between two populations (a thread with same_pop False):
 * every pair (a of A, b of B) with 0 < |x_b - x_a|^2 <= rc^2, no exclusions;
 * the force on a is coeff * (x_b - x_a), WCA_shift as in the pair law.
 * Accumulates into FA. */
void orc_cross_pair_forces(int dim, int nA, const double* posA, int nB, const double* posB,
                           double L, const double p[4], double* FA);

/* One population of a multi-population run. */
typedef struct orc_pop {
    orc_sys sys;              /* its physics; sys.wca is its self interaction */
    int32_t fixed;            /* pop.fixed: never moves and only pushes the others */
    int32_t pad;
    double* pos;
    double* pos1;
    double* tangent;
    double* forces;
    const int64_t* chain_id;
    const double* rest_lengths;
    const double* vtan;
    const double* xi;         /* [steps][n][2 dim] or NULL: orc_bead_noise(seed, ...) */
    uint64_t seed;
    /* a motor population (populations.py:617-660): its motors push beads of
     * the populations whose cross entry names it with kind 2 */
    int32_t motor;
    int32_t motor_still;      /* b_motors_dont_diffuse_while_pushing */
    double motor_max_dist_sq, motor_imparts, k_on, k_off;
    double motor_reach;       /* beads farther than this are not the motor's candidates */
    int32_t* motor_bead;      /* per motor: the bead it pushes, -1: free (where_active) */
} orc_pop;

/* The uniform draw of the motor rules: Philox4x32-10 keyed by the motor
 * population's seed, counter (motor, bead, step lo, step hi << 8 | 0x80 |
 * stage << 1 | kind), kind 0 the unbinding draw, 1 the binding draw; 53 bits
 * in [0, 1). Shared with the GPU path. */
double orc_motor_uniform(uint64_t seed, uint32_t motor, uint32_t bead, int64_t step, int stage,
                         int kind);

/* The SRK2 loop over several populations (updates.py:132-267 with the
 * same_pop False threads of SRK.py): each moving population P evaluates its
 * own forces (orc_force_eval), then adds the forces from every other
 * population Q in the order (Q - P) mod npop = 1, 2, ... (SRK.py:287-293):
 * cross[(P * npop + Q) * 5] == 1, the WCA law cross[... + 1 .. 4]; == 2, the
 * motor pushes of motor population Q (forces.py:248-283, below)
 * (interpop_forces, updates.py:274-312), Q's positions being pos1 in the
 * second stage unless Q is fixed; m is the largest sqrt(2 D_P) + 1/2 max|F_P|
 * / r_P over the moving populations, one dt for all; each population updates
 * with its own D, radius and noise. pair_mode 0: every pair of beads;
 * pair_mode 1 (the CPU baseline): each population's own pairs over its
 * Verlet list (grid_factor, rebuilt on time_vals' cadence) on n_threads
 * threads, the other populations through a cell grid of their beads.
 * Returns steps done. */
int64_t orc_multi_srk2_run(int npop, orc_pop* pops, const double* cross, const double time_vals[8],
                           double status_out[4], int64_t step0, int pair_mode, int n_threads,
                           double grid_factor);

/* forces.py:248-283 in one stage: for every bead b (index order) and every
 * motor m within motor_reach of it (index order): a bound motor stays bound
 * unless its draw is <= k_off, in which case it comes free and b's
 * where_active drops by one (whichever bead it pushed); a free motor binds to
 * b when its draw is < k_on and b lies within the motor range, adding
 * imparts * t_b to b's force and one to b's where_active. Draws from
 * orc_motor_uniform(seed of M, motor, bead, step, stage, kind). Motors are
 * independent of each other, so each motor's history is its beads in order. */
void orc_motor_push(int dim, int nb, const double* posb, const double* tb, int nm,
                    const double* posm, double L, const orc_pop* M, int64_t step, int stage,
                    double* F, int32_t* where_active_b);

#ifdef __cplusplus
}
#endif
#endif
