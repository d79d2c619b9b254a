/*
 * gcr.h -- C-ABI of the MI355X Brownian-dynamics hot path (libgcr.so).
 *
 * Plain C types only: caller-owned host buffers, opaque handles, int return
 * codes (0 = ok, < 0 = error, message from gcr_last_error()). All arrays are
 * C-contiguous float64 in the reference's layout: pos[n][dim],
 * tangent[n][dim], noise[steps][n][2*dim]; chain_id is int64 per bead as
 * populations.py:606 builds it (chains are runs of equal chain_id).
 *
 * Two layers:
 *  - the single-system API (gcr_create / gcr_upload / gcr_run /
 *    gcr_download / gcr_destroy) is what updates.SRK2_force_calc would bind
 *    in place of its SRK_jit call;
 *  - the pool API runs many replicas (independent simulations of one
 *    dimensionality) in the same launches; the ensemble and the serving
 *    loop are built on it.
 */
#ifndef GCR_H
#define GCR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCR_OK 0
#define GCR_ERR_ARG (-1)      /* bad argument (size, pointer, parameter) */
#define GCR_ERR_HIP (-2)      /* HIP runtime error, or no device */
#define GCR_ERR_CAPACITY (-3) /* pool capacity exhausted */
#define GCR_ERR_STATE (-4)    /* call out of order (e.g. run before begin) */

#define GCR_MAX_POPS 16 /* populations in one group (gcr_pool_group) */

#define GCR_FP64 0  /* forces and accumulation in FP64 */
#define GCR_MIXED 1 /* pair forces in FP32, accumulation and state in FP64 */

typedef struct gcr_pool gcr_pool;

/* Physics of one population, as SRK_jit receives it.
 * Fields map to: pop_consts = [D, Dr, radius, rigid_rod, n_nearby, Lmax]
 * (updates.py:148-162), force_params rows for "WCA" (populations.py:48-58),
 * "springs" (populations.py:345-353) and "self-propulsion"
 * (populations.py:355-358), and config.Lx. */
typedef struct gcr_replica_desc {
    int32_t n_beads;
    int32_t n_chains;
    int32_t n_nearby;    /* n_nearby_beads_dont_repel */
    int32_t has_springs; /* "springs" in the force list */
    int32_t has_wca;     /* "WCA" in the force list */
    int32_t has_propel;  /* "self-propulsion" and status.b_calc_active_forces */
    double L;            /* periodic box edge, config.Lx */
    double D, Dr, radius;
    double wca[4];       /* [12 eps r^6, 12 eps r^12, WCA_shift, LJ cutoff] */
    double spring[6];    /* [k, kappa, spontaneous angle, FENE, get_orient, FENE tol] */
    double v0;           /* active_params[0], used when v0_per_chain is NULL */
    double skin;         /* cell-list skin; <= 0 picks 0.3 * cutoff */
    uint64_t seed;       /* key of the in-kernel Philox noise */
    double rod;          /* rigid_rod_projection_factor (pop_consts[3]); used when the
                            longest chain has more than two beads (SRK.py:364) */
    int32_t srk_order;   /* config.SRK_order: 4, or 2 (also for 0) */
    int32_t walls;       /* config.walls as bits (x 1, y 2, z 4): WCA walls replace the
                            periodic faces (forces.py:405-435) */
    double wall_zone;    /* forbidden_zone_width of forces.py:417, grid_spacing *
                            grid_window / 2: the walls sit this far inside the faces */
    double wall_cutoff;  /* wall_params[3], the walls' range parameter; <= 0: wca[3] (they
                            differ while growing, when only the pair law is made purely
                            repulsive, updates.py:304-310) */
    double sliding_maxdist; /* active sliding (the "active_sliding_and_WCA" force,
                            forces.py:79-83, 114-120): partner range, 0: off */
    double sliding_v0;   /* push v0 t per antiparallel partner */
    int32_t motor;       /* a motor population (populations.py:617-660): its beads are motors
                            that bind to and push the beads of the population the group's
                            cross table names (kind 2, gcr_pool_group) */
    int32_t motor_still; /* b_motors_dont_diffuse_while_pushing */
    double motor_max_dist; /* motor_max_dist: range of a binding */
    double motor_imparts;  /* motor_imparts_force: push imparts * t of a binding */
    double k_on, k_off;  /* binding and unbinding probabilities per visit */
    double motor_reach;  /* beads within this of a motor are visited by it (its neighbour
                            list); <= 0: motor_max_dist */
} gcr_replica_desc;

/* ---------------------------------------------------------------- errors */
const char* gcr_last_error(void);
/* Device name, CU count and HBM bytes of the current device. */
int gcr_device_info(char* name, int32_t name_len, int32_t* n_cu, int64_t* hbm_bytes);
/* No counterpart in the reference: the cost of one grid-wide barrier on this
 * device, measured with a cooperative launch of blocks_per_cu workgroups of
 * 256 threads per CU (fewer if they cannot all be resident) running `rounds`
 * barriers back to back: out = [microseconds per barrier, workgroups]. The
 * price a persistent step kernel would pay at each boundary between its
 * stages, against the batch graphs' back-to-back launches. */
int gcr_grid_barrier_us(int32_t blocks_per_cu, int32_t rounds, double out[2]);

/* ------------------------------------------------------ single system API */

/* Replaces the argument marshalling of updates.SRK2_force_calc
 * (updates.py:132-247) and the state a SRK_jit call reads (SRK.py:14-29). */
int gcr_create(const gcr_replica_desc* desc, int32_t dim, int32_t precision, const double* pos,
               const double* tangent, const int64_t* chain_id, const double* rest_lengths,
               const double* v_tan_angle, gcr_pool** out);

/* Re-upload positions / propulsion directions (NULL leaves one unchanged),
 * e.g. after the host edited pop.pos between calls. */
int gcr_upload(gcr_pool* h, const double* pos, const double* tangent);

/* Replaces SRK.SRK_jit (SRK.py:13-484) for SRK_order = 2: runs steps until
 * time_vals[7] steps (max_steps) or time_vals[4] elapsed time (max_dt_usable)
 * is reached, as the loop at SRK.py:128-130. time_vals is the array of
 * updates.py:144-146. status_out = shared_arr[0] of updates.py:249-250:
 * [dt, elapsed (dt_used sum), steps, max_forces_strength, 0, flag].
 * xi: NULL -> in-kernel Philox noise; else xi[k][n][2*dim] for the k-th step
 * of this call (noise injection, k < xi_steps). */
int gcr_run(gcr_pool* h, const double time_vals[8], double status_out[6], const double* xi,
            int64_t xi_steps);

/* Copies back pos, pos1 (SRK2 predictor of the last step), tangent_list and
 * total forces (F0 + F1 of the last step), each [n][dim]; NULL skips one. */
int gcr_download(gcr_pool* h, double* pos, double* pos1, double* tangent, double* forces);

/* One force evaluation (springs, pair, propulsion; forces.py:57-402) at the
 * current positions, without integrating: forces[n][dim], tangent[n][dim]. */
int gcr_eval_forces(gcr_pool* h, double* forces, double* tangent);

int gcr_destroy(gcr_pool* h);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc):
 * the force evaluation of gcr_create + gcr_eval_forces without any
 * allocation. pos, tangent (NULL: zero directions), chain_id, rest_lengths,
 * v_tan_angle (NULL: zeros), forces and tangent_out (NULL: skipped) are
 * DEVICE pointers in the layout of gcr_create / gcr_eval_forces. The cell
 * list and the bead staging live in one fixed device array of the library
 * (GCR_NOALLOC_SCRATCH_BYTES per device, held until the process ends): at
 * most GCR_NOALLOC_MAX_BEADS beads and GCR_NOALLOC_MAX_CELLS cells, larger
 * systems return GCR_ERR_CAPACITY. The work is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and not waited for. Input that
 * gcr_create would reject on its contents (a position outside [0, L], chain
 * ids that are not contiguous runs of 0..n_chains-1) gives NaN forces. Every
 * call on a device shares that device's array: calls that overlap on one
 * device race. */
#define GCR_NOALLOC_MAX_BEADS 32768
#define GCR_NOALLOC_MAX_CELLS 131072
#define GCR_NOALLOC_SCRATCH_BYTES (16 << 20)
int gcr_eval_forces_noalloc(const gcr_replica_desc* desc, int32_t dim, int32_t precision,
                            const double* pos, const double* tangent, const int64_t* chain_id,
                            const double* rest_lengths, const double* v_tan_angle, double* forces,
                            double* tangent_out, void* stream);

/* -------------------------------------------------------------- pool API */

/* A pool of replicas of one dimensionality; capacities bound HBM use. */
int gcr_pool_create(int32_t dim, int32_t precision, int64_t bead_capacity, int64_t cell_capacity,
                    int64_t chain_capacity, int32_t replica_capacity, gcr_pool** out);
/* Steps per graph-captured batch and cell-list rebuild period (steps);
 * steps_per_batch is rounded up to a multiple of rebuild_every. */
int gcr_pool_set_schedule(gcr_pool* p, int32_t steps_per_batch, int32_t rebuild_every);
/* HBM bytes a replica of n beads, n_chains chains and ncell cells takes. */
int64_t gcr_replica_bytes(int32_t dim, int64_t n_beads, int64_t n_chains, int64_t n_cells);
/* Cells a replica would use (needed to size cell_capacity). */
int64_t gcr_replica_cells(int32_t dim, const gcr_replica_desc* desc);

int gcr_replica_add(gcr_pool* p, const gcr_replica_desc* desc, const double* pos,
                    const double* tangent, const int64_t* chain_id, const double* rest_lengths,
                    const double* v_tan_angle, const double* v0_per_chain, int32_t* rid);
int gcr_replica_remove(gcr_pool* p, int32_t rid);
int gcr_replica_upload(gcr_pool* p, int32_t rid, const double* pos, const double* tangent);
/* Changes a replica's physics in place between calls (forces, constants,
 * rigid-rod factor, rest lengths when rest_lengths is not NULL, speeds:
 * v0_per_chain or desc->v0): the parameter changes control.py's
 * early_time_procedures makes between chunks (control.py:153-219). Bead and
 * chain counts, box and noise key stay; the cutoff must fit the cell grid
 * made at gcr_replica_add (desc->skin there sets it). */
int gcr_replica_set_physics(gcr_pool* p, int32_t rid, const gcr_replica_desc* desc,
                            const double* rest_lengths, const double* v0_per_chain);
/* Makes replicas rids[0..n) of the pool (same box, n <= GCR_MAX_POPS,
 * SRK_order 2, no rigid rod) the populations of one system, in the order of
 * config.pop_name_list: the pops list SRK2_force_calc receives
 * (updates.py:132-267). cross[(a * n + b) * 5] == 1: population b acts on
 * population a with the WCA/LJ law cross[(a * n + b) * 5 + 1 .. 4] = {12 eps
 * r^6, 12 eps r^12, shift, cutoff}; == 2: the motors of motor population b
 * push the beads of a ("motor_push", forces.py:248-283; one pushed population
 * per motor population); as interpop_forces(pop_a, pop_b) gives them
 * (updates.py:274-312: interactions() averages the two WCA vectors,
 * populations.py:226-232; special_interactions override); each replica's own
 * desc->wca is its self law. fixed[a] nonzero (NULL: none): population a never
 * moves (pop.fixed) and only acts on the others. Every moving member then
 * steps with one dt, the largest sqrt(2 D) + max|F| / (2 r) of them
 * (SRK.py:323-353); begin each moving member with the same time_vals. Removing
 * a member dissolves the group. */
int gcr_pool_group(gcr_pool* p, const int32_t* rids, int32_t n, const double* cross,
                   const int32_t* fixed);
/* Thread chunk of each chain (chain_chunk[n_chains], NULL: all 0): the
 * reference's split of a population among its threads (updates.py:323-344),
 * which active sliding depends on: a partner in the bead's own chunk counts
 * only when its index is lower (forces.py:98-102, 114-120). */
int gcr_replica_set_chunks(gcr_pool* p, int32_t rid, const int32_t* chain_chunk);
/* where_active after the last step (int32[n], reference bead order): -1 plus
 * one per sliding partner in each stage of the step (SRK.py:132,
 * forces.py:119); -1 without active sliding. */
int gcr_replica_where_active(gcr_pool* p, int32_t rid, int32_t* where_active);
/* Arms the next call of a replica (same meaning of time_vals / xi as gcr_run). */
int gcr_replica_begin(gcr_pool* p, int32_t rid, const double time_vals[8], const double* xi,
                      int64_t xi_steps);
/* Advances every armed replica by up to max_batches batches (<= 0: until all
 * are done). n_running receives how many are still running. A batch in which
 * some replica's beads left their cells by more than half the skin is
 * replayed for that replica with a re-sort before every step. */
int gcr_pool_run(gcr_pool* p, int32_t max_batches, int32_t* n_running);
/* [dt, elapsed, steps, max_forces_strength, running, replays, step_global,
 *  failed]. failed = 1: even with a cell re-sort before every step a bead
 * moved more than half the skin in one step; the replica stopped there (its
 * pair set could be incomplete) while the rest of the pool went on. */
int gcr_replica_status(gcr_pool* p, int32_t rid, double out[8]);
int gcr_replica_download(gcr_pool* p, int32_t rid, double* pos, double* pos1, double* tangent,
                         double* forces);
int gcr_replica_eval_forces(gcr_pool* p, int32_t rid, double* forces, double* tangent);
/* Order parameters of a replica, computed on the device from its chains'
 * mean directions phi_c = atan2(sum t_y, sum t_x) (output_methods.py:492-510,
 * tangents as gcr_replica_download returns them): out = [polar order
 * |sum e^{i phi}|/C, nematic order S, director angle, C]
 * (analysis.py:294-328). */
int gcr_replica_order(gcr_pool* p, int32_t rid, double out[4]);
/* The collective observables of a replica as a time series kept on the
 * device (no host round trip per sample). gcr_replica_sample appends one row
 * [simulated time since the replica was added, polar order, nematic order,
 * director angle -- analysis.py:294-328 over the chains' mean directions of
 * output_methods.py:492-510, as gcr_replica_order --, mean square
 * displacement of the chain centres of mass since the origin, steps since
 * the replica was added]; reset != 0 empties the series and makes the current
 * centres the origin (also on the first call). A chain's centre is the mean
 * of its beads unwrapped bond by bond from its first bead; its displacement is
 * followed sample to sample by the minimum image, so samples must come before
 * a centre has moved half a box. At most GCR_SERIES_MAX rows (GCR_ERR_CAPACITY
 * beyond). gcr_replica_series copies min(rows, max_rows) rows of
 * GCR_SERIES_COLS doubles and sets *rows to the rows held. */
#define GCR_SERIES_COLS 6
#define GCR_SERIES_MAX 8192
int gcr_replica_sample(gcr_pool* p, int32_t rid, int32_t reset);
int gcr_replica_series(gcr_pool* p, int32_t rid, double* out, int32_t max_rows, int32_t* rows);
/* The cell list of a replica's last re-sort, the counterpart of grid_list
 * and where_is_bead (updates.py:502-583): *nc = cells per side (edge L/nc,
 * cell index (z nc + y) nc + x of the bead's position at the re-sort,
 * int(x nc / L) clamped into [0, nc)); cstart[c] (nc^dim + 1 entries, NULL
 * skips) = the first slot of cell c; slot[b] (n_beads entries, NULL skips) =
 * the slot of bead b. Slots are sorted by cell and, inside a cell, by bead
 * index (a stable sort of the beads by cell). A replica's slots count from 0. */
int gcr_replica_cell_list(gcr_pool* p, int32_t rid, int32_t* nc, int32_t* cstart, int32_t* slot);
/* Per-kernel timing: runs `steps` steps of every armed replica with the
 * kernels of the batch graphs launched directly; stage0, stage1 and update
 * record start/end HIP events in their own dispatch (hipExtLaunchKernelGGL)
 * on the pool's stream. ms_out = average duration of [stage0, stage1, update,
 * cell re-sort (all its kernels, per re-sort)]. */
int gcr_pool_time_kernels(gcr_pool* p, int32_t steps, double ms_out[4]);
/* Batch graphs of the pool so far (no counterpart in the reference; the
 * cost of admitting and retiring replicas): out = [captures, instantiations,
 * in-place updates of an executable graph, host seconds spent on them]. */
int gcr_pool_graph_stats(gcr_pool* p, double out[4]);
/* Tile edge (cells) of the LDS force kernels; 0 picks it per replica from
 * the cell occupancy. 2D: square tiles of about 512 owned beads, the tile
 * with its halo within 70% of the LDS capacity at the mean occupancy; at
 * most 16. 3D: cubes, the largest whose halo holds at most 70% of the 3D
 * tile's LDS capacity at the mean occupancy, at most 8 cells while the
 * pool's tiles still number two per compute unit (else at most 6); a side
 * set here is clamped to 8 and to the grid. No result bit depends on the
 * tiling. */
int gcr_pool_set_tile(gcr_pool* p, int32_t side);
/* 2D only: tiles of tile_x x tile_y cells (each in [1, 16]) for every
 * replica instead of squares; 0, 0 returns to gcr_pool_set_tile's rule.
 * Setting a side with gcr_pool_set_tile clears the rectangle. Like the side,
 * the same grid constants of config.py:147-200 as they bear on speed, with
 * no effect on any result bit. */
int gcr_pool_set_tile_shape(gcr_pool* p, int32_t tile_x, int32_t tile_y);
/* Threads per tile workgroup of the force kernels: 256, 512, 1024, or 0 to
 * pick about one owned bead per thread in 2D, and 512 (FP32) or 256 (FP64)
 * threads per 3D tile. Neither this nor the tile
 * edge changes any result bit: the pair sums keep one order and one
 * arithmetic whatever the tiling. */
int gcr_pool_set_tile_threads(gcr_pool* p, int32_t threads);
int gcr_pool_destroy(gcr_pool* p);

/* ------------------------------------------------- initial configuration */

/* Host-side initial condition (no device work): n_chains straight chains of
 * chain_len[c] beads at spacing `bond`, centres uniform in the box and
 * directions uniform on the circle / sphere, placed one after another and
 * rejected while any bead comes closer than the exclusion distance to a
 * placed bead. The distance starts at min_dist and drops by 10% (not below
 * min_dist_floor) when a chain finds no place in max_tries attempts; the
 * value reached is returned in *min_dist_used. Bead 0 of each chain leads;
 * tangent[b] is the chain direction, pointing towards bead 0 as the
 * reference's tangents do (forces.py:351-359). Plays the role of
 * random_initialization + the growth phase (populations.py:530-557,
 * control.py:153-219). GCR_ERR_STATE when even the floor distance fails. */
int gcr_init_chains(int32_t dim, double L, int64_t n_chains, const int32_t* chain_len, double bond,
                    double min_dist, double min_dist_floor, uint64_t seed, int64_t max_tries,
                    double* pos, double* tangent, double* min_dist_used);

/* The text table of a .dat snapshot, as the numpy.savetxt call of the
 * reference's writer (output_methods.py:420-485) leaves it: the header's
 * lines prefixed "# ", then one line per row, columns joined by a space;
 * column k (cols[k * n_rows + r]) printed "%d" when is_int[k] (its values
 * integral), else "%.8e". Rows are formatted on n_threads host threads.
 * GCR_ERR_ARG for bad arguments or a file that cannot be written. */
int gcr_write_table(const char* path, const char* header, int64_t n_rows, int32_t n_cols,
                    const double* cols, const int32_t* is_int, int32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
