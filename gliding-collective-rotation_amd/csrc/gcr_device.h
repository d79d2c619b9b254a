// gcr_device.h -- device-side data layout shared by the kernels and the C-ABI
// host code. One pool holds many replicas (independent simulations) of the
// same dimensionality; a single large system is a pool with one replica.
//
// Per-bead arrays are indexed by "slot". Slots of replica r occupy
// [bead_off, bead_off + n). Within a replica, beads are physically ordered by
// cell (row-major, x fastest) and, inside a cell, by original bead index; the
// order is rebuilt every `rebuild_every` steps by a counting sort. The
// original index ("orig", the bead's index in the reference's arrays) travels
// with each bead and keys the noise, the exclusion rule and the download.
#pragma once
#include <stdint.h>

#define GCR_F_WCA 1u        // "WCA" in the force list (forces.py:57-144)
#define GCR_F_SPRINGS 2u    // "springs" (forces.py:285-402)
#define GCR_F_PROPEL 4u     // "self-propulsion" (forces.py:206-223)
#define GCR_F_ORIENT 8u     // spring_params[4] != 0: tangents from bonds
#define GCR_F_BEND 16u      // kappa != 0
#define GCR_F_ROTDIFF 32u   // Dr > 0 and not ORIENT (SRK.py:399)
#define GCR_F_INJECT 64u    // noise read from an injected array
#define GCR_F_FENE 128u     // spring_params[3]: FENE springs (forces.py:331-341)
#define GCR_F_ROD 256u      // rigid-rod projection of the updates (updates.py:347-499)
#define GCR_F_SRK4 512u     // SRK_order 4 (SRK.py:186-215, 297-321)
#define GCR_F_WALLS 1024u   // WCA walls (forces.py:405-435)
#define GCR_F_GROUP 2048u   // a moving population of a group: other populations act on it
                            // (updates.py:132-267, SRK.py:287-293)
#define GCR_F_SLIDE 4096u   // active sliding (forces.py:79-83, 114-120)
#define GCR_F_MOTOR 8192u   // a motor population (populations.py:617-660)
#define GCR_F_MSTILL 16384u // its bound motors do not diffuse (SRK.py:389-396)

// The law of one population acting on another (interpop_forces,
// updates.py:274-312): a WCA/LJ pair law and the source replica, with the
// half-width in the source's cells of the scan that finds every source bead
// within range.
struct XPair {
    int32_t src;
    int32_t window;
    int32_t kind;  // 1: WCA/LJ law, 2: the motors of src push
    int32_t pad;
    double p0, p1, shift, rc;
};

struct RepParams {
    int32_t n;          // beads
    int32_t bead_off;   // first slot
    int32_t chain_off;  // first chain (global chain index)
    int32_t n_chains;
    int32_t cell_off;   // first cell
    int32_t nc;         // cells per side (cubic grid)
    int32_t ncell;      // nc^dim
    int32_t n_nearby;   // n_nearby_beads_dont_repel
    uint32_t flags;
    int32_t nblk;       // bead blocks of this replica
    int32_t blk_off;    // first bead block
    int32_t red_off;    // first block of the stage-1 max reduction
    int32_t red_n;      // blocks of the stage-1 max reduction
    int32_t tile_x, tile_y;  // cell tile of the LDS force kernels (3D: cubes of tile_x)
    int32_t tile_off, ntile; // first tile, tiles of this replica
    double L, halfL;
    double h, inv_h;    // cell edge and its inverse
    double skin_half;   // half the skin: allowed displacement since the last re-sort
                        // (2D, neighbour lists) or excursion outside the cell box (3D)
    double rl2;         // squared neighbour-list radius (2D): h^2, slightly padded
    double wca[4];      // make_WCA_params (populations.py:48-58)
    double rc2;
    double k, kappa, theta0;  // spring_params[0..2]
    double fene_tol;          // spring_params[5]: FENE maximal stretch / rest length
    double rod;               // rigid_rod_projection_factor
    int32_t group;            // group index, -1: none
    int32_t grp_first, grp_n; // the group's moving replicas in GRP (the dt is theirs)
    int32_t xin_off, xin_n;   // laws of the other populations acting on this one (XPAIRS)
    int32_t rod_lanes;        // lanes per chain of the rigid-rod kernel (64 / chains per wave)
    double wall_zone;         // distance of the walls from the box faces
    double wall_rc;           // the walls' range parameter (wall_params[3])
    double slide_d2, slide_v0; // active sliding: squared partner range, push per partner
    double motor_d2, motor_imparts, k_on, k_off, motor_reach;  // a motor population
    int32_t motor_target;     // the replica its motors push (set by the group)
    int32_t pad2;
    int32_t walls;            // bit i: walls normal to axis i
    int32_t pad0;
    double D, Dr, radius;     // pop_consts[0..2]
    double dtmax, mdps, cheat;
    uint64_t seed;
    const double* xi_inject;  // [steps][n][2*dim] in orig order, or null
    int64_t xi_steps;
};

// Mutable per-replica step state. Written by k_finalize (after the stage-1
// launch) and by the batch control kernels only.
struct RepState {
    double dt;        // most recent dt (the predictor of the next step uses it)
    double dt_used;   // dt_used of the most recent step (corrector step size)
    double elapsed;   // sum of dt_used in the current call
    double maxm;      // max_forces_strength of the most recent step
    double mpart;     // a group member's own sqrt(2D) + max force term (the group takes the max)
    int64_t steps;        // steps done in the current call
    int64_t step_global;  // steps done since the replica was created (noise key)
    int64_t batch_base;   // `steps` at the start of the running batch
    int64_t max_steps;    // < 0: unbounded
    double mdu;           // max_dt_usable, < 0: unbounded
    int32_t violation;    // a bead left its cell box by more than skin/2
    int32_t frozen;       // 1: kernels leave this replica untouched
    int32_t pred_far;     // safe batch: bit k, the virtual positions of SRK stage k
                          // left the skin
    int32_t fixed;        // a fixed population of a group: never moves (pop.fixed)
    int32_t sort_due;     // safe batch: the last update left the skin, re-sort first
    int32_t pad;
    double t_total;       // simulated time since the replica was added (sum of dt_used)
};

struct PoolArgs {
    // block tables for bead-parallel kernels
    const int32_t* blk_rep;
    const int32_t* blk_slot0;
    int32_t nblk;
    // chain-block tables of the replicas with rigid-rod projection
    const int32_t* chb_rep;
    const int32_t* chb_chain0;
    int32_t nchb;
    // chunk tables for the cell scan
    const int32_t* chk_rep;
    const int32_t* chk_cell0;
    int32_t nchk;
    const int32_t* rep_chk_off;  // first chunk of each replica (+1 entry)
    int32_t nrep;
    // replica tables
    const RepParams* P;
    RepState* S;
    double* BMAX;  // [block] max |F0+F1|^2 of the block (stage 1)
    int32_t* SCNT; // [slot] active sliding partners of the running force stage
    int32_t* TILES;        // [tile][4] (replica, cx0, cy0, ints of the tile's layout table to read:
                           // set by the list builder, 0 = the table's first TT_PERM) in 2D,
                           // (replica, cx0, cy0, cz0) in 3D
    int32_t* TT;           // [tile][TT_INTS] (3D: TT3_INTS) LDS layout of each tile, written at every re-sort
    int32_t ntile;
    // per-slot state
    double* X;     // [slot][dim] positions
    double* X1;    // [slot][dim] predictor positions
    double* F0;    // [slot][dim] stage-0 forces
    double* FT;    // [slot][dim] F0 + F1
    double* XI;    // [slot][dim] translational noise of the next / running
                   // step; moves with the bead at re-sorts
    double* XIR;   // [slot][dim] rotational noise (diffusing directions)
    double* T;     // [slot][dim] propulsion direction (ABP / diffusing state)
    int32_t* META; // [slot][2] (orig_global, chain_global)
    int32_t* LINK; // [slot][4] (prev2, prev1, next1, next2) slots, -1 if none
    int32_t* CELL; // [slot] cell (replica-local) the bead was binned into
    int32_t* SLOT; // [orig_global] -> slot
    // neighbour lists, rebuilt with every cell re-sort: [group][slot][4] u16
    // LDS indices in the order the bead's tile stages its halo
    uint16_t* NL;
    int32_t* NLEN; // [slot] list length, -1: no list (pairs from the cell scan)
    double* XR;    // [slot][dim] positions at the last re-sort
    int64_t nl_stride;  // slots per group (the pool's bead capacity)
    // per-chain constants
    const double* CHAIN_A;   // rest length
    const double* CHAIN_CS;  // [chain][3] (v0, cos v_tan_angle, sin v_tan_angle)
    const int32_t* CHAIN_FIRST;  // first orig_global of the chain
    const int32_t* CHAIN_LEN;
    // cells
    int32_t* CCOUNT;  // [cell] counts during a rebuild
    int32_t* CSTART;  // [cell + 1] first slot of each cell
    int32_t* CHKSUM;  // [chunk] partial sums for the scan
    // rebuild scratch
    int32_t* KEYOFF;  // [slot][2] (cell, offset within cell)
    int32_t* PAIRS;   // [slot][2] (orig_global, old slot)
    double* Xn;
    double* Tn;
    int32_t* METAn;
    double* XIn;
    double* XIRn;
    double* X1n;
    // rollback snapshot
    double* Xs;
    double* Ts;
    int32_t* METAs;
    RepState* Ss;
    // launch modes, fixed when a batch graph is captured:
    // safe  - the batch that re-sorts before every step: a corrector's skin
    //         check is moot there (the next step re-sorts first) and a
    //         predictor that leaves the skin sets pred_far, which re-sorts the
    //         replica on the predictor positions before the second stage;
    // on_x1 - that mid-step re-sort: cells, lists and XR from X1, and the
    //         live force array (SRK2's stage-0 forces, SRK4's running
    //         total) travels with the beads
    // active sliding: thread chunk per chain, where_active per slot (+ staging)
    const int32_t* CHAIN_CHUNK;
    int32_t* WA;      // where_active; for a motor population the bound bead, -1 free
    int32_t* WAn;
    int32_t* MC;      // bindings a bead received in the running stage
    // populations acting on each other
    const XPair* XPAIRS;
    const int32_t* GRP;
    int32_t safe;
    int32_t on_x1;
    // SRK stage of this launch: 0 the first force evaluation (at X), 1 and 2
    // the inner SRK4 evaluations, 3 the last (at X1); the re-sort on x1
    // keys on the virtual positions written by stage `sub`
    int32_t sub;
    // threads per tile of the 2D force stages (the list builder orders each
    // tile's owned beads for that many)
    int32_t nt_stage;
};
