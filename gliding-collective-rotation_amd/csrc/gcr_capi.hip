// gcr_capi.hip -- host side of libgcr.so: pools of replicas, device memory,
// graph-captured step batches, replay of a batch when the cell-list skin was
// exceeded, and the extern "C" entry points declared in include/gcr.h.
#include "gcr_kernels.hip"
#include "gcr_tile2d.hip"
#include "gcr_tile3d.hip"

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <utility>
#include <vector>

#include "gcr.h"

using namespace gcr;

namespace {

thread_local std::string g_err;

// RepState::frozen values: 0 running, 1 held by the host (replay or
// diagnostics), 2 marked for restore, FROZEN_FAILED stopped for good
constexpr int FROZEN_FAILED = 3;
// after a second replay in a row, the replica's next batches start in the
// safe form (while every running replica of the pool is in that state)
constexpr int SAFE_STICKY = 8;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

}  // namespace

// the message gcr_last_error returns, for the library's host-only sources
__attribute__((visibility("hidden"))) int gcr_internal_fail(int code, const char* msg)
{
    return fail(code, msg ? msg : "");
}

namespace {

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(GCR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

// First-fit allocator of index ranges (beads, cells, chains) inside a pool.
struct Segs {
    std::vector<std::pair<int64_t, int64_t>> fr;  // (start, length), sorted
    void init(int64_t cap) { fr.assign(1, {0, cap}); }
    int64_t alloc(int64_t n)
    {
        for (size_t i = 0; i < fr.size(); ++i)
            if (fr[i].second >= n) {
                int64_t s = fr[i].first;
                fr[i].first += n;
                fr[i].second -= n;
                if (fr[i].second == 0) fr.erase(fr.begin() + i);
                return s;
            }
        return -1;
    }
    void release(int64_t s, int64_t n)
    {
        auto it = std::lower_bound(fr.begin(), fr.end(), std::make_pair(s, (int64_t)0));
        it = fr.insert(it, {s, n});
        size_t i = it - fr.begin();
        if (i + 1 < fr.size() && fr[i].first + fr[i].second == fr[i + 1].first) {
            fr[i].second += fr[i + 1].second;
            fr.erase(fr.begin() + i + 1);
        }
        if (i > 0 && fr[i - 1].first + fr[i - 1].second == fr[i].first) {
            fr[i - 1].second += fr[i].second;
            fr.erase(fr.begin() + i);
        }
    }
};

template <typename T>
int dalloc(T** p, size_t count)
{
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
    return GCR_OK;
}

int64_t cells_per_side(double L, double rc, double skin)
{
    int64_t nc = (int64_t)std::floor(L / (rc + skin));
    return nc < 1 ? 1 : nc;
}

double default_skin(const gcr_replica_desc* d) { return d->skin > 0 ? d->skin : 0.3 * d->wca[3]; }

// Cells per side of a replica's grid. With a pair force the edge is at least
// the cutoff plus the skin. Without one the replica's own forces scan no
// cells, and the scans of other populations (cross forces, motors, which add
// their reach to the edge) and the re-sort's ranking within a cell want few
// beads per cell: an edge of about four beads per cell on average.
int64_t grid_side(int32_t dim, const gcr_replica_desc* d)
{
    if (d->has_wca) return cells_per_side(d->L, d->wca[3], default_skin(d));
    if (d->n_beads <= 8 || !(d->L > 0)) return 1;
    const double h = std::pow(4.0 * std::pow(d->L, dim) / (double)d->n_beads, 1.0 / dim);
    const int64_t cap = dim == 3 ? 512 : 16384;
    const int64_t nc = (int64_t)std::floor(d->L / h);
    return nc < 1 ? 1 : (nc > cap ? cap : nc);
}

}  // namespace

struct gcr_pool {
    int dim = 2, precision = GCR_FP64;
    int64_t capB = 0, capC = 0, capCh = 0;
    int capR = 0;
    Segs segB, segC, segCh;
    std::vector<RepParams> P;
    // host copy of the replicas' step states, in page-locked memory: it goes
    // to and from the device around every batch, and a copy from pageable
    // memory is staged through a bounce buffer (about 20 us for a few
    // hundred bytes)
    RepState* S = nullptr;
    std::vector<int> live;
    std::vector<int> replays;
    std::vector<int> safe_left;  // per replica: batches still started in the safe form
    std::vector<int> replayed;   // per replica: its last normal batch was replayed
    std::vector<double*> xi_dev;
    std::vector<double*> ser_dev;  // per replica: its observables series (GCR_SERIES_MAX rows)
    std::vector<int> ser_rows;
    int B = 32, K = 8;
    bool tables_dirty = true;
    hipGraphExec_t g_norm = nullptr, g_safe = nullptr;
    // the tables changed since a graph was captured: it is brought up to date
    // before its next launch (re-captured, and the executable graph updated in
    // place when the launch sequence kept its shape)
    bool norm_stale = false, safe_stale = false;
    int64_t n_capture = 0, n_instantiate = 0, n_update = 0;
    double graph_seconds = 0.0;  // host time spent capturing, instantiating, updating
    int nblk = 0, nchk = 0;
    hipStream_t stream = nullptr;
    PoolArgs a{};
    // device buffers owned by the pool
    int32_t *blk_rep = nullptr, *blk_slot0 = nullptr, *chk_rep = nullptr, *chk_cell0 = nullptr,
            *rep_chk_off = nullptr, *chb_rep = nullptr, *chb_chain0 = nullptr;
    int nchb = 0;  // chain blocks of the rigid-rod replicas
    RepParams* dP = nullptr;
    RepState *dS = nullptr, *dSs = nullptr;
    double *X = nullptr, *X1 = nullptr, *F0 = nullptr, *FT = nullptr, *XI = nullptr, *T = nullptr,
           *Xn = nullptr, *Tn = nullptr, *Xs = nullptr, *Ts = nullptr, *XIn = nullptr, *X1n = nullptr;
    int32_t *META = nullptr, *LINK = nullptr, *CELL = nullptr, *SLOT = nullptr, *KEYOFF = nullptr,
            *PAIRS = nullptr, *METAn = nullptr, *METAs = nullptr;
    double *CHAIN_A = nullptr, *CHAIN_CS = nullptr;
    int32_t *CHAIN_FIRST = nullptr, *CHAIN_LEN = nullptr;
    int32_t *CCOUNT = nullptr, *CSTART = nullptr, *CHKSUM = nullptr;
    double* BMAX = nullptr;
    int32_t* SCNT = nullptr;  // per slot: active sliding partners of the running stage
    int32_t* TILES = nullptr;
    int32_t* TT = nullptr;  // tile layout tables (TT_INTS / TT3_INTS per tile), sized to the tiles in use
    int64_t capTT = 0;
    uint16_t* NL = nullptr;
    int32_t* NLEN = nullptr;
    double* XR = nullptr;
    double* OBS = nullptr;  // order-parameter partial sums (5 per chain block + 5)
    double *COMW = nullptr, *COMU = nullptr;  // per chain: last centre, unwrapped displacement
    int64_t capChk = 0, capBlk = 0, capTile = 0, capChb = 0;
    int ntile = 0;
    int tile_side = 0;     // 0: chosen per replica from the cell occupancy
    int tile_w = 0, tile_h = 0;  // 2D: a rectangular tile (cells in x, y); 0: tile_side rules
    int n_cu = 256;        // compute units of the device (3D tile sides)
    int tile_threads = 0;  // threads per tile workgroup: 256, 512, 1024, 0 = from the tile sizes
    int nt_launch = BS;    // resolved by rebuild_tables
    int nt_list = BS;      // threads per tile of the list builder (rebuild_tables)
    bool has_srk4 = false; // a live replica integrates with SRK4: the step has 4 force stages
    bool tile_gen = false; // a live replica needs the general tile kernels (SRK4, walls, FENE)
    // populations acting on each other: replicas in population order, the law
    // of b on a at cross[(a * n + b) * 5] (flag, p0, p1, shift, cutoff), fixed
    // populations
    struct Group {
        std::vector<int> rids;
        std::vector<double> cross;
        std::vector<int> fixed;
    };
    std::vector<Group> groups;
    XPair* XPAIRS = nullptr;
    int32_t* GRP = nullptr;
    int32_t *CHAIN_CHUNK = nullptr, *WA = nullptr, *WAn = nullptr, *MC = nullptr;
    int64_t capX = 0;
    bool has_group = false;  // a group exists or a replica has active sliding: the
                             // second half of each force stage runs (k_group_stage)
    bool has_motor = false;  // a group has motors: k_motor runs in each force stage
    std::vector<int> motor_reps;  // the motor populations k_motor runs over
    std::vector<int> slide_reps;  // the replicas with active sliding
};

namespace {

void drop_graphs(gcr_pool* p)
{
    if (p->g_norm) (void)hipGraphExecDestroy(p->g_norm);
    if (p->g_safe) (void)hipGraphExecDestroy(p->g_safe);
    p->g_norm = p->g_safe = nullptr;
}

int upload_rep(gcr_pool* p, int r)
{
    HIPCHK(hipMemcpyAsync(p->dP + r, &p->P[r], sizeof(RepParams), hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->dS + r, &p->S[r], sizeof(RepState), hipMemcpyHostToDevice, p->stream));
    return GCR_OK;
}

int pull_states(gcr_pool* p)
{
    HIPCHK(hipMemcpyAsync(p->S, p->dS, sizeof(RepState) * p->capR, hipMemcpyDeviceToHost,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int push_states(gcr_pool* p)
{
    HIPCHK(hipMemcpyAsync(p->dS, p->S, sizeof(RepState) * p->capR, hipMemcpyHostToDevice,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

// Tile edge (cells) of the 2D LDS force kernel: about 2 BS owned beads per
// workgroup (fewer, larger tiles stage less halo and leave no partial last
// wave of workgroups on large grids), at most TILE_MAX_ROWS, the tile with its
// halo within TILE_MAX_BEADS at the mean occupancy, and never more than the
// grid.
int pick_tile(const RepParams& P, int forced)
{
    int side = forced;
    if (side <= 0) {
        const double occ = std::max((double)P.n / (double)P.ncell, 1e-3);
        side = (int)std::lround(std::sqrt(2.0 * BS / occ));
        while (side > 1 && (side + 2) * (side + 2) * occ > 0.7 * TILE_MAX_BEADS) --side;
    }
    side = std::max(1, std::min(side, TILE_MAX_ROWS));
    return std::min(side, P.nc);
}

// Tile side (cells) of the 3D LDS force kernel: the largest cube whose tile
// with its halo holds at most 70% of T3_MAX_BEADS at the mean occupancy
// (a 3D halo is 2-4x the owned cells, so tiles stay small), at most `cap`
// and the grid; a forced side is clamped to T3_MAX_SIDE and the grid.
int pick_tile3(const RepParams& P, int forced, int cap = T3_MAX_SIDE)
{
    const double occ = std::max((double)P.n / (double)P.ncell, 1e-3);
    int side = forced > 0 ? forced : cap;
    side = std::max(1, std::min(side, T3_MAX_SIDE));
    if (forced <= 0)
        while (side > 1 && (side + 2) * (side + 2) * (side + 2) * occ > 0.7 * T3_MAX_BEADS) --side;
    return std::min(side, P.nc);
}

// The largest side cap (8, 7 or 6) at which the pool's 3D tiles still number
// at least two per compute unit: cubes larger than 6 cells stage less halo
// per owned bead, but a launch of fewer tiles than that leaves compute units
// with one tile or none (106k 3D chain beads: 45 us per step at 6 cells, 512
// tiles; 64 us at 7 cells, 343 tiles; 197k beads: 74 us at 6 cells, 1000
// tiles; 62 us at 7 cells, 512 tiles). 6 when no cap reaches that.
int pick_cap3(const gcr_pool* p)
{
    for (int cap = T3_MAX_SIDE; cap > 6; --cap) {
        int64_t tiles = 0;
        for (int r = 0; r < p->capR; ++r) {
            if (!p->live[r]) continue;
            const RepParams& P = p->P[r];
            const int side = pick_tile3(P, 0, cap);
            const int64_t per = (P.nc + side - 1) / side;
            tiles += per * per * per;
        }
        if (tiles >= 2 * (int64_t)p->n_cu) return cap;
    }
    return 6;
}

// The dispatcher hands a launch's workgroups to the 8 XCDs of the MI355X
// round-robin by workgroup index, and each XCD has its own L2. order[b] is
// the tile (index into a row-major list of n tiles) that workgroup
// first + b runs: each XCD gets one contiguous run of tiles (a band of
// rows), so neighbouring tiles, which read each other's cells as halo, mostly
// share an L2.
// Ints of a tile's layout table.
int tt_ints(int dim) { return dim == 2 ? TT_INTS : TT3_INTS; }

std::vector<int> xcd_order(int first, int n)
{
    constexpr int NXCD = 8;
    int cnt[NXCD] = {0}, start[NXCD];
    for (int b = 0; b < n; ++b) ++cnt[(first + b) % NXCD];
    int acc = 0;
    for (int x = 0; x < NXCD; ++x) {
        start[x] = acc;
        acc += cnt[x];
    }
    std::vector<int> order(n);
    int seen[NXCD] = {0};
    for (int b = 0; b < n; ++b) {
        const int x = (first + b) % NXCD;
        order[b] = start[x] + seen[x]++;
    }
    return order;
}

// The physics fields of a replica's parameters (everything but its place in
// the pool, its cell grid and its noise key).
void set_physics(RepParams& P, const gcr_replica_desc* d, int longest_chain)
{
    P.n_nearby = d->n_nearby;
    // without a pair force the half skin is half the cell edge: what the
    // scans of other populations add to their reach (none on a single cell)
    P.skin_half = d->has_wca ? 0.5 * (P.h - d->wca[3]) : (P.nc > 1 ? 0.5 * P.h : 1e300);
    for (int i = 0; i < 4; ++i) P.wca[i] = d->wca[i];
    P.rc2 = d->wca[3] * d->wca[3];
    P.k = d->spring[0];
    P.kappa = d->spring[1];
    P.theta0 = d->spring[2];
    P.fene_tol = d->spring[5];
    P.rod = d->rod;
    P.D = d->D;
    P.Dr = d->Dr;
    P.radius = d->radius;
    uint32_t fl = P.flags & (GCR_F_INJECT | GCR_F_GROUP);
    if (d->has_wca) fl |= GCR_F_WCA;
    if (d->has_springs) fl |= GCR_F_SPRINGS;
    if (d->has_propel) fl |= GCR_F_PROPEL;
    bool orient = d->has_springs && d->spring[4] != 0.0;
    if (orient) fl |= GCR_F_ORIENT;
    if (d->has_springs && d->spring[1] != 0.0) fl |= GCR_F_BEND;
    if (d->has_springs && d->spring[3] != 0.0) fl |= GCR_F_FENE;
    // SRK.py:364: projection when the factor is set and the longest chain has
    // more than two beads
    if (d->rod != 0.0 && longest_chain > 2) fl |= GCR_F_ROD;
    {  // as many chains per wavefront of k_rod as fit (longer ones in segments)
        const int cpw = longest_chain >= 1 && longest_chain <= 64 ? 64 / longest_chain : 1;
        P.rod_lanes = 64 / std::min(cpw, ROD_MAX_CPW);
    }
    if (d->Dr > 0 && !orient) fl |= GCR_F_ROTDIFF;
    if (d->srk_order == 4) fl |= GCR_F_SRK4;
    P.walls = d->walls & 7;  // the kernels read the bits of their dimensions
    P.slide_d2 = d->sliding_maxdist * d->sliding_maxdist;
    P.slide_v0 = d->sliding_v0;
    if (d->sliding_maxdist > 0.0) fl |= GCR_F_SLIDE;
    P.motor_d2 = d->motor_max_dist * d->motor_max_dist;
    P.motor_imparts = d->motor_imparts;
    P.k_on = d->k_on;
    P.k_off = d->k_off;
    P.motor_reach = d->motor_reach > 0.0 ? d->motor_reach : d->motor_max_dist;
    if (d->motor) fl |= GCR_F_MOTOR;
    if (d->motor && d->motor_still) fl |= GCR_F_MSTILL;
    P.wall_zone = d->wall_zone;
    P.wall_rc = d->wall_cutoff > 0.0 ? d->wall_cutoff : d->wca[3];
    if (P.walls) fl |= GCR_F_WALLS;
    P.flags = fl;
}

bool srk_order_ok(const gcr_replica_desc* d)
{
    if (d->srk_order == 4 && d->sliding_maxdist > 0.0) return false;  // not combined
    return d->srk_order == 0 || d->srk_order == 2 || d->srk_order == 4;
}

// Block, chunk and tile tables of the live replicas.
int rebuild_tables(gcr_pool* p)
{
    std::vector<int32_t> br, bs, cr, cc, rco(p->capR + 1, 0), tl, hr, hc;
    bool has_srk4 = false, tile_gen = false;
    std::vector<XPair> xp;
    std::vector<int32_t> grp;
    bool has_motor = false;
    std::vector<int> motor_reps;
    for (auto& P : p->P) {
        P.group = -1;
        P.motor_target = -1;
    }
    for (int g = 0; g < (int)p->groups.size(); ++g) {
        // members in population order; each moving member takes the laws of
        // the others in the order (Q - P) mod n and the group's moving list
        const gcr_pool::Group& G = p->groups[g];
        const int n = (int)G.rids.size();
        const int32_t first = (int32_t)grp.size();
        for (int k = 0; k < n; ++k)
            if (!G.fixed[k]) grp.push_back(G.rids[k]);
        for (int k = 0; k < n; ++k) {
            RepParams& P = p->P[G.rids[k]];
            P.group = g;
            P.grp_first = first;
            P.grp_n = (int32_t)grp.size() - first;
            P.xin_off = (int32_t)xp.size();
            P.xin_n = 0;
            if (G.fixed[k]) continue;
            for (int ii = 1; ii < n; ++ii) {
                const int o = (k + ii) % n;
                const double* c = &G.cross[((size_t)k * n + o) * 5];
                if (c[0] == 0.0) continue;
                RepParams& Q = p->P[G.rids[o]];
                XPair e;
                e.src = G.rids[o];
                e.kind = c[0] == 2.0 ? 2 : 1;
                e.pad = 0;
                if (e.kind == 2) {  // motors of Q push this population
                    Q.motor_target = G.rids[k];
                    has_motor = true;
                    if (std::find(motor_reps.begin(), motor_reps.end(), G.rids[o]) == motor_reps.end())
                        motor_reps.push_back(G.rids[o]);
                }
                const double reach = c[4] + (Q.skin_half < Q.L ? Q.skin_half : Q.L);
                e.window = std::max(1, (int)std::ceil(reach / Q.h));
                e.p0 = c[1];
                e.p1 = c[2];
                e.shift = c[3];
                e.rc = c[4];
                xp.push_back(e);
                ++P.xin_n;
            }
        }
    }
    if ((int64_t)xp.size() > p->capX) return fail(GCR_ERR_CAPACITY, "population law table full");
    const int cap3 = p->dim == 3 ? pick_cap3(p) : 0;
    for (int r = 0; r < p->capR; ++r) {
        rco[r] = (int32_t)cr.size();
        RepParams& P = p->P[r];
        if (!p->live[r]) continue;
        P.blk_off = (int32_t)br.size();
        P.nblk = (P.n + BS - 1) / BS;
        for (int b = 0; b < P.nblk; ++b) {
            br.push_back(r);
            bs.push_back(P.bead_off + b * BS);
        }
        for (int c = 0; c < P.ncell; c += SCAN_CHUNK) {
            cr.push_back(r);
            cc.push_back(P.cell_off + c);
        }
        if (P.flags & GCR_F_SRK4) has_srk4 = true;
        if (P.flags & (GCR_F_SRK4 | GCR_F_WALLS | GCR_F_FENE | GCR_F_GROUP | GCR_F_SLIDE))
            tile_gen = true;
        if (P.flags & GCR_F_ROD)
            for (int c = 0; c < P.n_chains; c += BS) {
                hr.push_back(r);
                hc.push_back(P.chain_off + c);
            }
        if (p->dim == 2) {
            // a square tile unless the pool sets a rectangle (rows of more
            // cells in x are longer runs of slots in the halo's loads)
            const int side = pick_tile(P, p->tile_side);
            const int tx = p->tile_w > 0 ? std::min(p->tile_w, P.nc) : side;
            const int ty = p->tile_h > 0 ? std::min(p->tile_h, P.nc) : side;
            P.tile_x = tx;
            P.tile_y = ty;
            P.tile_off = (int32_t)(tl.size() / 4);
            std::vector<int32_t> rowmajor;  // (x, y) of the tiles, row by row
            for (int y = 0; y < P.nc; y += ty)
                for (int x = 0; x < P.nc; x += tx) {
                    rowmajor.push_back(x);
                    rowmajor.push_back(y);
                }
            const int n = (int)rowmajor.size() / 2;
            const std::vector<int> order = xcd_order(P.tile_off, n);
            for (int b = 0; b < n; ++b) {
                tl.push_back(r);
                tl.push_back(rowmajor[2 * order[b]]);
                tl.push_back(rowmajor[2 * order[b] + 1]);
                tl.push_back(0);
            }
            P.ntile = n;
            P.red_off = P.tile_off;
            P.red_n = P.ntile;
        } else {  // cubic tiles (replica, cx0, cy0, cz0), XCD bands of (x, y, z) order
            const int side = pick_tile3(P, p->tile_side, cap3);
            P.tile_x = P.tile_y = side;
            P.tile_off = (int32_t)(tl.size() / 4);
            std::vector<int32_t> order3;
            for (int z = 0; z < P.nc; z += side)
                for (int y = 0; y < P.nc; y += side)
                    for (int x = 0; x < P.nc; x += side) {
                        order3.push_back(x);
                        order3.push_back(y);
                        order3.push_back(z);
                    }
            const int n = (int)order3.size() / 3;
            const std::vector<int> order = xcd_order(P.tile_off, n);
            for (int b = 0; b < n; ++b) {
                tl.push_back(r);
                tl.push_back(order3[3 * order[b]]);
                tl.push_back(order3[3 * order[b] + 1]);
                tl.push_back(order3[3 * order[b] + 2]);
            }
            P.ntile = n;
            P.red_off = P.tile_off;
            P.red_n = P.ntile;
        }
        // k_group_stage writes the maxima of its blocks (GROUP_LANES per bead
        // block) after the tile and stage-block maxima
        if (P.flags & (GCR_F_GROUP | GCR_F_SLIDE)) {
            P.red_off = (int32_t)p->capTile + GROUP_LANES * P.blk_off;
            P.red_n = GROUP_LANES * P.nblk;
        }
    }
    rco[p->capR] = (int32_t)cr.size();
    {
        // about one owned bead per thread: the per-bead chain of dependent
        // loads is long, and only more waves per CU hide it
        int64_t beads = 0;
        for (int r = 0; r < p->capR; ++r)
            if (p->live[r]) beads += p->P[r].n;
        const int64_t nt = (int64_t)tl.size() / 4;
        const int64_t per = nt ? beads / nt : 0;
        if (p->tile_threads)
            p->nt_launch = p->tile_threads;
        else if (p->dim == 2)
            p->nt_launch = per >= 768 ? 1024 : per >= 384 ? 512 : 256;
        else  // 3D: the halo is 3-5x the owned beads, and staging it is the
              // longer part of the prologue; FP64 records (40 B) leave room
              // for two tiles per CU, whose 512 threads would then idle
              // (100k 3D chain beads: FP32 25 vs 37 us per stage at 512 vs
              // 256 threads, FP64 41 vs 46 us at 256 vs 512)
            p->nt_launch = p->precision == GCR_MIXED ? 512 : 256;
        // the 2D list builder: one owned bead per thread whatever the force
        // stages run at (each bead's 3 x 3 scan is a long chain of LDS reads)
        p->nt_list = p->dim == 2 ? (per >= 768 ? 1024 : per >= 384 ? 512 : 256) : p->nt_launch;
    }
    if ((int64_t)br.size() > p->capBlk || (int64_t)cr.size() > p->capChk ||
        (int64_t)tl.size() / 4 > p->capTile || (int64_t)hr.size() > p->capChb)
        return fail(GCR_ERR_CAPACITY, "block table capacity exceeded");
    p->nchb = (int)hr.size();
    p->has_srk4 = has_srk4;
    p->tile_gen = tile_gen;
    p->has_group = !p->groups.empty();
    p->has_motor = has_motor;
    p->motor_reps = motor_reps;
    p->slide_reps.clear();
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && (p->P[r].flags & GCR_F_SLIDE)) {
            p->has_group = true;
            p->slide_reps.push_back(r);
        }
    if (!xp.empty())
        HIPCHK(hipMemcpyAsync(p->XPAIRS, xp.data(), xp.size() * sizeof(XPair), hipMemcpyHostToDevice,
                              p->stream));
    if (!grp.empty())
        HIPCHK(hipMemcpyAsync(p->GRP, grp.data(), grp.size() * 4, hipMemcpyHostToDevice, p->stream));
    if (p->nchb) {
        HIPCHK(hipMemcpyAsync(p->chb_rep, hr.data(), hr.size() * 4, hipMemcpyHostToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(p->chb_chain0, hc.data(), hc.size() * 4, hipMemcpyHostToDevice, p->stream));
    }
    p->nblk = (int)br.size();
    p->nchk = (int)cr.size();
    p->ntile = (int)(tl.size() / 4);
    if (p->nblk) {
        HIPCHK(hipMemcpyAsync(p->blk_rep, br.data(), br.size() * 4, hipMemcpyHostToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(p->blk_slot0, bs.data(), bs.size() * 4, hipMemcpyHostToDevice, p->stream));
    }
    if (p->nchk) {
        HIPCHK(hipMemcpyAsync(p->chk_rep, cr.data(), cr.size() * 4, hipMemcpyHostToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(p->chk_cell0, cc.data(), cc.size() * 4, hipMemcpyHostToDevice, p->stream));
    }
    if (p->ntile)
        HIPCHK(hipMemcpyAsync(p->TILES, tl.data(), tl.size() * 4, hipMemcpyHostToDevice, p->stream));
    if (p->ntile > p->capTT) {  // the tile layout tables grow with the tiles in use
        HIPCHK(hipStreamSynchronize(p->stream));
        if (p->TT) HIPCHK(hipFree(p->TT));
        p->TT = nullptr;
        p->capTT = 0;
        HIPCHK(hipMalloc((void**)&p->TT, sizeof(int32_t) * tt_ints(p->dim) * (size_t)p->ntile));
        p->capTT = p->ntile;
        p->a.TT = p->TT;
    }
    // tables of the new tiles are filled by the next re-sort; until then a
    // tile takes the path that computes its layout in the kernel
    if (p->ntile) {
        const int n = p->ntile * tt_ints(p->dim);
        k_fill_i32<<<(n + BS - 1) / BS, BS, 0, p->stream>>>(p->TT, n, 0);
    }
    HIPCHK(hipMemcpyAsync(p->rep_chk_off, rco.data(), rco.size() * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->dP, p->P.data(), sizeof(RepParams) * p->capR, hipMemcpyHostToDevice,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->a.nblk = p->nblk;
    p->a.nchk = p->nchk;
    p->a.nchb = p->nchb;
    p->a.ntile = p->ntile;
    p->a.nt_stage = p->nt_launch;
    p->tables_dirty = false;
    p->norm_stale = p->safe_stale = true;
    return GCR_OK;
}

// Kernel sequences (launched directly or captured into a graph).
// Cell re-sort, neighbour lists and XR; on_x1: the mid-step re-sort of the
// safe batch, keyed on the virtual positions of SRK stage `sub`.
void launch_rebuild(gcr_pool* p, int j, bool on_x1 = false, int sub = 0)
{
    PoolArgs a = p->a;
    a.on_x1 = on_x1 ? 1 : 0;
    a.sub = sub;
    hipStream_t st = p->stream;
    if (p->dim == 2)
        k_bin<2><<<p->nblk, BS, 0, st>>>(a, j);
    else
        k_bin<3><<<p->nblk, BS, 0, st>>>(a, j);
    k_scan_chunks<<<p->nchk, BS, 0, st>>>(a, j);
    k_scan_reps<<<p->capR, BS, 0, st>>>(a, j);
    k_scan_final<<<p->nchk, BS, 0, st>>>(a, j);
    k_scatter<<<p->nblk, BS, 0, st>>>(a, j);
    if (p->dim == 2) {
        k_gather<2><<<p->nblk, BS, 0, st>>>(a, 1, j);
        k_links<2><<<p->nblk, BS, 0, st>>>(a, 1, j);
        if (p->ntile) {
            const int nt = p->nt_list;
            if (p->precision == GCR_MIXED) {
                if (nt == 1024)
                    k_nlist_tile2d<float, 1024><<<p->ntile, 1024, 0, st>>>(a, j);
                else if (nt == 512)
                    k_nlist_tile2d<float, 512><<<p->ntile, 512, 0, st>>>(a, j);
                else
                    k_nlist_tile2d<float, 256><<<p->ntile, 256, 0, st>>>(a, j);
            } else {
                if (nt == 1024)
                    k_nlist_tile2d<double, 1024><<<p->ntile, 1024, 0, st>>>(a, j);
                else if (nt == 512)
                    k_nlist_tile2d<double, 512><<<p->ntile, 512, 0, st>>>(a, j);
                else
                    k_nlist_tile2d<double, 256><<<p->ntile, 256, 0, st>>>(a, j);
            }
        }
    } else {
        k_gather<3><<<p->nblk, BS, 0, st>>>(a, 1, j);
        k_links<3><<<p->nblk, BS, 0, st>>>(a, 1, j);
        if (p->ntile) {
            const int nt = p->nt_launch;
            if (p->precision == GCR_MIXED) {
                if (nt == 1024)
                    k_nlist_tile3d<float, 1024><<<p->ntile, 1024, 0, st>>>(a, j);
                else if (nt == 512)
                    k_nlist_tile3d<float, 512><<<p->ntile, 512, 0, st>>>(a, j);
                else
                    k_nlist_tile3d<float, 256><<<p->ntile, 256, 0, st>>>(a, j);
            } else {
                if (nt == 1024)
                    k_nlist_tile3d<double, 1024><<<p->ntile, 1024, 0, st>>>(a, j);
                else if (nt == 512)
                    k_nlist_tile3d<double, 512><<<p->ntile, 512, 0, st>>>(a, j);
                else
                    k_nlist_tile3d<double, 256><<<p->ntile, 256, 0, st>>>(a, j);
            }
        }
    }
}

// Launch one kernel of a step; with events (ev0, ev1) the dispatch itself
// records its start and end (hipExtLaunchKernelGGL), which times the kernel
// alone on the pool's stream.
template <typename K>
void launch_k(K kernel, int grid, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, const PoolArgs& a, int j,
              int block = BS)
{
    if (grid <= 0) return;
    if (ev0)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, ev0, ev1, 0, a, j);
    else
        kernel<<<grid, block, 0, st>>>(a, j);
}

// 2D tile kernels at the pool's threads per tile, in their lean variant
// unless a replica needs the general one.
template <typename FT, int STAGE, bool GEN>
void launch_tile_v(gcr_pool* p, const PoolArgs& a, int j, hipEvent_t ev0, hipEvent_t ev1)
{
    if (p->nt_launch == 1024)
        launch_k(k_stage_tile2d<FT, STAGE, 1024, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 1024);
    else if (p->nt_launch == 512)
        launch_k(k_stage_tile2d<FT, STAGE, 512, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 512);
    else
        launch_k(k_stage_tile2d<FT, STAGE, 256, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 256);
}

template <typename FT, int STAGE>
void launch_tile(gcr_pool* p, const PoolArgs& a, int j, hipEvent_t ev0, hipEvent_t ev1)
{
    if (p->tile_gen)
        launch_tile_v<FT, STAGE, true>(p, a, j, ev0, ev1);
    else
        launch_tile_v<FT, STAGE, false>(p, a, j, ev0, ev1);
}

// 3D tile kernels at the pool's threads per tile, in their lean variant
// unless a replica needs the general one.
template <typename FT, int STAGE, bool GEN>
void launch_tile3_v(gcr_pool* p, const PoolArgs& a, int j, hipEvent_t ev0, hipEvent_t ev1)
{
    if (p->nt_launch == 1024)
        launch_k(k_stage_tile3d<FT, STAGE, 1024, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 1024);
    else if (p->nt_launch == 512)
        launch_k(k_stage_tile3d<FT, STAGE, 512, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 512);
    else
        launch_k(k_stage_tile3d<FT, STAGE, 256, GEN>, p->ntile, p->stream, ev0, ev1, a, j, 256);
}

template <typename FT, int STAGE>
void launch_tile3(gcr_pool* p, const PoolArgs& a, int j, hipEvent_t ev0, hipEvent_t ev1)
{
    if (p->tile_gen)
        launch_tile3_v<FT, STAGE, true>(p, a, j, ev0, ev1);
    else
        launch_tile3_v<FT, STAGE, false>(p, a, j, ev0, ev1);
}

// One launch of a step: which = 0 the first force stage, 1 a later force
// stage (SRK stage `sub`: 1 and 2 exist in SRK4 only, 3 is the last), 2 the
// update, 3 the max-force reduction and dt.
void launch_stage(gcr_pool* p, int which, int j, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                  int sub = 3)
{
    PoolArgs a = p->a;
    a.sub = which == 0 ? 0 : sub;
    hipStream_t st = p->stream;
    const bool mixed = p->precision == GCR_MIXED;
    if (which == 3) {
        launch_k(k_finalize, p->capR, st, ev0, ev1, a, j);
    } else if (p->dim == 2) {
        if (which == 0)
            mixed ? launch_tile<float, 0>(p, a, j, ev0, ev1) : launch_tile<double, 0>(p, a, j, ev0, ev1);
        else if (which == 1)
            mixed ? launch_tile<float, 1>(p, a, j, ev0, ev1) : launch_tile<double, 1>(p, a, j, ev0, ev1);
        else
            launch_k(k_update<2>, p->nblk, st, ev0, ev1, a, j);
    } else {
        if (which == 0)
            mixed ? launch_tile3<float, 0>(p, a, j, ev0, ev1) : launch_tile3<double, 0>(p, a, j, ev0, ev1);
        else if (which == 1)
            mixed ? launch_tile3<float, 1>(p, a, j, ev0, ev1) : launch_tile3<double, 1>(p, a, j, ev0, ev1);
        else
            launch_k(k_update<3>, p->nblk, st, ev0, ev1, a, j);
    }
}

// Rigid-rod projection of a stage's update (only pools with rod replicas):
// stage 0 a virtual update after the force stage of SRK stage `sub`, stage 1
// the real update.
void launch_rod(gcr_pool* p, int stage, int j, int sub = 0)
{
    if (!p->nchb) return;
    PoolArgs a = p->a;
    a.sub = stage == 0 ? sub : 3;
    // one wavefront per chain: an entry of BS chains takes 64 blocks of 4 waves
    static_assert(BS == 256, "k_rod maps BS chains onto 64 blocks of 4 wavefronts");
    const int nb = p->nchb * 64;
    if (p->dim == 2)
        stage == 0 ? k_rod<2, 0><<<nb, BS, 0, p->stream>>>(a, j)
                   : k_rod<2, 1><<<nb, BS, 0, p->stream>>>(a, j);
    else
        stage == 0 ? k_rod<3, 0><<<nb, BS, 0, p->stream>>>(a, j)
                   : k_rod<3, 1><<<nb, BS, 0, p->stream>>>(a, j);
}

// Forces between the populations of groups, after a force stage.
void launch_group(gcr_pool* p, int stage, int j)
{
    for (int r : p->slide_reps) {  // the sliding partners, SLIDE_LANES blocks per bead block
        const RepParams& P = p->P[r];
        const int nb = P.nblk * SLIDE_LANES;
        if (p->dim == 2)
            stage == 0 ? k_slide_count<2, 0><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off)
                       : k_slide_count<2, 1><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off);
        else
            stage == 0 ? k_slide_count<3, 0><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off)
                       : k_slide_count<3, 1><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off);
    }
    const int nb = p->nblk * GROUP_LANES;
    if (p->dim == 2)
        stage == 0 ? k_group_stage<2, 0><<<nb, BS, 0, p->stream>>>(p->a, j)
                   : k_group_stage<2, 1><<<nb, BS, 0, p->stream>>>(p->a, j);
    else
        stage == 0 ? k_group_stage<3, 0><<<nb, BS, 0, p->stream>>>(p->a, j)
                   : k_group_stage<3, 1><<<nb, BS, 0, p->stream>>>(p->a, j);
}

// Motor bindings of a force stage (before the stage's second half).
// The motor rule, over the bead blocks of each motor population (MOTOR_LANES
// blocks per bead block).
void launch_motor(gcr_pool* p, int stage, int j)
{
    for (int r : p->motor_reps) {
        const RepParams& P = p->P[r];
        if (!p->live[r] || P.nblk == 0) continue;
        const int nb = P.nblk * MOTOR_LANES;
        if (p->dim == 2)
            stage == 0 ? k_motor<2, 0><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off)
                       : k_motor<2, 1><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off);
        else
            stage == 0 ? k_motor<3, 0><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off)
                       : k_motor<3, 1><<<nb, BS, 0, p->stream>>>(p->a, j, P.blk_off);
    }
}

// SRK4 virtual update of stage `sub` (1 or 2) for the replicas without rod.
void launch_virtual(gcr_pool* p, int j, int sub)
{
    PoolArgs a = p->a;
    a.sub = sub;
    if (p->dim == 2)
        k_virtual<2><<<p->nblk, BS, 0, p->stream>>>(a, j);
    else
        k_virtual<3><<<p->nblk, BS, 0, p->stream>>>(a, j);
}

void launch_begin(gcr_pool* p)
{
    if (p->dim == 2)
        k_batch_begin<2><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
    else
        k_batch_begin<3><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
}

// One batch of B steps with a re-sort every K steps. The safe batch (K = 1,
// the replay of a batch that left the skin) also re-sorts a replica on its
// virtual positions when those left the skin, so every force evaluation sees
// its complete pair set whatever a step's displacements. Pools with SRK4
// replicas run the two inner SRK4 stages between the first and the last
// force stage; the SRK2 replicas skip them.
void launch_batch(gcr_pool* p, int K, bool safe)
{
    p->a.safe = safe ? 1 : 0;
    launch_begin(p);
    for (int j = 0; j < p->B; ++j) {
        if (j % K == 0) launch_rebuild(p, j);
        launch_stage(p, 0, j);
        if (p->has_motor) launch_motor(p, 0, j);
        if (p->has_group) launch_group(p, 0, j);
        launch_rod(p, 0, j, 0);
        if (safe) launch_rebuild(p, j, true, 0);
        for (int sub = 1; p->has_srk4 && sub <= 2; ++sub) {
            launch_stage(p, 1, j, nullptr, nullptr, sub);
            launch_virtual(p, j, sub);
            launch_rod(p, 0, j, sub);
            if (safe) launch_rebuild(p, j, true, sub);
        }
        launch_stage(p, 1, j);
        if (p->has_motor) launch_motor(p, 1, j);
        if (p->has_group) launch_group(p, 1, j);
        launch_stage(p, 3, j);
        if (p->has_group) k_group_dt<<<(p->capR + 63) / 64, 64, 0, p->stream>>>(p->a, j);
        launch_rod(p, 1, j);
        launch_stage(p, 2, j);
    }
}

double now_seconds()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// Captures the batch graph (K, safe) and makes *exec run it: an existing
// executable graph is updated in place when the new capture has its shape
// (replicas admitted or retired change launch sizes and arguments, not the
// sequence of kernels), else a new one is instantiated.
int capture(gcr_pool* p, int K, bool safe, hipGraphExec_t* exec)
{
    const double t0 = now_seconds();
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    launch_batch(p, K, safe);
    p->a.safe = 0;
    HIPCHK(hipStreamEndCapture(p->stream, &g));
    ++p->n_capture;
    bool updated = false;
    if (*exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        const hipError_t e = hipGraphExecUpdate(*exec, g, &bad, &res);
        updated = e == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();  // a refused update is not an error of the pool
            (void)hipGraphExecDestroy(*exec);
            *exec = nullptr;
        }
    }
    if (updated) {
        ++p->n_update;
    } else {
        const hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(GCR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        ++p->n_instantiate;
    }
    HIPCHK(hipGraphDestroy(g));
    p->graph_seconds += now_seconds() - t0;
    return GCR_OK;
}

// The normal batch graph, up to date with the tables.
int ensure_ready(gcr_pool* p)
{
    int rc;
    if (p->tables_dirty && (rc = rebuild_tables(p)) != GCR_OK) return rc;
    if ((!p->g_norm || p->norm_stale) && (rc = capture(p, p->K, false, &p->g_norm)) != GCR_OK) return rc;
    p->norm_stale = false;
    return GCR_OK;
}

// The safe batch graph, captured when a replay or a safe batch first needs it.
int ensure_safe(gcr_pool* p)
{
    int rc;
    if (p->tables_dirty && (rc = rebuild_tables(p)) != GCR_OK) return rc;
    if ((!p->g_safe || p->safe_stale) && (rc = capture(p, 1, true, &p->g_safe)) != GCR_OK) return rc;
    p->safe_stale = false;
    return GCR_OK;
}

bool running(const RepState& s)
{
    if (s.frozen || s.fixed) return false;
    bool a = s.max_steps < 0 || s.steps < s.max_steps;
    bool b = s.mdu < 0 || s.elapsed < s.mdu;
    return a && b;
}

int count_running(gcr_pool* p)
{
    int n = 0;
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && running(p->S[r])) ++n;
    return n;
}

// Run the safe (rebuild every step) batch for the replicas that exceeded
// their skin, after restoring them to the start of the batch.
int replay(gcr_pool* p)
{
    std::vector<int> which;
    // populations of a group move together: one member's violation replays
    // the whole group
    for (const auto& G : p->groups) {
        bool any = false;
        for (int r : G.rids) any = any || p->S[r].violation;
        if (any)
            for (size_t k = 0; k < G.rids.size(); ++k)
                if (!G.fixed[k]) p->S[G.rids[k]].violation = 1;
    }
    for (int r = 0; r < p->capR; ++r) {
        if (!p->live[r]) continue;
        if (p->S[r].fixed) continue;
        if (p->S[r].violation) {
            which.push_back(r);
            p->S[r].frozen = 2;
        } else {
            p->S[r].frozen = 1;
        }
    }
    int rc = ensure_safe(p);
    if (rc) return rc;
    if ((rc = push_states(p))) return rc;
    if (p->dim == 2)
        k_restore<2><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
    else
        k_restore<3><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
    k_restore_state<<<(p->capR + 63) / 64, 64, 0, p->stream>>>(p->a);
    HIPCHK(hipGraphLaunch(p->g_safe, p->stream));
    if ((rc = pull_states(p))) return rc;
    for (int r : which) {
        p->replays[r] += 1;
        // a second replay in a row: the next batches start in the safe form
        if (p->replayed[r]) p->safe_left[r] = SAFE_STICKY;
        p->replayed[r] = 1;
        // the safe batch re-sorts wherever a step left the skin, so this is
        // not expected; if it happens the pair set may have been incomplete,
        // and the replica stops here, marked failed, while the others go on
        if (p->S[r].violation) p->S[r].frozen = FROZEN_FAILED;
    }
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && p->S[r].frozen == 1) p->S[r].frozen = 0;
    return push_states(p);
}

int with_only(gcr_pool* p, int rid, bool on)
{
    // freeze every other live replica (diagnostics must not re-sort them)
    int rc = pull_states(p);
    if (rc) return rc;
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && r != rid && p->S[r].frozen != FROZEN_FAILED) p->S[r].frozen = on ? 1 : 0;
    return push_states(p);
}

int check_rid(gcr_pool* p, int rid)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    if (rid < 0 || rid >= p->capR || !p->live[rid]) return fail(GCR_ERR_ARG, "no such replica");
    return GCR_OK;
}

int download_one(gcr_pool* p, int rid, const double* src, double* host, int width)
{
    const RepParams& P = p->P[rid];
    k_unpermute<<<(P.n + BS - 1) / BS, BS, 0, p->stream>>>(src, p->Xn, p->SLOT, P.bead_off, P.n, width);
    HIPCHK(hipMemcpyAsync(host, p->Xn, sizeof(double) * (size_t)P.n * width, hipMemcpyDeviceToHost,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int upload_perm(gcr_pool* p, int rid, const double* host, double* dst)
{
    const RepParams& P = p->P[rid];
    HIPCHK(hipMemcpyAsync(p->Xn, host, sizeof(double) * (size_t)P.n * p->dim, hipMemcpyHostToDevice,
                          p->stream));
    k_permute_in<<<(P.n + BS - 1) / BS, BS, 0, p->stream>>>(p->Xn, dst, p->SLOT, P.bead_off, P.n, p->dim);
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

}  // namespace

// =================================================================== C-ABI

extern "C" {

const char* gcr_last_error(void) { return g_err.c_str(); }

// Grid-wide barriers of a cooperative launch, `rounds` of them back to back.
__global__ void __launch_bounds__(BS) k_grid_barrier_probe(int rounds, int* sink)
{
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    for (int i = 0; i < rounds; ++i) g.sync();
    if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] = rounds;
}

int gcr_grid_barrier_us(int32_t blocks_per_cu, int32_t rounds, double out[2])
{
    if (!out || blocks_per_cu < 1 || rounds < 1) return fail(GCR_ERR_ARG, "need blocks_per_cu >= 1, rounds >= 1");
    int dev = 0;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (!prop.cooperativeLaunch) return fail(GCR_ERR_STATE, "the device has no cooperative launch");
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_grid_barrier_probe, BS, 0));
    const int blocks = std::min(blocks_per_cu, per_cu) * prop.multiProcessorCount;
    int* sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GCR_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == GCR_OK) rc = fail(GCR_ERR_HIP, hipGetErrorString(e));
        return rc == GCR_OK;
    };
    float ms[2] = {0.f, 0.f};
    if (ok(hipMalloc((void**)&sink, sizeof(int))) && ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
        ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1))) {
        // a warm-up launch, then the launch without barriers and with `rounds`
        const int counts[3] = {rounds, 0, rounds};
        for (int k = 0; k < 3 && rc == GCR_OK; ++k) {
            int n = counts[k];
            void* args[] = {&n, &sink};
            if (ok(hipEventRecord(e0, st)) &&
                ok(hipLaunchCooperativeKernel((const void*)k_grid_barrier_probe, dim3(blocks), dim3(BS), args, 0,
                                              st)) &&
                ok(hipEventRecord(e1, st)) && ok(hipEventSynchronize(e1)) && k > 0)
                ok(hipEventElapsedTime(&ms[k - 1], e0, e1));
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (sink) (void)hipFree(sink);
    if (rc) return rc;
    out[0] = 1e3 * (double)(ms[1] - ms[0]) / rounds;
    out[1] = (double)blocks;
    return GCR_OK;
}

int gcr_device_info(char* name, int32_t name_len, int32_t* n_cu, int64_t* hbm_bytes)
{
    int dev = 0;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return GCR_OK;
}

int64_t gcr_replica_bytes(int32_t dim, int64_t n, int64_t n_chains, int64_t n_cells)
{
    // per bead: 10 dim-vectors + noise (2 x 2 dim) of doubles, 5 int pairs, links,
    // cell, slot and two where_active ints; the neighbour list, its length and
    // the re-sort positions; per chain: 4 doubles + 3 ints; per cell: 2 ints
    const int64_t nl = 2 * (dim == 3 ? NL3_MAX : NL_MAX) + 4 + 8 * dim;
    return n * (8 * (10 * dim + 4 * dim) + 4 * (5 * 2 + 4 + 2 + 2) + nl) + n_chains * (8 * 4 + 12) +
           (n_cells + 1) * 8;
}

int64_t gcr_replica_cells(int32_t dim, const gcr_replica_desc* d)
{
    if (!d) return -1;
    const int64_t nc = grid_side(dim, d);
    return dim == 3 ? nc * nc * nc : nc * nc;
}

int gcr_pool_create(int32_t dim, int32_t precision, int64_t capB, int64_t capC, int64_t capCh,
                    int32_t capR, gcr_pool** out)
{
    if (!out) return fail(GCR_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(GCR_ERR_ARG, "dim must be 2 or 3");
    if (precision != GCR_FP64 && precision != GCR_MIXED) return fail(GCR_ERR_ARG, "bad precision");
    if (capB <= 0 || capC <= 0 || capCh <= 0 || capR <= 0 || capB > (1ll << 30) || capC > (1ll << 30))
        return fail(GCR_ERR_ARG, "capacities must be positive and below 2^30");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GCR_ERR_HIP, "no HIP device");
    gcr_pool* p = new gcr_pool();
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            p->n_cu = std::max(1, prop.multiProcessorCount);
    }
    p->dim = dim;
    p->precision = precision;
    p->capB = capB;
    p->capC = capC;
    p->capCh = capCh;
    p->capR = capR;
    p->segB.init(capB);
    p->segC.init(capC);
    p->segCh.init(capCh);
    p->P.assign(capR, RepParams{});
    if (hipHostMalloc((void**)&p->S, sizeof(RepState) * capR, hipHostMallocDefault) != hipSuccess) {
        p->S = nullptr;
        gcr_pool_destroy(p);
        return fail(GCR_ERR_HIP, "page-locked host memory for the replica states");
    }
    for (int r = 0; r < capR; ++r) {
        p->S[r] = RepState{};
        p->S[r].frozen = 1;
    }
    p->live.assign(capR, 0);
    p->replays.assign(capR, 0);
    p->safe_left.assign(capR, 0);
    p->replayed.assign(capR, 0);
    p->xi_dev.assign(capR, nullptr);
    p->ser_dev.assign(capR, nullptr);
    p->ser_rows.assign(capR, 0);
    p->capBlk = capB / BS + capR + 1;
    p->capChk = capC / SCAN_CHUNK + capR + 1;
    p->capTile = capC + (int64_t)capR + 1;  // tiles of one cell at worst
    p->capChb = capCh / BS + capR + 1;
    int rc = GCR_OK;
    auto A = [&](int r) {
        if (r != GCR_OK && rc == GCR_OK) rc = r;
    };
    const size_t nb = (size_t)capB;
    A(dalloc(&p->blk_rep, p->capBlk));
    A(dalloc(&p->blk_slot0, p->capBlk));
    A(dalloc(&p->chk_rep, p->capChk));
    A(dalloc(&p->chk_cell0, p->capChk));
    A(dalloc(&p->rep_chk_off, capR + 1));
    A(dalloc(&p->chb_rep, p->capChb));
    A(dalloc(&p->chb_chain0, p->capChb));
    A(dalloc(&p->dP, capR));
    A(dalloc(&p->dS, capR));
    A(dalloc(&p->dSs, capR));
    for (double** q : {&p->X, &p->X1, &p->F0, &p->FT, &p->T, &p->Xn, &p->Tn, &p->Xs, &p->Ts, &p->X1n})
        A(dalloc(q, nb * dim));
    A(dalloc(&p->XI, nb * 2 * dim));
    A(dalloc(&p->XIn, nb * 2 * dim));
    for (int32_t** q : {&p->META, &p->KEYOFF, &p->PAIRS, &p->METAn, &p->METAs}) A(dalloc(q, nb * 2));
    A(dalloc(&p->LINK, nb * 4));
    A(dalloc(&p->CELL, nb));
    A(dalloc(&p->SLOT, nb));
    A(dalloc(&p->CHAIN_A, capCh));
    A(dalloc(&p->CHAIN_CS, capCh * 3));
    A(dalloc(&p->CHAIN_FIRST, capCh));
    A(dalloc(&p->CHAIN_LEN, capCh));
    A(dalloc(&p->CCOUNT, capC + 1));
    A(dalloc(&p->CSTART, capC + 1));
    A(dalloc(&p->CHKSUM, p->capChk));
    // tile maxima of the force stages, then those of k_group_stage's blocks
    // (GROUP_LANES per bead block)
    A(dalloc(&p->BMAX, p->capTile + GROUP_LANES * p->capBlk));
    p->capX = (int64_t)capR * GCR_MAX_POPS;
    A(dalloc(&p->XPAIRS, p->capX));
    A(dalloc(&p->GRP, capR));
    A(dalloc(&p->CHAIN_CHUNK, capCh));
    A(dalloc(&p->WA, nb));
    A(dalloc(&p->WAn, nb));
    A(dalloc(&p->MC, nb));
    A(dalloc(&p->SCNT, nb));
    A(dalloc(&p->TILES, p->capTile * 4));
    A(dalloc(&p->OBS, 5 * (capCh / BS + 2)));
    A(dalloc(&p->COMW, 3 * capCh));
    A(dalloc(&p->COMU, 3 * capCh));
    // neighbour lists: u16 LDS indices (NL_MAX, 128 B per bead in 2D;
    // NL3_MAX, 256 B per bead in 3D)
    A(dalloc(&p->NL, nb * (dim == 3 ? NL3_MAX : NL_MAX)));
    A(dalloc(&p->NLEN, nb));
    A(dalloc(&p->XR, nb * dim));
    if (rc == GCR_OK && hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(GCR_ERR_HIP, "stream creation failed");
    if (rc == GCR_OK && hipMemset(p->CCOUNT, 0, sizeof(int32_t) * (capC + 1)) != hipSuccess)
        rc = fail(GCR_ERR_HIP, "memset failed");
    if (rc == GCR_OK && hipMemset(p->X1, 0, sizeof(double) * nb * dim) != hipSuccess)
        rc = fail(GCR_ERR_HIP, "memset failed");
    if (rc == GCR_OK && hipMemset(p->FT, 0, sizeof(double) * nb * dim) != hipSuccess)
        rc = fail(GCR_ERR_HIP, "memset failed");
    if (rc == GCR_OK && hipMemcpy(p->dS, p->S, sizeof(RepState) * capR, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(GCR_ERR_HIP, "state upload failed");
    if (rc != GCR_OK) {
        gcr_pool_destroy(p);
        return rc;
    }
    PoolArgs& a = p->a;
    a.blk_rep = p->blk_rep;
    a.blk_slot0 = p->blk_slot0;
    a.chk_rep = p->chk_rep;
    a.chk_cell0 = p->chk_cell0;
    a.rep_chk_off = p->rep_chk_off;
    a.chb_rep = p->chb_rep;
    a.chb_chain0 = p->chb_chain0;
    a.nrep = capR;
    a.P = p->dP;
    a.S = p->dS;
    a.X = p->X;
    a.X1 = p->X1;
    a.F0 = p->F0;
    a.FT = p->FT;
    a.XI = p->XI;
    a.T = p->T;
    a.META = p->META;
    a.LINK = p->LINK;
    a.CELL = p->CELL;
    a.SLOT = p->SLOT;
    a.CHAIN_A = p->CHAIN_A;
    a.CHAIN_CS = p->CHAIN_CS;
    a.CHAIN_FIRST = p->CHAIN_FIRST;
    a.CHAIN_LEN = p->CHAIN_LEN;
    a.CCOUNT = p->CCOUNT;
    a.CSTART = p->CSTART;
    a.CHKSUM = p->CHKSUM;
    a.BMAX = p->BMAX;
    a.SCNT = p->SCNT;
    a.TILES = p->TILES;
    a.NL = p->NL;
    a.NLEN = p->NLEN;
    a.XR = p->XR;
    a.nl_stride = (int64_t)capB;
    a.KEYOFF = p->KEYOFF;
    a.PAIRS = p->PAIRS;
    a.Xn = p->Xn;
    a.Tn = p->Tn;
    a.METAn = p->METAn;
    a.XIn = p->XIn;
    a.XIRn = p->XIn + nb * dim;
    a.XIR = p->XI + nb * dim;
    a.X1n = p->X1n;
    a.Xs = p->Xs;
    a.Ts = p->Ts;
    a.METAs = p->METAs;
    a.Ss = p->dSs;
    a.XPAIRS = p->XPAIRS;
    a.GRP = p->GRP;
    a.CHAIN_CHUNK = p->CHAIN_CHUNK;
    a.WA = p->WA;
    a.WAn = p->WAn;
    a.MC = p->MC;
    *out = p;
    return GCR_OK;
}

int gcr_pool_set_schedule(gcr_pool* p, int32_t steps_per_batch, int32_t rebuild_every)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    if (steps_per_batch < 1 || rebuild_every < 1 || steps_per_batch > 4096)
        return fail(GCR_ERR_ARG, "steps_per_batch in [1, 4096] and rebuild_every >= 1 required");
    if (rebuild_every > steps_per_batch) rebuild_every = steps_per_batch;
    p->K = rebuild_every;
    p->B = ((steps_per_batch + rebuild_every - 1) / rebuild_every) * rebuild_every;
    drop_graphs(p);
    return GCR_OK;
}

int gcr_replica_add(gcr_pool* p, const gcr_replica_desc* d, const double* pos, const double* tangent,
                    const int64_t* chain_id, const double* rest_lengths, const double* v_tan_angle,
                    const double* v0_per_chain, int32_t* rid)
{
    if (!p || !d || !pos || !chain_id || !rid) return fail(GCR_ERR_ARG, "null argument");
    const int dim = p->dim;
    const int64_t n = d->n_beads, nch = d->n_chains;
    if (n < 1 || nch < 1 || nch > n) return fail(GCR_ERR_ARG, "need 1 <= n_chains <= n_beads");
    if (!(d->L > 0) || !(d->radius > 0)) return fail(GCR_ERR_ARG, "L and radius must be positive");
    if (d->has_wca && !(d->wca[3] > 0 && 2 * d->wca[3] < d->L))
        return fail(GCR_ERR_ARG, "WCA cutoff must be in (0, L/2)");
    if (d->D < 0 || d->Dr < 0) return fail(GCR_ERR_ARG, "D and Dr must be >= 0 (resolve defaults first)");
    if (!srk_order_ok(d)) return fail(GCR_ERR_ARG, "srk_order must be 2 or 4 (2 with active sliding)");
    // chains: runs of equal chain_id with ids 0..nch-1 in order of first appearance
    std::vector<int32_t> first(nch, -1), len(nch, 0);
    int64_t prev = -1;
    for (int64_t b = 0; b < n; ++b) {
        int64_t c = chain_id[b];
        if (c < 0 || c >= nch) return fail(GCR_ERR_ARG, "chain_id out of [0, n_chains)");
        if (c != prev) {
            if (first[c] >= 0) return fail(GCR_ERR_ARG, "chains must be contiguous runs of chain_id");
            first[c] = (int32_t)b;
        }
        len[c] += 1;
        prev = c;
    }
    for (int64_t c = 0; c < nch; ++c)
        if (first[c] < 0) return fail(GCR_ERR_ARG, "every chain id needs at least one bead");
    for (int64_t b = 0; b < n * dim; ++b)
        if (!(pos[b] >= 0.0 && pos[b] <= d->L)) return fail(GCR_ERR_ARG, "positions must lie in [0, L]");
    const int64_t nc = grid_side(dim, d);
    const int64_t ncell = dim == 3 ? nc * nc * nc : nc * nc;
    if (ncell > (1 << 28)) return fail(GCR_ERR_ARG, "cell grid too large");
    int r = -1;
    for (int i = 0; i < p->capR; ++i)
        if (!p->live[i]) {
            r = i;
            break;
        }
    if (r < 0) return fail(GCR_ERR_CAPACITY, "replica capacity exhausted");
    int64_t b0 = p->segB.alloc(n);
    if (b0 < 0) return fail(GCR_ERR_CAPACITY, "bead capacity exhausted");
    int64_t c0 = p->segC.alloc(ncell + 1);
    if (c0 < 0) {
        p->segB.release(b0, n);
        return fail(GCR_ERR_CAPACITY, "cell capacity exhausted");
    }
    int64_t h0 = p->segCh.alloc(nch);
    if (h0 < 0) {
        p->segB.release(b0, n);
        p->segC.release(c0, ncell + 1);
        return fail(GCR_ERR_CAPACITY, "chain capacity exhausted");
    }
    RepParams& P = p->P[r];
    P = RepParams{};
    P.n = (int32_t)n;
    P.bead_off = (int32_t)b0;
    P.chain_off = (int32_t)h0;
    P.n_chains = (int32_t)nch;
    P.cell_off = (int32_t)c0;
    P.nc = (int32_t)nc;
    P.ncell = (int32_t)ncell;
    P.L = d->L;
    P.halfL = d->L / 2.0;
    P.h = d->L / (double)nc;
    P.inv_h = (double)nc / d->L;
    P.rl2 = P.h * P.h * (1.0 + 1e-5);  // list radius rc + 2 skin_half = h
    P.seed = d->seed;
    set_physics(P, d, *std::max_element(len.begin(), len.end()));
    // host staging of per-bead and per-chain data
    std::vector<int32_t> meta(2 * n), slot(n);
    for (int64_t b = 0; b < n; ++b) {
        meta[2 * b] = (int32_t)(b0 + b);
        meta[2 * b + 1] = (int32_t)(h0 + chain_id[b]);
        slot[b] = (int32_t)(b0 + b);
    }
    std::vector<double> ca(nch), cs(3 * nch);
    std::vector<int32_t> cf(nch), cl(nch);
    for (int64_t c = 0; c < nch; ++c) {
        ca[c] = rest_lengths ? rest_lengths[c] : 0.0;
        double ang = v_tan_angle ? v_tan_angle[c] : 0.0;
        cs[3 * c] = v0_per_chain ? v0_per_chain[c] : d->v0;
        if (ang == 0.0 || dim != 2) {  // forces.py:214-217: plain v0 * t
            cs[3 * c + 1] = 1.0;
            cs[3 * c + 2] = 0.0;
        } else {
            cs[3 * c + 1] = std::cos(ang);
            cs[3 * c + 2] = std::sin(ang);
        }
        cf[c] = (int32_t)(b0 + first[c]);
        cl[c] = len[c];
    }
    std::vector<double> tang(n * dim, 0.0);
    if (tangent)
        std::memcpy(tang.data(), tangent, sizeof(double) * n * dim);
    hipStream_t st = p->stream;
    HIPCHK(hipMemcpyAsync(p->X + b0 * dim, pos, sizeof(double) * n * dim, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->X1 + b0 * dim, pos, sizeof(double) * n * dim, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->T + b0 * dim, tang.data(), sizeof(double) * n * dim, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->META + 2 * b0, meta.data(), 8 * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->SLOT + b0, slot.data(), 4 * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->CHAIN_A + h0, ca.data(), 8 * nch, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->CHAIN_CS + 3 * h0, cs.data(), 24 * nch, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->CHAIN_FIRST + h0, cf.data(), 4 * nch, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->CHAIN_LEN + h0, cl.data(), 4 * nch, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(p->FT + b0 * dim, 0, sizeof(double) * n * dim, st));
    HIPCHK(hipMemsetAsync(p->CHAIN_CHUNK + h0, 0, 4 * nch, st));
    k_fill_i32<<<(int)((n + BS - 1) / BS), BS, 0, st>>>(p->WA + b0, (int)n, -1);
    RepState& S = p->S[r];
    S = RepState{};
    S.max_steps = 0;  // not armed
    S.mdu = -1.0;
    S.frozen = 0;
    p->live[r] = 1;
    p->replays[r] = 0;
    p->safe_left[r] = 0;
    p->replayed[r] = 0;
    HIPCHK(hipStreamSynchronize(st));
    int rc = upload_rep(p, r);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    p->tables_dirty = true;
    // first cell sort of the new replica, so links, cells and the layout are
    // valid before its first step (other replicas are left untouched)
    if ((rc = rebuild_tables(p))) return rc;
    if ((rc = with_only(p, r, true))) return rc;
    launch_rebuild(p, -1);
    HIPCHK(hipGetLastError());
    if ((rc = with_only(p, r, false))) return rc;
    *rid = r;
    return GCR_OK;
}

// Members of group g become independent replicas again.
void dissolve_group(gcr_pool* p, int g)
{
    for (int r : p->groups[g].rids) {
        p->P[r].flags &= ~GCR_F_GROUP;
        p->P[r].group = -1;
        p->S[r].fixed = 0;
    }
    p->groups.erase(p->groups.begin() + g);
    p->tables_dirty = true;
}

int group_of(const gcr_pool* p, int rid)
{
    for (int g = 0; g < (int)p->groups.size(); ++g)
        for (int r : p->groups[g].rids)
            if (r == rid) return g;
    return -1;
}

int gcr_pool_group(gcr_pool* p, const int32_t* rids, int32_t n, const double* cross,
                   const int32_t* fixed)
{
    if (!p || !rids) return fail(GCR_ERR_ARG, "null argument");
    if (n < 2 || n > GCR_MAX_POPS) return fail(GCR_ERR_ARG, "a group has 2 to GCR_MAX_POPS populations");
    gcr_pool::Group G;
    int moving = 0;
    for (int k = 0; k < n; ++k) {
        int rc = check_rid(p, rids[k]);
        if (rc) return rc;
        for (int q = 0; q < k; ++q)
            if (rids[q] == rids[k]) return fail(GCR_ERR_ARG, "a replica appears twice in the group");
        if (group_of(p, rids[k]) >= 0) return fail(GCR_ERR_ARG, "replica already in a group");
        const RepParams& P = p->P[rids[k]];
        if (P.L != p->P[rids[0]].L) return fail(GCR_ERR_ARG, "the populations of a group share one box");
        if (P.flags & (GCR_F_SRK4 | GCR_F_ROD))
            return fail(GCR_ERR_ARG, "groups run SRK_order 2 without the rigid rod");
        G.rids.push_back(rids[k]);
        G.fixed.push_back(fixed && fixed[k] ? 1 : 0);
        moving += G.fixed.back() ? 0 : 1;
    }
    if (!moving) return fail(GCR_ERR_ARG, "a group needs a moving population");
    G.cross.assign((size_t)n * n * 5, 0.0);
    if (cross) std::memcpy(G.cross.data(), cross, sizeof(double) * G.cross.size());
    for (int b = 0; b < n; ++b) {  // motors: a motor population, one pushed population
        int pushed = 0;
        for (int a2 = 0; a2 < n; ++a2) {
            const double kind = G.cross[((size_t)a2 * n + b) * 5];
            if (kind == 2.0) {
                if (!(p->P[rids[b]].flags & GCR_F_MOTOR))
                    return fail(GCR_ERR_ARG, "kind 2 needs a motor population as the source");
                if (G.fixed[a2]) return fail(GCR_ERR_ARG, "motors push a moving population");
                ++pushed;
            } else if (kind != 0.0 && kind != 1.0) {
                return fail(GCR_ERR_ARG, "cross kinds are 0, 1 (law) or 2 (motors)");
            }
        }
        if (pushed > 1) return fail(GCR_ERR_ARG, "a motor population pushes one population");
    }
    int rc = pull_states(p);
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        RepParams& P = p->P[rids[k]];
        if (G.fixed[k]) {
            P.flags &= ~GCR_F_GROUP;
            p->S[rids[k]].fixed = 1;
        } else {
            P.flags |= GCR_F_GROUP;
            p->S[rids[k]].fixed = 0;
        }
    }
    p->groups.push_back(G);
    p->tables_dirty = true;
    if ((rc = push_states(p))) return rc;
    return rebuild_tables(p);
}

int gcr_replica_remove(gcr_pool* p, int32_t rid)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    const int g = group_of(p, rid);
    if (g >= 0) {
        if ((rc = pull_states(p))) return rc;
        dissolve_group(p, g);
        if ((rc = push_states(p))) return rc;
    }
    RepParams& P = p->P[rid];
    p->segB.release(P.bead_off, P.n);
    p->segC.release(P.cell_off, P.ncell + 1);
    p->segCh.release(P.chain_off, P.n_chains);
    if (p->xi_dev[rid]) {
        (void)hipFree(p->xi_dev[rid]);
        p->xi_dev[rid] = nullptr;
    }
    if (p->ser_dev[rid]) {
        (void)hipFree(p->ser_dev[rid]);
        p->ser_dev[rid] = nullptr;
    }
    p->ser_rows[rid] = 0;
    p->live[rid] = 0;
    p->S[rid] = RepState{};
    p->S[rid].frozen = 1;
    P = RepParams{};
    rc = upload_rep(p, rid);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));
    p->tables_dirty = true;
    return GCR_OK;
}

int gcr_replica_set_physics(gcr_pool* p, int32_t rid, const gcr_replica_desc* d, const double* rest_lengths,
                            const double* v0_per_chain)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (!d) return fail(GCR_ERR_ARG, "null descriptor");
    if (!srk_order_ok(d)) return fail(GCR_ERR_ARG, "srk_order must be 2 or 4 (2 with active sliding)");
    const int grp = group_of(p, rid);
    if (grp >= 0 && (d->srk_order == 4 || (d->rod != 0.0)))
        return fail(GCR_ERR_ARG, "groups run SRK_order 2 without the rigid rod");
    RepParams& P = p->P[rid];
    if (d->n_beads != P.n || d->n_chains != P.n_chains || d->L != P.L)
        return fail(GCR_ERR_ARG, "beads, chains and box must stay the same");
    if (d->has_wca && !(d->wca[3] > 0 && d->wca[3] < P.h))
        return fail(GCR_ERR_ARG, "cutoff does not fit the replica's cell grid (re-create the replica)");
    // the longest chain, from the chain lengths on the device
    std::vector<int32_t> cl(P.n_chains);
    HIPCHK(hipMemcpy(cl.data(), p->CHAIN_LEN + P.chain_off, 4 * cl.size(), hipMemcpyDeviceToHost));
    const int longest = cl.empty() ? 0 : *std::max_element(cl.begin(), cl.end());
    const uint32_t old = P.flags;
    set_physics(P, d, longest);
    if (rest_lengths)
        HIPCHK(hipMemcpyAsync(p->CHAIN_A + P.chain_off, rest_lengths, 8 * (size_t)P.n_chains,
                              hipMemcpyHostToDevice, p->stream));
    if (!v0_per_chain) {  // d->v0 for every chain (as gcr_replica_add), set on the device
        if (P.n_chains > 0)
            k_set_speed<<<(P.n_chains + BS - 1) / BS, BS, 0, p->stream>>>(
                p->CHAIN_CS + 3 * (size_t)P.chain_off, P.n_chains, d->v0);
    } else {  // speeds per chain, beside the chains' angle terms
        std::vector<double> cs(3 * (size_t)P.n_chains);
        HIPCHK(hipMemcpy(cs.data(), p->CHAIN_CS + 3 * (size_t)P.chain_off, 8 * cs.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < P.n_chains; ++c) cs[3 * c] = v0_per_chain[c];
        HIPCHK(hipMemcpyAsync(p->CHAIN_CS + 3 * (size_t)P.chain_off, cs.data(), 8 * cs.size(),
                              hipMemcpyHostToDevice, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    // the parameters only: the device state (dt, counters) stays as it is
    HIPCHK(hipMemcpyAsync(p->dP + rid, &P, sizeof(RepParams), hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    // rod, pair and SRK4 flags change the launch tables or the stages of a
    // step; lists are rebuilt at the start of every batch, so a new exclusion
    // rule takes effect there
    if ((old ^ P.flags) & (GCR_F_ROD | GCR_F_WCA | GCR_F_SRK4 | GCR_F_WALLS | GCR_F_FENE | GCR_F_SLIDE))
        p->tables_dirty = true;
    if (grp >= 0) p->tables_dirty = true;  // the scan windows follow the skin
    return GCR_OK;
}

int gcr_replica_set_chunks(gcr_pool* p, int32_t rid, const int32_t* chain_chunk)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    const RepParams& P = p->P[rid];
    if (chain_chunk)
        HIPCHK(hipMemcpyAsync(p->CHAIN_CHUNK + P.chain_off, chain_chunk, 4 * (size_t)P.n_chains,
                              hipMemcpyHostToDevice, p->stream));
    else
        HIPCHK(hipMemsetAsync(p->CHAIN_CHUNK + P.chain_off, 0, 4 * (size_t)P.n_chains, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int gcr_replica_where_active(gcr_pool* p, int32_t rid, int32_t* out)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (!out) return fail(GCR_ERR_ARG, "null output");
    const RepParams& P = p->P[rid];
    int32_t* tmp = reinterpret_cast<int32_t*>(p->Xn);  // staging, free between calls
    k_unpermute_i32<<<(P.n + BS - 1) / BS, BS, 0, p->stream>>>(p->WA, tmp, p->SLOT, P.bead_off, P.n);
    HIPCHK(hipMemcpyAsync(out, tmp, 4 * (size_t)P.n, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int gcr_replica_upload(gcr_pool* p, int32_t rid, const double* pos, const double* tangent)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    const RepParams& P = p->P[rid];
    if (pos) {
        for (int64_t b = 0; b < (int64_t)P.n * p->dim; ++b)
            if (!(pos[b] >= 0.0 && pos[b] <= P.L)) return fail(GCR_ERR_ARG, "positions must lie in [0, L]");
        if ((rc = upload_perm(p, rid, pos, p->X))) return rc;
    }
    if (tangent && (rc = upload_perm(p, rid, tangent, p->T))) return rc;
    return GCR_OK;
}

int gcr_replica_begin(gcr_pool* p, int32_t rid, const double tv[8], const double* xi, int64_t xi_steps)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (!tv) return fail(GCR_ERR_ARG, "null time_vals");
    RepParams& P = p->P[rid];
    RepState& S = p->S[rid];
    if (!(tv[1] > 0) || !(tv[2] > 0) || !(tv[3] > 0)) return fail(GCR_ERR_ARG, "dtmax, max_dist_per_step, RK_cheat_factor must be > 0");
    if (tv[7] < 0 && tv[4] < 0) return fail(GCR_ERR_ARG, "need max_steps >= 0 or max_dt_usable > 0");
    P.dtmax = tv[1];
    P.mdps = tv[2];
    P.cheat = tv[3];
    if (p->xi_dev[rid]) {
        HIPCHK(hipFree(p->xi_dev[rid]));
        p->xi_dev[rid] = nullptr;
    }
    P.flags &= ~GCR_F_INJECT;
    P.xi_inject = nullptr;
    P.xi_steps = 0;
    if (xi) {
        if (tv[7] < 0 || xi_steps < (int64_t)tv[7])
            return fail(GCR_ERR_ARG, "injected noise must cover max_steps steps");
        size_t bytes = sizeof(double) * (size_t)xi_steps * P.n * 2 * p->dim;
        HIPCHK(hipMalloc((void**)&p->xi_dev[rid], bytes));
        HIPCHK(hipMemcpy(p->xi_dev[rid], xi, bytes, hipMemcpyHostToDevice));
        P.flags |= GCR_F_INJECT;
        P.xi_inject = p->xi_dev[rid];
        P.xi_steps = xi_steps;
    }
    if ((rc = pull_states(p))) return rc;
    S.dt = tv[0];
    S.dt_used = 0.0;
    S.elapsed = 0.0;
    S.steps = 0;
    S.batch_base = 0;
    S.max_steps = tv[7] < 0 ? -1 : (int64_t)tv[7];
    S.mdu = tv[4];
    S.violation = 0;
    S.frozen = 0;
    rc = upload_rep(p, rid);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int gcr_pool_run(gcr_pool* p, int32_t max_batches, int32_t* n_running)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    int rc = ensure_ready(p);
    if (rc) return rc;
    if ((rc = pull_states(p))) return rc;
    for (int b = 0; max_batches <= 0 || b < max_batches; ++b) {
        if (count_running(p) == 0) break;
        // every running replica had a batch replayed lately (a growth phase
        // does, chunk after chunk): start in the safe form at once, which
        // the normal batch would most likely need again. Its snapshot and
        // per-step re-sorts are those of a replay; a violation there stops
        // the replica as after a replay.
        int nrun = 0, nsafe = 0;
        for (int r = 0; r < p->capR; ++r)
            if (p->live[r] && running(p->S[r])) {
                ++nrun;
                if (p->safe_left[r] > 0) ++nsafe;
            }
        if (nrun > 0 && nsafe == nrun) {
            if ((rc = ensure_safe(p))) return rc;
            HIPCHK(hipGraphLaunch(p->g_safe, p->stream));
            if ((rc = pull_states(p))) return rc;
            bool failed = false;
            for (int r = 0; r < p->capR; ++r) {
                if (!p->live[r]) continue;
                if (p->safe_left[r] > 0) --p->safe_left[r];
                if (p->S[r].violation && !p->S[r].fixed) {
                    p->S[r].frozen = FROZEN_FAILED;
                    failed = true;
                }
            }
            if (failed && (rc = push_states(p))) return rc;
            continue;
        }
        HIPCHK(hipGraphLaunch(p->g_norm, p->stream));
        if ((rc = pull_states(p))) return rc;
        bool viol = false;
        for (int r = 0; r < p->capR; ++r) {
            if (!p->live[r]) continue;
            if (p->S[r].violation)
                viol = true;
            else
                p->replayed[r] = 0;  // this normal batch went through
        }
        if (viol && (rc = replay(p))) return rc;
    }
    if (n_running) *n_running = count_running(p);
    return GCR_OK;
}

int gcr_replica_status(gcr_pool* p, int32_t rid, double out[8])
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (!out) return fail(GCR_ERR_ARG, "null output");
    if ((rc = pull_states(p))) return rc;
    const RepState& S = p->S[rid];
    out[0] = S.dt;
    out[1] = S.elapsed;
    out[2] = (double)S.steps;
    out[3] = S.maxm;
    out[4] = running(S) ? 1.0 : 0.0;
    out[5] = (double)p->replays[rid];
    out[6] = (double)S.step_global;
    out[7] = S.frozen == FROZEN_FAILED ? 1.0 : 0.0;
    return GCR_OK;
}

int gcr_replica_download(gcr_pool* p, int32_t rid, double* pos, double* pos1, double* tangent,
                         double* forces)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (p->tables_dirty && (rc = rebuild_tables(p))) return rc;
    const int dim = p->dim;
    if (pos && (rc = download_one(p, rid, p->X, pos, dim))) return rc;
    if (pos1 && (rc = download_one(p, rid, p->X1, pos1, dim))) return rc;
    if (forces && (rc = download_one(p, rid, p->FT, forces, dim))) return rc;
    if (tangent) {
        if (p->P[rid].flags & GCR_F_ORIENT) {
            // tangents at x1, from the bonds (needs valid links: built by a run)
            if (dim == 2)
                k_tangents<2><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
            else
                k_tangents<3><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
            // k_tangents writes Tn; move it out of the staging buffer first
            HIPCHK(hipMemcpyAsync(p->Ts, p->Tn, sizeof(double) * (size_t)p->capB * dim,
                                  hipMemcpyDeviceToDevice, p->stream));
            if ((rc = download_one(p, rid, p->Ts, tangent, dim))) return rc;
        } else if ((rc = download_one(p, rid, p->T, tangent, dim))) {
            return rc;
        }
    }
    return GCR_OK;
}

int gcr_replica_eval_forces(gcr_pool* p, int32_t rid, double* forces, double* tangent)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (p->tables_dirty && (rc = rebuild_tables(p))) return rc;
    if ((rc = with_only(p, rid, true))) return rc;
    launch_rebuild(p, -1);
    const bool mixed = p->precision == GCR_MIXED;
    if (p->dim == 2)
        mixed ? launch_tile<float, 2>(p, p->a, rid, nullptr, nullptr)
              : launch_tile<double, 2>(p, p->a, rid, nullptr, nullptr);
    else
        mixed ? launch_tile3<float, 2>(p, p->a, rid, nullptr, nullptr)
              : launch_tile3<double, 2>(p, p->a, rid, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p->Ts, p->Tn, sizeof(double) * (size_t)p->capB * p->dim,
                          hipMemcpyDeviceToDevice, p->stream));
    if ((rc = with_only(p, rid, false))) return rc;
    if (forces && (rc = download_one(p, rid, p->FT, forces, p->dim))) return rc;
    if (tangent && (rc = download_one(p, rid, p->Ts, tangent, p->dim))) return rc;
    return GCR_OK;
}

int gcr_replica_order(gcr_pool* p, int32_t rid, double out[4])
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (!out) return fail(GCR_ERR_ARG, "null output");
    if (p->tables_dirty && (rc = rebuild_tables(p))) return rc;
    const RepParams& P = p->P[rid];
    const double* tang = p->T;
    if (P.flags & GCR_F_ORIENT) {
        // tangents at x1 from the bonds, as the download returns them
        if (p->dim == 2)
            k_tangents<2><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
        else
            k_tangents<3><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
        tang = p->Tn;
    }
    const int nblk = (P.n_chains + BS - 1) / BS;
    k_chain_order<<<nblk, BS, 0, p->stream>>>(p->a, rid, tang, p->dim, p->OBS);
    k_order_final<<<1, 64, 0, p->stream>>>(p->OBS, nblk);
    HIPCHK(hipGetLastError());
    double sums[4];
    HIPCHK(hipMemcpyAsync(sums, p->OBS + 4 * nblk, sizeof(sums), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    // analysis.py:294-328: polar = |sum e^{i phi}| / C; S = |(q11, q12)| / C,
    // director = arctan((|q| - q11) / q12) (0 when q12 = 0)
    const double C = (double)P.n_chains;
    const double q11 = sums[2], q12 = sums[3], qf = std::hypot(q11, q12);
    out[0] = std::hypot(sums[0], sums[1]) / C;
    out[1] = qf / C;
    out[2] = q12 != 0.0 ? std::atan((qf - q11) / q12) : 0.0;
    out[3] = C;
    return GCR_OK;
}

int gcr_replica_sample(gcr_pool* p, int32_t rid, int32_t reset)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    if (p->tables_dirty && (rc = rebuild_tables(p))) return rc;
    if (!p->ser_dev[rid]) {
        HIPCHK(hipMalloc((void**)&p->ser_dev[rid], sizeof(double) * GCR_SERIES_COLS * GCR_SERIES_MAX));
        p->ser_rows[rid] = 0;
        reset = 1;  // no origin yet
    }
    if (reset) p->ser_rows[rid] = 0;
    if (p->ser_rows[rid] >= GCR_SERIES_MAX) return fail(GCR_ERR_CAPACITY, "observables series full");
    const RepParams& P = p->P[rid];
    const double* tang = p->T;
    if (P.flags & GCR_F_ORIENT) {
        // tangents at x1 from the bonds, as the download returns them
        if (p->dim == 2)
            k_tangents<2><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
        else
            k_tangents<3><<<p->nblk, BS, 0, p->stream>>>(p->a, rid);
        tang = p->Tn;
    }
    const int nblk = (P.n_chains + BS - 1) / BS;
    if (p->dim == 2)
        k_chain_sample<2><<<nblk, BS, 0, p->stream>>>(p->a, rid, tang, reset ? 1 : 0, p->COMW, p->COMU, p->OBS);
    else
        k_chain_sample<3><<<nblk, BS, 0, p->stream>>>(p->a, rid, tang, reset ? 1 : 0, p->COMW, p->COMU, p->OBS);
    k_sample_final<<<1, 64, 0, p->stream>>>(p->a, rid, p->OBS, nblk,
                                            p->ser_dev[rid] + (size_t)GCR_SERIES_COLS * p->ser_rows[rid]);
    HIPCHK(hipGetLastError());
    p->ser_rows[rid] += 1;
    return GCR_OK;
}

int gcr_replica_series(gcr_pool* p, int32_t rid, double* out, int32_t max_rows, int32_t* rows)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    const int n = p->ser_rows[rid];
    if (rows) *rows = n;
    if (out && n > 0 && max_rows > 0) {
        const int m = n < max_rows ? n : max_rows;
        HIPCHK(hipMemcpyAsync(out, p->ser_dev[rid], sizeof(double) * GCR_SERIES_COLS * (size_t)m,
                              hipMemcpyDeviceToHost, p->stream));
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    return GCR_OK;
}

int gcr_replica_cell_list(gcr_pool* p, int32_t rid, int32_t* nc, int32_t* cstart, int32_t* slot)
{
    int rc = check_rid(p, rid);
    if (rc) return rc;
    const RepParams& P = p->P[rid];
    if (nc) *nc = P.nc;
    if (cstart) {
        HIPCHK(hipMemcpyAsync(cstart, p->CSTART + P.cell_off, sizeof(int32_t) * ((size_t)P.ncell + 1),
                              hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        for (int c = 0; c <= P.ncell; ++c) cstart[c] -= P.bead_off;
    }
    if (slot) {
        HIPCHK(hipMemcpyAsync(slot, p->SLOT + P.bead_off, sizeof(int32_t) * (size_t)P.n,
                              hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        for (int b = 0; b < P.n; ++b) slot[b] -= P.bead_off;
    }
    return GCR_OK;
}

int gcr_pool_time_kernels(gcr_pool* p, int32_t steps, double ms_out[4])
{
    if (!p || !ms_out || steps < 1) return fail(GCR_ERR_ARG, "bad arguments");
    int rc;
    if (p->tables_dirty && (rc = rebuild_tables(p))) return rc;
    if (p->has_srk4 || p->has_group)
        return fail(GCR_ERR_ARG, "kernel timing covers pools of independent SRK2 replicas");
    // The step sequence of the batch graphs, launched directly; the three
    // per-step kernels record start/end events in their own dispatch packets
    // (hipExtLaunchKernelGGL) on the pool's stream; the re-sort (several
    // kernels) is bracketed by stream events.
    hipEvent_t ev[8];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double acc[4] = {0, 0, 0, 0};
    int nreb = 0;
    float t;
    launch_begin(p);
    for (int j = 0; j < steps; ++j) {
        const bool withr = (j % p->K == 0);
        if (withr) {
            HIPCHK(hipEventRecord(ev[6], p->stream));
            launch_rebuild(p, j);
            HIPCHK(hipEventRecord(ev[7], p->stream));
        }
        launch_stage(p, 0, j, ev[0], ev[1]);
        launch_rod(p, 0, j);
        launch_stage(p, 1, j, ev[2], ev[3]);
        launch_stage(p, 3, j);
        launch_rod(p, 1, j);
        launch_stage(p, 2, j, ev[4], ev[5]);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(p->stream));
        if (withr) {
            HIPCHK(hipEventElapsedTime(&t, ev[6], ev[7]));
            acc[3] += t;
            ++nreb;
        }
        for (int k = 0; k < 3; ++k) {
            HIPCHK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
            acc[k] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    ms_out[0] = acc[0] / steps;
    ms_out[1] = acc[1] / steps;
    ms_out[2] = acc[2] / steps;
    ms_out[3] = nreb ? acc[3] / nreb : 0.0;
    if ((rc = pull_states(p))) return rc;
    bool viol = false;
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && p->S[r].violation) viol = true;
    // put every timed replica back to the snapshot launch_begin took
    // (positions, directions, bead order, step state) and re-sort it there:
    // the timing steps leave no trace in later runs
    for (int r = 0; r < p->capR; ++r)
        if (p->live[r] && p->S[r].frozen == 0) p->S[r].frozen = 2;
    if ((rc = push_states(p))) return rc;
    if (p->dim == 2)
        k_restore<2><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
    else
        k_restore<3><<<p->nblk, BS, 0, p->stream>>>(p->a, 1);
    k_restore_state<<<(p->capR + 63) / 64, 64, 0, p->stream>>>(p->a);
    launch_rebuild(p, -1);
    HIPCHK(hipGetLastError());
    if ((rc = pull_states(p))) return rc;
    if (viol) return fail(GCR_ERR_STATE, "skin exceeded during kernel timing; timings discarded");
    return GCR_OK;
}

int gcr_pool_graph_stats(gcr_pool* p, double out[4])
{
    if (!p || !out) return fail(GCR_ERR_ARG, "null argument");
    out[0] = (double)p->n_capture;
    out[1] = (double)p->n_instantiate;
    out[2] = (double)p->n_update;
    out[3] = p->graph_seconds;
    return GCR_OK;
}

int gcr_pool_set_tile(gcr_pool* p, int32_t side)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    if (side < 0 || side > TILE_MAX_ROWS) return fail(GCR_ERR_ARG, "tile side must be in [0, 16]");
    p->tile_side = side;
    p->tile_w = p->tile_h = 0;
    p->tables_dirty = true;
    return GCR_OK;
}

int gcr_pool_set_tile_shape(gcr_pool* p, int32_t tile_x, int32_t tile_y)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    if (p->dim != 2) return fail(GCR_ERR_ARG, "rectangular tiles are 2D");
    const bool clear = tile_x == 0 && tile_y == 0;
    if (!clear && (tile_x < 1 || tile_x > TILE_MAX_ROWS || tile_y < 1 || tile_y > TILE_MAX_ROWS))
        return fail(GCR_ERR_ARG, "tile edges must be in [1, 16] (or both 0)");
    p->tile_w = tile_x;
    p->tile_h = tile_y;
    if (!clear) p->tile_side = 0;
    p->tables_dirty = true;
    return GCR_OK;
}

int gcr_pool_set_tile_threads(gcr_pool* p, int32_t threads)
{
    if (!p) return fail(GCR_ERR_ARG, "null pool");
    if (threads != 0 && threads != 256 && threads != 512 && threads != 1024)
        return fail(GCR_ERR_ARG, "threads per tile must be 0, 256, 512 or 1024");
    p->tile_threads = threads;
    p->tables_dirty = true;
    return GCR_OK;
}

int gcr_pool_destroy(gcr_pool* p)
{
    if (!p) return GCR_OK;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    drop_graphs(p);
    for (double* q : p->xi_dev)
        if (q) (void)hipFree(q);
    for (double* q : p->ser_dev)
        if (q) (void)hipFree(q);
    void* bufs[] = {p->blk_rep, p->blk_slot0, p->chk_rep, p->chk_cell0, p->rep_chk_off, p->chb_rep,
                    p->chb_chain0, p->dP, p->dS,
                    p->dSs, p->X, p->X1, p->F0, p->FT, p->XI, p->T, p->Xn, p->Tn, p->Xs, p->Ts,
                    p->XIn, p->X1n,
                    p->META, p->LINK, p->CELL, p->SLOT, p->KEYOFF, p->PAIRS, p->METAn, p->METAs,
                    p->CHAIN_A, p->CHAIN_CS, p->CHAIN_FIRST, p->CHAIN_LEN, p->CCOUNT, p->CSTART,
                    p->CHKSUM, p->BMAX, p->TILES, p->TT, p->NL, p->NLEN, p->XR, p->OBS, p->COMW, p->COMU,
                    p->XPAIRS, p->GRP, p->CHAIN_CHUNK, p->WA, p->WAn, p->MC, p->SCNT};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (p->S) (void)hipHostFree(p->S);
    delete p;
    return GCR_OK;
}

// ------------------------------------------------------ single system API

int gcr_create(const gcr_replica_desc* d, int32_t dim, int32_t precision, const double* pos,
               const double* tangent, const int64_t* chain_id, const double* rest_lengths,
               const double* v_tan_angle, gcr_pool** out)
{
    if (!d || !out) return fail(GCR_ERR_ARG, "null argument");
    int64_t ncell = gcr_replica_cells(dim, d);
    gcr_pool* p = nullptr;
    int rc = gcr_pool_create(dim, precision, d->n_beads, ncell + 1, d->n_chains, 1, &p);
    if (rc) return rc;
    int32_t rid;
    rc = gcr_replica_add(p, d, pos, tangent, chain_id, rest_lengths, v_tan_angle, nullptr, &rid);
    if (rc) {
        std::string m = g_err;
        gcr_pool_destroy(p);
        return fail(rc, m);
    }
    *out = p;
    return GCR_OK;
}

int gcr_upload(gcr_pool* h, const double* pos, const double* tangent)
{
    return gcr_replica_upload(h, 0, pos, tangent);
}

int gcr_run(gcr_pool* h, const double tv[8], double status_out[6], const double* xi, int64_t xi_steps)
{
    int rc = gcr_replica_begin(h, 0, tv, xi, xi_steps);
    if (rc) return rc;
    int32_t nr = 0;
    rc = gcr_pool_run(h, 0, &nr);
    if (rc) return rc;
    if ((rc = pull_states(h))) return rc;
    if (h->S[0].frozen == FROZEN_FAILED)
        return fail(GCR_ERR_STATE, "a bead moved more than skin/2 within one step (pair set no "
                                   "longer complete); create the system with a larger skin");
    if (status_out) {
        double st[8];
        if ((rc = gcr_replica_status(h, 0, st))) return rc;
        status_out[0] = st[0];
        status_out[1] = st[1];
        status_out[2] = st[2];
        status_out[3] = st[3];
        status_out[4] = 0.0;
        status_out[5] = 0.0;
    }
    return GCR_OK;
}

int gcr_download(gcr_pool* h, double* pos, double* pos1, double* tangent, double* forces)
{
    return gcr_replica_download(h, 0, pos, pos1, tangent, forces);
}

int gcr_eval_forces(gcr_pool* h, double* forces, double* tangent)
{
    return gcr_replica_eval_forces(h, 0, forces, tangent);
}

int gcr_destroy(gcr_pool* h) { return gcr_pool_destroy(h); }

}  // extern "C"

// The force evaluation without device allocation (gcr_eval_forces_noalloc).
#include "gcr_noalloc.hip"
