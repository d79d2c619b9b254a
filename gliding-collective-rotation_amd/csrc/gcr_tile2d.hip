// gcr_tile2d.hip -- 2D force stages over tiles of cells staged in LDS.
//
// One workgroup owns a tile of tile_x x tile_y cells of one replica. It stages
// the beads of the tile and of a one-cell halo into LDS (positions relative
// to the centre of the cell each bead is binned in, plus (orig, chain)), then
// every owned bead sums its pair forces over the 3 x 3 surrounding cells out
// of LDS. Relative coordinates make the periodic image implicit (the
// neighbour cell's offset is -h, 0 or +h), so the inner loop has no wrap and,
// in the mixed form, runs entirely in FP32 with separations accurate to the
// FP32 ulp of a cell edge. Bonded terms, propulsion and the SRK2 stage update
// are the same device functions as k_stage; outputs are one write per bead.
// A tile whose beads do not fit in LDS (or a box of fewer than 3 cells per
// side) sums its pairs from global memory with pair_sum instead.
#define TILE_MAX_HALO 400    // (tile_x + 2) * (tile_y + 2) cells
#define TILE_MAX_BEADS 1024  // beads of a tile with its halo held in LDS (16-24 B each)
#define TILE_MAX_ROWS 16

namespace gcr {

// LDS record of a staged bead: position relative to the centre of its cell
// and (original index, chain). FP32: one 16-byte record; FP64: 24 bytes.
template <typename FT>
struct alignas(16) LBead {
    FT x, y;
    int orig, chain;
};

// Scalars of the pair law, held in registers for the whole loop.
template <typename FT>
struct PairConsts {
    FT h, rc2, p0, p1, shift;
    int nn;
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
    k.nn = P.n_nearby;
    k.shifted = P.wca[2] != 0.0;
    return k;
}

// Pair forces of one bead from the LDS tile: cell c0 = hy * W + hx is its
// cell in the tile frame, (xi, yi) its position relative to that cell's
// centre. Branch-free body: every candidate of the 3 x 3 cells is evaluated
// with a mask (cutoff, self, same-chain exclusion of updates.py:645-650), so
// all lanes follow one path and the LDS reads pipeline. The forces of one
// neighbour cell are summed in FT and added to the FP64 accumulator.
template <typename FT>
__device__ __forceinline__ void tile_pair_sum(const LBead<FT>* lb, const int* cst, int W, int c0, FT xi,
                                              FT yi, int2 me, const PairConsts<FT>& k, double* acc)
{
    for (int dy = -1; dy <= 1; ++dy) {
        const FT oy = (FT)dy * k.h - yi;
        for (int dx = -1; dx <= 1; ++dx) {
            const FT ox = (FT)dx * k.h - xi;
            const int c = c0 + dy * W + dx;
            const int t1 = cst[c + 1];
            FT fx = (FT)0, fy = (FT)0;
            if (!k.shifted) {
#pragma unroll 4
                for (int t = cst[c]; t < t1; ++t) {
                    const LBead<FT> q = lb[t];
                    const FT sx = q.x + ox, sy = q.y + oy;
                    const FT ds = sx * sx + sy * sy;
                    const int d = q.orig - me.x;
                    // bitwise, not short-circuit: keeps the body free of branches
                    const int excl = (q.chain == me.y) & ((k.nn < 0) | ((d < 0 ? -d : d) <= k.nn));
                    const bool use = (ds <= k.rc2) & (ds > (FT)0) & !excl;
                    const FT cf = PairLaw<FT>::coeff(use ? ds : (FT)1, k.p0, k.p1);
                    const FT w = use ? cf : (FT)0;
                    fx += sx * w;
                    fy += sy * w;
                }
            } else {
                for (int t = cst[c]; t < t1; ++t) {
                    const LBead<FT> q = lb[t];
                    FT sx = q.x + ox, sy = q.y + oy;
                    FT ds = sx * sx + sy * sy;
                    if (!(ds <= k.rc2 && ds > (FT)0)) continue;
                    if (q.chain == me.y) {
                        if (k.nn < 0) continue;
                        const int d = q.orig - me.x;
                        if ((d < 0 ? -d : d) <= k.nn) continue;
                    }
                    // WCA_shift, forces.py:123-131
                    FT mag = sqrt(ds);
                    FT nm = mag + k.shift;
                    nm = nm < (FT)0 ? (FT)0 : nm;
                    FT rs = nm / mag;
                    sx *= rs;
                    sy *= rs;
                    ds = nm * nm;
                    const FT cf = PairLaw<FT>::coeff(ds, k.p0, k.p1);
                    fx += sx * cf;
                    fy += sy * cf;
                }
            }
            acc[0] += (double)fx;
            acc[1] += (double)fy;
        }
    }
}

// STAGE 0 / 1 as k_stage; STAGE 2: force evaluation of replica `j` only at
// the positions X (forces into FT, tangents into Tn; no noise, no update).
template <typename FT, int STAGE>
__global__ void __launch_bounds__(BS) k_stage_tile2d(PoolArgs a, int j)
{
    const int4 tile = reinterpret_cast<const int4*>(a.TILES)[blockIdx.x];
    const int r = tile.x;
    const RepParams& P = a.P[r];
    RepState* S = a.S + r;
    if (STAGE == 2) {
        if (r != j) return;
    } else if (!step_runs(S, j)) {
        return;
    }
    const double* Xa = STAGE == 1 ? a.X1 : a.X;
    const int nc = P.nc, TX = P.tile_x, TY = P.tile_y, W = TX + 2, H = TY + 2;
    const int cx0 = tile.y, cy0 = tile.z;
    const int own_x = min(TX, nc - cx0), own_y = min(TY, nc - cy0);
    const int nhalo = W * H;

    __shared__ int cst[TILE_MAX_HALO + 1];  // LDS offset of each halo cell
    __shared__ int gst[TILE_MAX_HALO];      // first global slot of each halo cell
    __shared__ int own_pref[TILE_MAX_ROWS + 1];
    __shared__ LBead<FT> lb[TILE_MAX_BEADS];

    // 1. bead counts of the halo cells -> LDS offsets (two cells per thread)
    int cnt[2], g0[2], sum = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = threadIdx.x * 2 + k;
        cnt[k] = 0;
        g0[k] = 0;
        if (c < nhalo) {
            const int gx = (cx0 - 1 + c % W + nc) % nc;
            const int gy = (cy0 - 1 + c / W + nc) % nc;
            const int gc = P.cell_off + gy * nc + gx;
            g0[k] = a.CSTART[gc];
            cnt[k] = a.CSTART[gc + 1] - g0[k];
        }
        sum += cnt[k];
    }
    int total;
    int ex = block_exclusive_scan(sum, &total);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = threadIdx.x * 2 + k;
        if (c < nhalo) {
            cst[c] = ex;
            gst[c] = g0[k];
        }
        ex += cnt[k];
    }
    if (threadIdx.x == 0) cst[nhalo] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        // owned beads of tile row ry are the LDS range of cells 1..own_x of halo row ry+1
        int p = 0;
        for (int ry = 0; ry < own_y; ++ry) {
            own_pref[ry] = p;
            const int row = (ry + 1) * W;
            p += cst[row + own_x + 1] - cst[row + 1];
        }
        own_pref[own_y] = p;
    }
    const bool in_lds = (P.flags & GCR_F_WCA) && total <= TILE_MAX_BEADS && nc >= 3;

    // 2. stage the halo: positions relative to the centre of their cell
    if (in_lds) {
        for (int i = threadIdx.x; i < total; i += BS) {
            int lo = 0, hi = nhalo;  // cst[lo] <= i < cst[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (cst[mid] <= i)
                    lo = mid;
                else
                    hi = mid;
            }
            const int gx = (cx0 - 1 + lo % W + nc) % nc;
            const int gy = (cy0 - 1 + lo / W + nc) % nc;
            const int g = gst[lo] + (i - cst[lo]);
            const double2 x = *reinterpret_cast<const double2*>(Xa + 2 * g);
            const int2 m = *reinterpret_cast<const int2*>(a.META + 2 * g);
            LBead<FT> q;
            q.x = (FT)wrap1(x.x - (gx + 0.5) * P.h, P.L, P.halfL);
            q.y = (FT)wrap1(x.y - (gy + 0.5) * P.h, P.L, P.halfL);
            q.orig = m.x;
            q.chain = m.y;
            lb[i] = q;
        }
    }
    __syncthreads();

    // 3. owned beads
    const PairConsts<FT> kc = pair_consts<FT>(P);
    const int n_own = own_pref[own_y];
    double fsq = 0.0;
    for (int o = threadIdx.x; o < n_own; o += BS) {
        int ry = 0;
        while (own_pref[ry + 1] <= o) ++ry;
        const int row = (ry + 1) * W;
        const int li = cst[row + 1] + (o - own_pref[ry]);
        int hx = 1;
        while (cst[row + hx + 1] <= li) ++hx;
        const int c0 = row + hx;
        const int s = gst[c0] + (li - cst[c0]);
        double xs[2], acc[2], tang[2];
        load_vec<2>(Xa, s, xs);
        const int2 me = *reinterpret_cast<const int2*>(a.META + 2 * s);
        const bool springs = (P.flags & GCR_F_SPRINGS) != 0;
        ChainNb<2> cn;
        if (springs) load_chain_nb<2>(Xa, a.LINK, s, cn);  // in flight during the pair loop
        double pa[2] = {0.0, 0.0};
        if (P.flags & GCR_F_WCA) {
            if (in_lds) {
                const LBead<FT> q = lb[li];
                tile_pair_sum<FT>(lb, cst, W, c0, q.x, q.y, me, kc, pa);
            } else {
                pair_sum<2, FT>(Xa, a.META, a.CSTART, xs, me.x, me.y, a.CELL[s], P, pa);
            }
        }
        if (STAGE == 1) {
            load_vec<2>(a.F0, s, acc);
        } else {
            acc[0] = 0.0;
            acc[1] = 0.0;
        }
        // bonded terms first, then the pair sum, then propulsion (SRK.py:228-253)
        bool tang_set = false;
        if (springs) bonded<2, FT>(cn, xs, a.CHAIN_A[me.y], P, acc, tang, tang_set);
        acc[0] += pa[0];
        acc[1] += pa[1];
        if (STAGE == 2) {
            if (!tang_set) load_vec<2>(a.T, s, tang);
            if (P.flags & GCR_F_PROPEL) propel<2>(a, me.y, tang, acc);
            store_vec<2>(a.FT, s, acc);
            store_vec<2>(a.Tn, s, tang);
        } else {
            if (P.flags & GCR_F_PROPEL) {
                if (!tang_set) load_vec<2>(a.T, s, tang);
                propel<2>(a, me.y, tang, acc);
            }
            const double q = stage_store<2, STAGE>(a, P, S, s, me.x, xs, acc);
            fsq = q > fsq ? q : fsq;
        }
    }
    if (STAGE == 1) {
        const double m = block_max(fsq);
        if (threadIdx.x == 0) a.BMAX[blockIdx.x] = m;
    }
}

}  // namespace gcr
