// gcr_tile3d.hip -- 3D force stages over boxes of cells staged in LDS, with
// per-bead neighbour lists of LDS indices.
//
// One workgroup owns a cubic tile of side^3 cells of one replica (side =
// RepParams.tile_x) and stages the beads of the tile and of a one-cell halo
// into LDS in a fixed order: halo cells in (z, y, x) order, a cell's beads in
// slot order. That order depends on the cell contents only, so it is the
// same at every step between two re-sorts. After each re-sort
// k_nlist_tile3d stores for every owned bead the LDS indices of the beads
// within rc + skin that the same-chain rule of updates.py:645-650 does not
// exclude, in the order of the 3 x 3 x 3 cell scan (z, y, x, then slots):
// the three cells of a halo row are one LDS range, so the scan is nine
// ranges. The stage kernels then sum pair forces over the list from LDS.
//
// LDS records hold each coordinate relative to the centre of the bead's cell
// (rounded to FT) and the halo coordinates of that cell, so a separation is
// fma(dc, h, rel_j - rel_i) with the cell offset dc in {-1, 0, 1} -- the
// operations the global cell scan (pair_sum) performs, in the same order, so
// a bead gets the same bits from its LDS list, from the cell scan, or from a
// different tiling. A bead whose list overflowed, and every bead of a tile
// that does not fit in LDS (or of a box of fewer than 3 cells per side), sums
// its pairs by the cell scan from global memory.
#define T3_MAX_SIDE 8
#define T3_MAX_HALO 1000   // (side + 2)^3
#define T3_MAX_ROWS 64     // owned (y, z) rows of a tile
#define T3_MAX_BEADS 1536  // beads of a tile with its halo held in LDS
// neighbour-list entries per 3D bead (32 groups of 4): a bead inside an LJ
// cluster has about 90 beads within rc + skin, and one that overflows its
// list sums its pairs by the slow global cell scan, which sets the launch's
// length
#define NL3_MAX 128
// Layout table of a 3D tile (TT3_INTS ints at TT[tile * TT3_INTS], written by
// the list builder at every re-sort): [0] 1 when the stage kernels may use it
// (the tile fits in LDS and no halo axis wraps onto itself), [1] staged
// beads, [2] owned beads, [3] owned rows; from TT3_HALO, per halo row
// hz W + hy: (first slot of its first run, beads in that run, first slot of
// the run after the x wrap, LDS offset of the row), a halo row being one or
// two runs of slots; from TT3_OWN, per owned row oz own_y + oy: (first slot,
// beads, LDS index of its first bead).
#define TT3_INTS 640
#define TT3_HALO 8
#define TT3_OWN (TT3_HALO + 4 * (T3_MAX_SIDE + 2) * (T3_MAX_SIDE + 2))
static_assert(TT3_OWN + 3 * T3_MAX_ROWS <= TT3_INTS, "3D tile table size");

namespace gcr {

template <typename FT>
struct LPos3;

template <>
struct alignas(8) LPos3<float> {
    float x, y, z;
    float cx, cy, cz;
};

template <>
struct alignas(8) LPos3<double> {
    double x, y, z;
    float cx, cy, cz, pad;
};

template <typename FT>
__device__ __forceinline__ LPos3<FT> make_lpos3(const double* rel, int hx, int hy, int hz)
{
    LPos3<FT> q;
    q.x = (FT)rel[0];
    q.y = (FT)rel[1];
    q.z = (FT)rel[2];
    q.cx = (float)hx;
    q.cy = (float)hy;
    q.cz = (float)hz;
    return q;
}

// Separation q - me (cell offsets -1, 0 or +1 for every listed pair).
template <typename FT>
__device__ __forceinline__ void separation3(const LPos3<FT>& q, const LPos3<FT>& me, FT h, FT* sep)
{
    sep[0] = cell_sep<FT>((FT)(q.cx - me.cx), h, q.x, me.x);
    sep[1] = cell_sep<FT>((FT)(q.cy - me.cy), h, q.y, me.y);
    sep[2] = cell_sep<FT>((FT)(q.cz - me.cz), h, q.z, me.z);
}

// Pair forces of one bead over its list of LDS indices (groups of 4 u16 at
// G[g * stride], the last group padded with the bead's own index, which the
// zero-distance mask drops), summed in FT. The groups are read from global
// memory four at a time, the next four in flight while the current ones are
// summed: a 3D launch keeps only two to four waves per SIMD, too few to hide
// one load's latency behind another wave's pairs.
template <typename FT>
__device__ __forceinline__ void list3_pair_sum(const LPos3<FT>* lp, const uint2* G, int64_t stride, int len,
                                               const LPos3<FT>& me, const PairConsts<FT>& k, FT* f)
{
    const int ng = (len + 3) >> 2;
    const uint2 none = make_uint2(0u, 0u);
    uint2 e[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) e[u] = u < ng ? G[(int64_t)u * stride] : none;
    for (int g0 = 0; g0 < ng; g0 += 4) {
        uint2 n[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) n[u] = g0 + 4 + u < ng ? G[(int64_t)(g0 + 4 + u) * stride] : none;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (g0 + v >= ng) break;
            const uint2 ev = e[v];
            const LPos3<FT> q[4] = {lp[ev.x & 0xffffu], lp[ev.x >> 16], lp[ev.y & 0xffffu], lp[ev.y >> 16]};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                FT sep[3];
                separation3<FT>(q[u], me, k.h, sep);
                if (k.shifted)
                    pair_term_shifted<3, FT>(sep, k, f);
                else
                    pair_term<3, FT>(sep, k, f);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = n[u];
    }
}

// Geometry of a 3D tile and its halo: halo cell c = (hz V + hy) W + hx with
// W = V = U = side + 2; owned rows (oy, oz) in order oz * own_y + oy.
struct T3Layout {
    int cx0, cy0, cz0, W, nhalo, own_x, own_y, own_z, total;
    bool in_lds;
};

struct T3Shared {
    int cst[T3_MAX_HALO + 1];  // LDS offset of each halo cell
    int gst[T3_MAX_HALO];      // first global slot of each halo cell
    int own_pref[T3_MAX_ROWS + 1];
};

// Halo bead counts -> LDS offsets, owned-row prefix; ends with a barrier.
template <int NT>
__device__ __forceinline__ T3Layout tile3_layout(const PoolArgs& a, const RepParams& P, int4 tile,
                                                 T3Shared& sh)
{
    constexpr int CPT = (T3_MAX_HALO + NT - 1) / NT;  // halo cells per thread
    T3Layout t;
    const int nc = P.nc, side = P.tile_x;
    t.W = side + 2;
    t.cx0 = tile.y;
    t.cy0 = tile.z;
    t.cz0 = tile.w;
    t.own_x = min(side, nc - t.cx0);
    t.own_y = min(side, nc - t.cy0);
    t.own_z = min(side, nc - t.cz0);
    t.nhalo = t.W * t.W * t.W;
    int cnt[CPT], g0[CPT], sum = 0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = threadIdx.x * CPT + k;
        cnt[k] = 0;
        g0[k] = 0;
        if (c < t.nhalo) {
            const int gx = (t.cx0 - 1 + c % t.W + nc) % nc;
            const int gy = (t.cy0 - 1 + (c / t.W) % t.W + nc) % nc;
            const int gz = (t.cz0 - 1 + c / (t.W * t.W) + nc) % nc;
            const int gc = P.cell_off + (gz * nc + gy) * nc + gx;
            g0[k] = a.CSTART[gc];
            cnt[k] = a.CSTART[gc + 1] - g0[k];
        }
        sum += cnt[k];
    }
    int ex = block_exclusive_scan<NT>(sum, &t.total);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = threadIdx.x * CPT + k;
        if (c < t.nhalo) {
            sh.cst[c] = ex;
            sh.gst[c] = g0[k];
        }
        ex += cnt[k];
    }
    if (threadIdx.x == 0) sh.cst[t.nhalo] = t.total;
    __syncthreads();
    if (threadIdx.x == 0) {
        // owned row (oy, oz): the LDS range of cells 1..own_x of halo row (oy + 1, oz + 1)
        int p = 0;
        for (int oz = 0; oz < t.own_z; ++oz)
            for (int oy = 0; oy < t.own_y; ++oy) {
                sh.own_pref[oz * t.own_y + oy] = p;
                const int c = ((oz + 1) * t.W + oy + 1) * t.W + 1;
                p += sh.cst[c + t.own_x] - sh.cst[c];
            }
        sh.own_pref[t.own_y * t.own_z] = p;
    }
    t.in_lds = (P.flags & GCR_F_WCA) && t.total <= T3_MAX_BEADS && nc >= 3;
    __syncthreads();
    return t;
}

// Halo cell of staged bead i: a binary search over the cells' LDS offsets.
__device__ __forceinline__ int staged_cell3(const T3Shared& sh, const T3Layout& t, int i)
{
    int lo = 0, hi = t.nhalo;  // cst[lo] <= i < cst[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sh.cst[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Stage the halo's positions (and, for the list builder, (orig, chain)), and
// the halo cell of every staged bead (FP32: marked per cell range; FP64: by
// a search, which keeps registers free). Block-uniform call.
template <typename FT, int NT>
__device__ __forceinline__ void stage3(const PoolArgs& a, const RepParams& P, const double* Xa,
                                       const T3Shared& sh, const T3Layout& t, LPos3<FT>* lp, int2* lm,
                                       uint16_t* lcell)
{
    const int nc = P.nc;
    constexpr bool mark = sizeof(FT) == 4;
    if (mark) {
        for (int c = threadIdx.x; c < t.nhalo; c += NT) {
            const int i1 = sh.cst[c + 1];
            for (int i = sh.cst[c]; i < i1; ++i) lcell[i] = (uint16_t)c;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < t.total; i += NT) {
        const int hc = mark ? (int)lcell[i] : staged_cell3(sh, t, i);
        if (!mark) lcell[i] = (uint16_t)hc;
        const int g = sh.gst[hc] + (i - sh.cst[hc]);
        const int hx = hc % t.W, hy = (hc / t.W) % t.W, hz = hc / (t.W * t.W);
        const int gc[3] = {(t.cx0 - 1 + hx + nc) % nc, (t.cy0 - 1 + hy + nc) % nc, (t.cz0 - 1 + hz + nc) % nc};
        double x[3], rel[3];
        load_vec<3>(Xa, g, x);
#pragma unroll
        for (int d = 0; d < 3; ++d) rel[d] = cell_rel(x[d], gc[d], P);
        lp[i] = make_lpos3<FT>(rel, hx, hy, hz);
        if (lm) lm[i] = *reinterpret_cast<const int2*>(a.META + 2 * g);
    }
}

// Owned bead o of the tile: its LDS index li, halo cell c0 and slot s. A
// thread visits its owned beads in increasing o, so the owned row ro carries
// over from the previous call (start at 0). staged: the cell table lcell is
// filled (tile in LDS); otherwise the cell is found by scanning the row.
__device__ __forceinline__ void owned_bead3(const T3Shared& sh, const T3Layout& t, const uint16_t* lcell,
                                            bool staged, int o, int& ro, int& li, int& c0, int& s)
{
    while (sh.own_pref[ro + 1] <= o) ++ro;
    const int oz = ro / t.own_y, oy = ro - oz * t.own_y;
    const int c = ((oz + 1) * t.W + oy + 1) * t.W + 1;
    li = sh.cst[c] + (o - sh.own_pref[ro]);
    if (staged) {
        c0 = lcell[li];
    } else {
        int hx = 0;
        while (sh.cst[c + hx + 1] <= li) ++hx;
        c0 = c + hx;
    }
    s = sh.gst[c0] + (li - sh.cst[c0]);
}

// The tile's layout table from the layout just computed. Block-uniform call.
__device__ __forceinline__ void write_tile_table3(const PoolArgs& a, const RepParams& P, const T3Layout& t,
                                                  const T3Shared& sh)
{
    int* T = a.TT + (size_t)blockIdx.x * TT3_INTS;
    const int W = t.W, rows = W * W, orows = t.own_y * t.own_z;
    const bool fast = t.in_lds && W <= P.nc;
    if (threadIdx.x == 0) {
        T[0] = fast ? 1 : 0;
        T[1] = t.total;
        T[2] = sh.own_pref[orows];
        T[3] = orows;
    }
    if (!fast) return;
    // halo cells before the x wrap: 1 if the left one wraps, nc - xa if the
    // right ones do, else the whole row
    const int xa = t.cx0 - 1;
    const int k = xa < 0 ? 1 : (xa + W > P.nc ? P.nc - xa : W);
    for (int hr = threadIdx.x; hr < rows; hr += blockDim.x) {
        const int c = hr * W;  // the row's first halo cell
        int* R = T + TT3_HALO + 4 * hr;
        R[0] = sh.gst[c];
        R[1] = sh.cst[c + k] - sh.cst[c];
        R[2] = k < W ? sh.gst[c + k] : 0;
        R[3] = sh.cst[c];
    }
    for (int ro = threadIdx.x; ro < orows; ro += blockDim.x) {
        const int oz = ro / t.own_y, oy = ro - oz * t.own_y;
        const int c = ((oz + 1) * W + oy + 1) * W + 1;  // owned cells 1 .. own_x of halo row (oy+1, oz+1)
        int* R = T + TT3_OWN + 3 * ro;
        R[0] = sh.gst[c];
        R[1] = sh.cst[c + t.own_x] - sh.cst[c];
        R[2] = sh.cst[c];
    }
}

// Owned bead o of a tile from its layout table: (slot, LDS index), or s = -1.
__device__ __forceinline__ void table_owned3(const int* tt, int o, int& s, int& li)
{
    s = -1;
    li = 0;
    int before = 0;
    const int orows = tt[3];
    for (int ro = 0; ro < orows; ++ro) {
        const int* R = tt + TT3_OWN + 3 * ro;
        if (o < before + R[1]) {
            s = R[0] + (o - before);
            li = R[2] + (o - before);
            return;
        }
        before += R[1];
    }
}

// Neighbour lists of the tile's owned beads after a cell re-sort (same run
// condition as the re-sort), in the cell scan's order; NLEN = -1 marks a
// bead without a list (more than NL3_MAX entries, or a tile outside LDS).
template <typename FT, int NT>
__global__ void __launch_bounds__(NT) k_nlist_tile3d(PoolArgs a, int j)
{
    const int4 tile = reinterpret_cast<const int4*>(a.TILES)[blockIdx.x];
    if (!rebuild_runs(a, tile.x, j)) return;
    if (!(a.P[tile.x].flags & GCR_F_WCA)) return;
    // the replica's parameters in LDS: read from global memory after the
    // kernel's first store, each would be a load waiting on memory
    constexpr int PW = (int)(sizeof(RepParams) / 4);
    static_assert(PW <= NT, "one word per thread");
    __shared__ RepParams Ps;
    if ((int)threadIdx.x < PW)
        reinterpret_cast<uint32_t*>(&Ps)[threadIdx.x] = reinterpret_cast<const uint32_t*>(a.P + tile.x)[threadIdx.x];
    __syncthreads();
    const RepParams& P = Ps;
    __shared__ T3Shared sh;
    __shared__ LPos3<FT> lp[T3_MAX_BEADS];
    __shared__ int2 lm[T3_MAX_BEADS];
    __shared__ uint16_t lcell[T3_MAX_BEADS];
    const T3Layout t = tile3_layout<NT>(a, P, tile, sh);
    write_tile_table3(a, P, t, sh);
    if (t.in_lds) stage3<FT, NT>(a, P, a.on_x1 ? a.X1 : a.X, sh, t, lp, lm, lcell);
    __syncthreads();
    const FT rl2 = (FT)P.rl2, h = (FT)P.h;
    const int nn = P.n_nearby;
    const int n_own = sh.own_pref[t.own_y * t.own_z];
    const int WW = t.W * t.W;
    uint2* NL = reinterpret_cast<uint2*>(a.NL);
    int ro = 0;
    for (int o = threadIdx.x; o < n_own; o += NT) {
        int li, c0, s;
        owned_bead3(sh, t, lcell, t.in_lds, o, ro, li, c0, s);
        if (!t.in_lds) {
            a.NLEN[s] = -1;
            continue;
        }
        const LPos3<FT> me = lp[li];
        const int2 mm = lm[li];
        int n = 0;
        uint32_t w0 = 0u, w1 = 0u;  // the group being filled
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy) {
                const int c = c0 + dz * WW + dy * t.W;
                const int q1 = sh.cst[c + 2];
                for (int q = sh.cst[c - 1]; q < q1; ++q) {
                    FT sep[3];
                    separation3<FT>(lp[q], me, h, sep);
                    const int2 m = lm[q];
                    const int d = m.x - mm.x;
                    const bool excl = (m.y == mm.y) & ((nn < 0) | ((d < 0 ? -d : d) <= nn));
                    FT ds = sep[0] * sep[0];
                    ds = fma(sep[1], sep[1], ds);
                    ds = fma(sep[2], sep[2], ds);
                    const bool ok = (q != li) & (ds <= rl2) & !excl;
                    const bool put = ok & (n < NL3_MAX);
                    const uint32_t v = put ? (uint32_t)q << (16 * (n & 1)) : 0u;
                    w0 |= (n & 2) ? 0u : v;
                    w1 |= (n & 2) ? v : 0u;
                    if (put & ((n & 3) == 3)) {
                        NL[(int64_t)(n >> 2) * a.nl_stride + s] = make_uint2(w0, w1);
                        w0 = w1 = 0u;
                    }
                    n += ok ? 1 : 0;
                }
            }
        if (n <= NL3_MAX) {
            if (n & 3) {  // pad the last group with the bead's own index
                for (int sl = n & 3; sl < 4; ++sl) {
                    const uint32_t v = (uint32_t)li << (16 * (sl & 1));
                    if (sl < 2)
                        w0 |= v;
                    else
                        w1 |= v;
                }
                NL[(int64_t)(n >> 2) * a.nl_stride + s] = make_uint2(w0, w1);
            }
            a.NLEN[s] = n;
        } else {
            a.NLEN[s] = -1;
        }
    }
}

// The inputs of one owned bead that come from global memory and that the
// pair loop and the bonded terms need: position, (orig, chain), list length
// and chain links.
struct Head3 {
    int s;
    double x[3];
    int2 me;
    int len;
    int4 lk;
};

__device__ __forceinline__ void load_head3(const PoolArgs& a, const RepParams& P, const double* Xa, int s,
                                           bool listed, Head3& b)
{
    b.s = s;
    load_vec<3>(Xa, s, b.x);
    b.me = *reinterpret_cast<const int2*>(a.META + 2 * s);
    b.len = (listed && (P.flags & GCR_F_WCA)) ? a.NLEN[s] : -1;
    b.lk = (P.flags & GCR_F_SPRINGS) ? *reinterpret_cast<const int4*>(a.LINK + 4 * s) : make_int4(-1, -1, -1, -1);
}

// The 3D force stages: STAGE 0 / 1 as the 2D tile kernels (STAGE 1 for SRK
// stage a.sub), STAGE 2 a force evaluation of replica j only at X (forces
// into FT, tangents into Tn). Per bead: the bonded terms, the pair sum, the
// propulsion and walls, then stage_store -- every feature of the step.
// Prologue: with the tile's layout table (written at the last re-sort) each
// thread finds its first owned bead and its share of the halo from the table
// and issues their loads at once; without one the layout is computed here
// (tile3_layout) and the halo staged after it. In the bead loop the chain
// neighbours' positions are in flight during the pair loop, and the pair sum
// (in FT) is added after the bonded terms, as the reference orders them.
// GEN = false is the lean variant for pools without SRK4, walls, FENE
// springs, sliding and interacting populations: their code stays out of the
// kernel, and with it the registers it would hold (the general FP32 variant
// spills at 128 VGPRs).
template <typename FT, int STAGE, int NT, bool GEN>
__global__ void __launch_bounds__(NT, sizeof(FT) == 4 ? 4 : 2) k_stage_tile3d(PoolArgs a, int j)
{
    const int4 tile = reinterpret_cast<const int4*>(a.TILES)[blockIdx.x];
    const int r = tile.x;
    RepState* S = a.S + r;
    {
        // the run test's reads (state, and the stage's SRK order) issue together
        const uint32_t flags = a.P[r].flags;
        const bool runs = STAGE == 2 ? r == j : step_runs(S, j);
        const bool stage = STAGE != 1 || later_stage_applies(flags, a.sub);
        if (!(runs & stage)) return;
    }
    // the tile's layout table and the replica's parameters in LDS (the
    // parameters: read from global memory after the kernel's first store,
    // each would be a load waiting on memory)
    constexpr int PW = (int)(sizeof(RepParams) / 4);
    __shared__ RepParams Ps;
    __shared__ int tt[TT3_INTS];
    __shared__ T3Shared sh;
    __shared__ LPos3<FT> lp[T3_MAX_BEADS];
    __shared__ uint16_t lcell[T3_MAX_BEADS];
    for (int i = threadIdx.x; i < TT3_INTS + PW; i += NT) {
        if (i < TT3_INTS)
            tt[i] = a.TT[(size_t)blockIdx.x * TT3_INTS + i];
        else
            reinterpret_cast<uint32_t*>(&Ps)[i - TT3_INTS] = reinterpret_cast<const uint32_t*>(a.P + r)[i - TT3_INTS];
    }
    __syncthreads();
    const RepParams& P = Ps;
    const double* Xa = STAGE == 1 ? a.X1 : a.X;
    const bool fast = tt[0] != 0;
    constexpr bool early = sizeof(FT) == 4;  // the FP64 kernels hold more registers
    T3Layout t;
    int n_own, li0 = 0;
    Head3 h0;
    h0.s = -1;
    if (fast) {
        t.in_lds = true;
        n_own = tt[2];
        const int total = tt[1];
        const int W = P.tile_x + 2, rows = W * W, nc = P.nc;
        if ((int)threadIdx.x < n_own) {
            int s0;
            table_owned3(tt, (int)threadIdx.x, s0, li0);
            load_head3(a, P, Xa, s0, true, h0);
        }
        // the halo: staged bead i lies in the last halo row whose LDS offset
        // is <= i, in the first run of that row or after its x wrap
        constexpr int KS = (T3_MAX_BEADS + NT - 1) / NT;
        double xv[KS][3];
        int cv[KS], rv[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            rv[k] = -1;
            if (i < total) {
                int lo = 0, hi = rows;  // row offset[lo] <= i < row offset[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (tt[TT3_HALO + 4 * mid + 3] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                const int* R = tt + TT3_HALO + 4 * lo;
                const int off = i - R[3];
                const int g = off < R[1] ? R[0] + off : R[2] + (off - R[1]);
                load_vec<3>(Xa, g, xv[k]);
                cv[k] = a.CELL[g];
                rv[k] = lo;
            }
        }
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            if (rv[k] < 0) continue;
            const int i = (int)threadIdx.x + k * NT;
            const int hz = rv[k] / W, hy = rv[k] - hz * W;
            const int gz = (tile.w - 1 + hz + nc) % nc;
            const int gy = (tile.z - 1 + hy + nc) % nc;
            const int gx = cv[k] - (gz * nc + gy) * nc;  // the cell the bead was binned into
            // halo column: x - (cx0 - 1), folded into [0, W) across the wrap
            int hx = gx - (tile.y - 1);
            hx += hx < 0 ? nc : (hx >= nc ? -nc : 0);
            double rel[3];
            rel[0] = cell_rel(xv[k][0], gx, P);
            rel[1] = cell_rel(xv[k][1], gy, P);
            rel[2] = cell_rel(xv[k][2], gz, P);
            lp[i] = make_lpos3<FT>(rel, hx, hy, hz);
        }
    } else {
        t = tile3_layout<NT>(a, P, tile, sh);
        n_own = sh.own_pref[t.own_y * t.own_z];
        if (t.in_lds) stage3<FT, NT>(a, P, Xa, sh, t, lp, nullptr, lcell);
    }
    __syncthreads();
    const PairConsts<FT> kc = pair_consts<FT>(P);
    const uint2* NL = reinterpret_cast<const uint2*>(a.NL);
    double fsq = 0.0;
    int ro = 0;
    for (int o = threadIdx.x; o < n_own; o += NT) {
        // each bead reads the parameters it needs from LDS again rather than
        // holding them in registers across the loop
        asm volatile("" ::: "memory");
        int li;
        Head3 hd;
        if (fast && o == (int)threadIdx.x) {  // loaded in the prologue
            li = li0;
            hd = h0;
        } else {
            int sx;
            if (fast) {
                table_owned3(tt, o, sx, li);
            } else {
                int c0;
                owned_bead3(sh, t, lcell, t.in_lds, o, ro, li, c0, sx);
            }
            load_head3(a, P, Xa, sx, t.in_lds, hd);
        }
        const int s = hd.s;
        const int2 me = hd.me;
        const double* xs = hd.x;
        double acc[3], tang[3];
        // SRK2 accumulates into F0: loaded after the pair loop, in flight
        // during the bonded terms (held across the pair loop, its registers
        // made the FP32 stage spill)
        const bool from_f0 = STAGE == 1 && !(GEN && (P.flags & GCR_F_SRK4));
        // the chain neighbours' positions: FP32, in flight during the pair
        // loop; FP64, after it (the registers)
        ChainNb<3> cn;
        cn.lk = hd.lk;
        auto load_nb = [&]() {
            if (cn.lk.x >= 0) load_vec<3>(Xa, cn.lk.x, cn.p2);
            if (cn.lk.y >= 0) load_vec<3>(Xa, cn.lk.y, cn.p1);
            if (cn.lk.z >= 0) load_vec<3>(Xa, cn.lk.z, cn.n1);
            if (cn.lk.w >= 0) load_vec<3>(Xa, cn.lk.w, cn.n2);
        };
        if (early) load_nb();
        FT f[3] = {(FT)0, (FT)0, (FT)0};
        if (P.flags & GCR_F_WCA) {
            if (hd.len >= 0)
                list3_pair_sum<FT>(lp, NL + s, a.nl_stride, hd.len, lp[li], kc, f);
            else
                pair_sum<3, FT>(Xa, a.META, a.CSTART, xs, me.x, me.y, a.CELL[s], P, f);
        }
        if (!early) load_nb();
        if (from_f0) load_vec<3>(a.F0, s, acc);
        if (!from_f0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[i] = 0.0;
        }
        // bonded terms first, then the pair sum, then propulsion (SRK.py:228-253)
        bool tang_set = false;
        if (P.flags & GCR_F_SPRINGS) bonded<3, FT, GEN>(cn, xs, a.CHAIN_A[me.y], P, acc, tang, tang_set);
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += (double)f[i];
        if (STAGE == 2) {
            if (!tang_set) load_vec<3>(a.T, s, tang);
            if (P.flags & GCR_F_PROPEL) propel<3>(a, me.y, tang, acc);
            if (GEN && (P.flags & GCR_F_WALLS)) add_walls<3>(P, xs, acc);
            store_vec<3>(a.FT, s, acc);
            store_vec<3>(a.Tn, s, tang);
            continue;
        }
        if (P.flags & GCR_F_PROPEL) {
            if (!tang_set) load_vec<3>(a.T, s, tang);
            propel<3>(a, me.y, tang, acc);
        }
        if (GEN && (P.flags & GCR_F_WALLS)) {  // at the step's start position in every stage
            double x0[3];
            load_vec<3>(a.X, s, x0);
            add_walls<3>(P, x0, acc);
        }
        if (GEN && (P.flags & (GCR_F_SLIDE | GCR_F_GROUP))) {  // tangents at this stage, for k_group_stage
            if (!tang_set) load_vec<3>(a.T, s, tang);
            store_vec<3>(a.Tn, s, tang);
        }
        const double q = stage_store<3, STAGE>(a, P, S, s, me.x, xs, acc);
        fsq = q > fsq ? q : fsq;
    }
    if (STAGE == 1 && a.sub == 3) {
        const double m = block_max<NT>(fsq);
        if (threadIdx.x == 0) a.BMAX[blockIdx.x] = m;
    }
}

}  // namespace gcr
