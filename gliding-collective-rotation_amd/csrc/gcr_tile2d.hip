// gcr_tile2d.hip -- 2D force stages over tiles of cells staged in LDS, with
// per-bead neighbour lists.
//
// One workgroup owns a tile of tile_x x tile_y cells of one replica. It stages
// the beads of the tile and of a one-cell halo into LDS in a fixed order
// (halo cells row-major, a cell's beads in slot order). That order depends on
// the cell contents only, so it is the same at every step between two cell
// re-sorts. At each re-sort, k_nlist_tile2d stores for every bead the LDS
// indices of the beads within rc + skin of it that are not excluded by the
// same-chain rule of updates.py:645-650 (a Verlet list in tile coordinates).
// The stage kernels then sum pair forces over the list only: about a third of
// the candidates of the 3 x 3 cell scan, no exclusion test, and every load
// from LDS. The list stays complete while no bead has moved more than half
// the skin since the re-sort; the stages check that and flag the replica for
// a replay otherwise.
//
// LDS coordinates: position relative to the centre of the bead's cell (FP32
// in the mixed form, FP64 in the plain one) plus the halo coordinates of that
// cell, so a separation is (dcell * h) + (dx), accurate to the ulp of a cell
// edge, and identical to the one the global cell scan forms.
// Bonded terms, propulsion and the SRK2 stage update are the same device
// functions as the 3D tile kernels; outputs are one write per bead. A bead whose list
// overflowed, and every bead of a tile that does not fit in LDS (or of a box
// of fewer than 3 cells per side), sums its pairs from global memory with
// pair_sum instead.
#define TILE_MAX_HALO 400    // (tile_x + 2) * (tile_y + 2) cells
#define TILE_MAX_BEADS 1536  // beads of a tile with its halo held in LDS
#define TILE_MAX_ROWS 16
#define NL_MAX 64            // neighbour-list entries per bead (16 groups of 4)
// Layout table of a tile (TT_INTS ints, written by the list builder at every
// re-sort): [0] bit 0 set when the stage kernels may use it (the tile fits in
// LDS and its halo rows and columns do not wrap onto themselves), bit 1 set
// when the processing order below is filled, [1] staged beads, [2] owned
// beads, [3] owned rows; from TT_HALO, per halo row (first slot of its first
// run, beads in that run, first slot of the run after the x wrap, LDS offset
// of the row), a halo row being one or two runs of slots; from TT_OWN, per
// owned row (first slot, beads, LDS index of its first bead); from TT_PERM,
// as u16, the owned bead each pass of the stage kernels' bead loop takes
// (TT_PERM_MAX beads at most): the beads sorted by list length, longest
// first, every other pass of NT threads reversed, so that the lanes of a
// wave walk lists of about one length and every wave gets a long and a
// short one.
#define TT_HALO 4
#define TT_OWN (TT_HALO + 4 * (TILE_MAX_ROWS + 2))
#define TT_PERM 128
#define TT_PERM_MAX 1024
#define TT_INTS (TT_PERM + TT_PERM_MAX / 2)
static_assert(TT_OWN + 3 * TILE_MAX_ROWS <= TT_PERM, "tile table size");

namespace gcr {

// LDS position record of a staged bead: position relative to the centre of
// the bead's cell (rounded to FT) and the halo coordinates of that cell
// (small integers, as floats). Separations are formed by cell_sep from these,
// as the global cell scan forms them, so the sums do not depend on the tile
// shape or on which path computed them.
template <typename FT>
struct LPos;

template <>
struct alignas(16) LPos<float> {
    float x, y;
    float cx, cy;
};

template <>
struct alignas(8) LPos<double> {
    double x, y;
    float cx, cy;
};

template <typename FT>
__device__ __forceinline__ LPos<FT> make_lpos(double rx, double ry, int hx, int hy)
{
    LPos<FT> q;
    q.x = (FT)rx;
    q.y = (FT)ry;
    q.cx = (float)hx;
    q.cy = (float)hy;
    return q;
}

// Separation q - me; the cell offsets are -1, 0 or +1 for every listed pair.
template <typename FT>
__device__ __forceinline__ void separation(const LPos<FT>& q, const LPos<FT>& me, FT h, FT* sep)
{
    sep[0] = cell_sep<FT>((FT)(q.cx - me.cx), h, q.x, me.x);
    sep[1] = cell_sep<FT>((FT)(q.cy - me.cy), h, q.y, me.y);
}

template <typename FT>
struct FT2 {
    FT x, y;
};

// Pair forces of one bead over its list (len entries, groups of 4 LDS
// indices at NL[g * stride + s]; the last group is padded with the bead's own
// index, which the zero-distance mask drops). Summed in FT. first: the first
// group, loaded with the bead's other inputs.
template <typename FT, bool SHIFTED>
__device__ __forceinline__ void list_loop(const LPos<FT>* lp, const uint2* G, int64_t stride, int ng,
                                          uint2 first, const LPos<FT>& me, const PairConsts<FT>& k, FT* f)
{
    uint32_t e0 = first.x, e1 = first.y;
    for (int g = 0; g < ng; ++g) {
        uint32_t n0 = e0, n1 = e1;
        if (g + 1 < ng) {  // next group in flight during this one
            const uint2 en = G[(int64_t)(g + 1) * stride];
            n0 = en.x;
            n1 = en.y;
        }
        // the four LDS reads issue together, then the four terms
        const LPos<FT> q[4] = {lp[e0 & 0xffffu], lp[e0 >> 16], lp[e1 & 0xffffu], lp[e1 >> 16]};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            FT sep[2];
            separation<FT>(q[u], me, k.h, sep);
            if (SHIFTED)
                pair_term_shifted<2, FT>(sep, k, f);
            else
                pair_term<2, FT>(sep, k, f);
        }
        e0 = n0;
        e1 = n1;
    }
}

template <typename FT>
__device__ __forceinline__ FT2<FT> list_pair_sum(const LPos<FT>* lp, const uint2* G, int64_t stride, int len,
                                                 uint2 first, const LPos<FT> me, const PairConsts<FT>& k)
{
    FT f[2] = {(FT)0, (FT)0};
    const int ng = (len + 3) >> 2;
    if (k.shifted)
        list_loop<FT, true>(lp, G, stride, ng, first, me, k, f);
    else
        list_loop<FT, false>(lp, G, stride, ng, first, me, k, f);
    return FT2<FT>{f[0], f[1]};
}

// Geometry of a tile and its halo in LDS, shared by the list builder and the
// stage kernels so that both see the same LDS order.
struct TileLayout {
    int cx0, cy0, W, nhalo, own_x, own_y, total;
    bool in_lds;
};

struct TileShared {
    int cst[TILE_MAX_HALO + 1];  // LDS offset of each halo cell
    int gst[TILE_MAX_HALO];      // first global slot of each halo cell
    int own_pref[TILE_MAX_ROWS + 1];
};

// Halo bead counts -> LDS offsets (two cells per thread), owned-bead row
// prefix; ends with a barrier.
template <int NT>
__device__ __forceinline__ TileLayout tile_layout(const PoolArgs& a, const RepParams& P, int4 tile,
                                                  TileShared& sh)
{
    TileLayout t;
    const int nc = P.nc, TX = P.tile_x, TY = P.tile_y;
    t.W = TX + 2;
    t.cx0 = tile.y;
    t.cy0 = tile.z;
    t.own_x = min(TX, nc - t.cx0);
    t.own_y = min(TY, nc - t.cy0);
    t.nhalo = t.W * (TY + 2);
    int cnt[2], g0[2], sum = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = threadIdx.x * 2 + k;
        cnt[k] = 0;
        g0[k] = 0;
        if (c < t.nhalo) {
            const int gx = (t.cx0 - 1 + c % t.W + nc) % nc;
            const int gy = (t.cy0 - 1 + c / t.W + nc) % nc;
            const int gc = P.cell_off + gy * nc + gx;
            g0[k] = a.CSTART[gc];
            cnt[k] = a.CSTART[gc + 1] - g0[k];
        }
        sum += cnt[k];
    }
    int ex = block_exclusive_scan<NT>(sum, &t.total);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = threadIdx.x * 2 + k;
        if (c < t.nhalo) {
            sh.cst[c] = ex;
            sh.gst[c] = g0[k];
        }
        ex += cnt[k];
    }
    if (threadIdx.x == 0) sh.cst[t.nhalo] = t.total;
    __syncthreads();
    if (threadIdx.x == 0) {
        // owned beads of tile row ry are the LDS range of cells 1..own_x of halo row ry+1
        int p = 0;
        for (int ry = 0; ry < t.own_y; ++ry) {
            sh.own_pref[ry] = p;
            const int row = (ry + 1) * t.W;
            p += sh.cst[row + t.own_x + 1] - sh.cst[row + 1];
        }
        sh.own_pref[t.own_y] = p;
    }
    t.in_lds = (P.flags & GCR_F_WCA) && t.total <= TILE_MAX_BEADS && nc >= 3;
    __syncthreads();
    return t;
}

// Halo cell of staged bead i: a binary search over the cells' LDS offsets.
__device__ __forceinline__ int staged_cell(const TileShared& sh, const TileLayout& t, int i)
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

// Stage the halo's positions (and, for the list builder, the original index), and
// the halo cell of every staged bead. FP32 positions: each thread first marks
// the LDS range of its cells and every staged bead reads its cell from there
// (one LDS read instead of a search). FP64 positions: the search, which keeps
// the FP64 stage kernels within 128 VGPRs (four waves per SIMD at 512 threads;
// the marking pass costs them 14 registers). Block-uniform call.
template <typename FT, int NT>
__device__ __forceinline__ void stage_positions(const PoolArgs& a, const RepParams& P, const double* Xa,
                                                const TileShared& sh, const TileLayout& t, LPos<FT>* lp,
                                                int* lm, uint16_t* lcell)
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
        const int hc = mark ? (int)lcell[i] : staged_cell(sh, t, i);
        if (!mark) lcell[i] = (uint16_t)hc;
        const int g = sh.gst[hc] + (i - sh.cst[hc]);
        const int hx = hc % t.W, hy = hc / t.W;
        const int gx = (t.cx0 - 1 + hx + nc) % nc;
        const int gy = (t.cy0 - 1 + hy + nc) % nc;
        const double2 x = *reinterpret_cast<const double2*>(Xa + 2 * g);
        lp[i] = make_lpos<FT>(cell_rel(x.x, gx, P), cell_rel(x.y, gy, P), hx, hy);
        if (lm) lm[i] = a.META[2 * g];
    }
}

// Owned bead o of the tile: its LDS index li, halo cell c0 and slot s. A
// thread visits its owned beads in increasing o, so the tile row ry carries
// over from the previous call (start at 0).
__device__ __forceinline__ void owned_bead(const TileShared& sh, const TileLayout& t, const uint16_t* lcell,
                                           int o, int& ry, int& li, int& c0, int& s)
{
    while (sh.own_pref[ry + 1] <= o) ++ry;
    const int row = (ry + 1) * t.W;
    li = sh.cst[row + 1] + (o - sh.own_pref[ry]);
    c0 = lcell[li];
    s = sh.gst[c0] + (li - sh.cst[c0]);
}

// The same without the staged cell table (tile outside LDS): the cell is
// found by scanning the row's cell offsets.
__device__ __forceinline__ void owned_bead_scan(const TileShared& sh, const TileLayout& t, int o, int& ry,
                                                int& li, int& c0, int& s)
{
    while (sh.own_pref[ry + 1] <= o) ++ry;
    const int row = (ry + 1) * t.W;
    li = sh.cst[row + 1] + (o - sh.own_pref[ry]);
    int hx = 1;
    while (sh.cst[row + hx + 1] <= li) ++hx;
    c0 = row + hx;
    s = sh.gst[c0] + (li - sh.cst[c0]);
}

// The tile's layout table (TT_INTS ints at TT[blockIdx.x]) from the layout
// just computed: halo row hy is halo cells hy W .. hy W + W - 1, whose slots
// are one run, or two where the row's x coordinates wrap round the box (the
// left halo cell at nc - 1, or the right ones from 0). Block-uniform call.
__device__ __forceinline__ bool write_tile_table(const PoolArgs& a, const RepParams& P, const TileLayout& t,
                                                 const TileShared& sh)
{
    int* T = a.TT + (size_t)blockIdx.x * TT_INTS;
    const int W = t.W, rows = P.tile_y + 2;
    const bool fast = t.in_lds && W <= P.nc && rows <= P.nc;
    if (threadIdx.x == 0) {
        T[0] = fast ? 1 : 0;
        T[1] = t.total;
        T[2] = sh.own_pref[t.own_y];
        T[3] = t.own_y;
    }
    if (!fast) return false;
    const int tid = (int)threadIdx.x;
    if (tid < rows) {
        const int c = tid * W, xa = t.cx0 - 1;
        // halo cells before the wrap: 1 if the left one wraps, nc - xa if the
        // right ones do, else the whole row
        const int k = xa < 0 ? 1 : (xa + W > P.nc ? P.nc - xa : W);
        int* R = T + TT_HALO + 4 * tid;
        R[0] = sh.gst[c];
        R[1] = sh.cst[c + k] - sh.cst[c];
        R[2] = k < W ? sh.gst[c + k] : 0;
        R[3] = sh.cst[c];
    }
    if (tid < t.own_y) {
        const int c = (tid + 1) * W + 1;  // owned cells 1 .. own_x of halo row tid + 1
        int* R = T + TT_OWN + 3 * tid;
        R[0] = sh.gst[c];
        R[1] = sh.cst[c + t.own_x] - sh.cst[c];
        R[2] = sh.cst[c];
    }
    return true;
}

// Neighbour lists of the tile's owned beads, after a cell re-sort (same run
// condition as the re-sort). Each bead scans the 3 x 3 cells around its own
// in LDS order (row by row; a row's three cells are one LDS range) and keeps the beads within rc + skin that the same-chain rule
// does not exclude; the list order is that scan order, so it depends on the
// positions alone. NLEN = -1 marks a bead without a list (overflow, or a
// tile outside LDS). The positions at the re-sort are kept in XR for the
// displacement check of the stages.
template <typename FT, int NT>
__global__ void __launch_bounds__(NT) k_nlist_tile2d(PoolArgs a, int j)
{
    const int4 tile = reinterpret_cast<const int4*>(a.TILES)[blockIdx.x];
    if (!rebuild_runs(a, tile.x, j)) return;
    if (!(a.P[tile.x].flags & GCR_F_WCA)) return;
    // the replica's parameters in LDS (reads of global memory after the
    // table's stores would each wait on memory)
    constexpr int PW = (int)(sizeof(RepParams) / 4);
    static_assert(PW <= NT, "one word per thread");
    __shared__ RepParams Ps;
    __shared__ TileShared sh;
    __shared__ LPos<FT> lp[TILE_MAX_BEADS];
    // the staged beads' original indices (the chain rule needs no more: a
    // chain's beads are one run of original indices, so a bead knows its
    // chain's run; keeping the chain per staged bead as well would hold the
    // FP32 kernel to three tiles per CU)
    __shared__ int lm[TILE_MAX_BEADS];
    __shared__ uint16_t lcell[TILE_MAX_BEADS];
    __shared__ uint8_t okey[TT_PERM_MAX];     // list groups of each owned bead (17: no list)
    __shared__ uint16_t osort[TT_PERM_MAX];   // owned beads by list length, longest first
    __shared__ int obin[NL_MAX / 4 + 2];
    if ((int)threadIdx.x < PW)
        reinterpret_cast<uint32_t*>(&Ps)[threadIdx.x] = reinterpret_cast<const uint32_t*>(a.P + tile.x)[threadIdx.x];
    __syncthreads();
    const RepParams& P = Ps;
    const TileLayout t = tile_layout<NT>(a, P, tile, sh);
    const bool write_ok = write_tile_table(a, P, t, sh);
    if (t.in_lds) stage_positions<FT, NT>(a, P, a.on_x1 ? a.X1 : a.X, sh, t, lp, lm, lcell);
    __syncthreads();
    const FT rl2 = (FT)P.rl2, h = (FT)P.h;
    const int nn = P.n_nearby;
    const int n_own = sh.own_pref[t.own_y];
    uint2* NL = reinterpret_cast<uint2*>(a.NL);
    int ry = 0;
    for (int o = threadIdx.x; o < n_own; o += NT) {
        int li, c0, s;
        if (!t.in_lds) {
            owned_bead_scan(sh, t, o, ry, li, c0, s);
            a.NLEN[s] = -1;
            continue;
        }
        owned_bead(sh, t, lcell, o, ry, li, c0, s);
        const LPos<FT> me = lp[li];
        const int2 mm = *reinterpret_cast<const int2*>(a.META + 2 * s);
        const int cf = a.CHAIN_FIRST[mm.y], cl = a.CHAIN_LEN[mm.y];
        // the original indices the list leaves out, one run: the chain's
        // beads within n_nearby of this one (itself included), or the whole
        // chain for n_nearby < 0 (updates.py:645-650)
        int ex_lo = cf, ex_hi = cf + cl;
        if (nn >= 0) {
            const int w = nn < cl ? nn : cl;
            ex_lo = max(cf, mm.x - w);
            ex_hi = min(cf + cl, mm.x + w + 1);
        }
        const uint32_t ex_n = (uint32_t)(ex_hi - ex_lo);
        int n = 0;
        // the group being filled, as one 64-bit word into which each entry
        // shifts from the top: after four the first sits in the low half of
        // w0 and the fourth in the high half of w1
        uint32_t w0 = 0u, w1 = 0u;
        // the three cells of a halo row are consecutive in LDS: one range per row
        for (int dy = -1; dy <= 1; ++dy) {
            const int c = c0 + dy * t.W;
            const int q1 = sh.cst[c + 2];
            for (int q = sh.cst[c - 1]; q < q1; ++q) {
                FT sep[2];
                separation<FT>(lp[q], me, h, sep);
                const bool excl = (uint32_t)(lm[q] - ex_lo) < ex_n;
                const bool ok = (fma(sep[1], sep[1], sep[0] * sep[0]) <= rl2) & !excl;
                const bool put = ok & (n < NL_MAX);
                const uint32_t s0 = __builtin_amdgcn_alignbit(w1, w0, 16);
                const uint32_t s1 = __builtin_amdgcn_alignbit((uint32_t)q, w1, 16);
                w0 = put ? s0 : w0;
                w1 = put ? s1 : w1;
                if (put & ((n & 3) == 3)) NL[(int64_t)(n >> 2) * a.nl_stride + s] = make_uint2(w0, w1);
                n += ok ? 1 : 0;
            }
        }
        if (n <= NL_MAX) {
            if (n & 3) {  // pad the last group with the bead's own index
                for (int sl = n & 3; sl < 4; ++sl) {
                    w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
                    w1 = __builtin_amdgcn_alignbit((uint32_t)li, w1, 16);
                }
                NL[(int64_t)(n >> 2) * a.nl_stride + s] = make_uint2(w0, w1);
            }
            a.NLEN[s] = n;
        } else {
            a.NLEN[s] = -1;
        }
        if (o < TT_PERM_MAX) okey[o] = (uint8_t)(n <= NL_MAX ? (n + 3) >> 2 : NL_MAX / 4 + 1);
    }
    // The processing order of the stage kernels' bead loop (a.nt_stage
    // threads per tile): a counting sort of the owned beads by list groups,
    // longest first (which bead of a key comes first does not matter: the
    // order decides only which thread computes a bead, not its result). Only
    // for lists of three groups or more on average: the sorted order spreads
    // a wave's beads over the tile, so its loads of lists and positions touch
    // more cache lines, which short lists do not win back.
    if (threadIdx.x == 0) a.TILES[4 * (size_t)blockIdx.x + 3] = 0;  // the table without an order
    if (!write_ok || n_own > TT_PERM_MAX || a.nt_stage <= 0) return;
    constexpr int NK = NL_MAX / 4 + 2;
    for (int k = threadIdx.x; k < NK; k += NT) obin[k] = 0;
    __syncthreads();
    for (int o = threadIdx.x; o < n_own; o += NT) atomicAdd(&obin[okey[o]], 1);
    __syncthreads();
    {
        int groups = 0;
        for (int k = 1; k < NK; ++k) groups += k * obin[k];
        if (groups < 3 * n_own) return;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = NK - 1; k >= 0; --k) {
            const int c = obin[k];
            obin[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int o = threadIdx.x; o < n_own; o += NT) osort[atomicAdd(&obin[okey[o]], 1)] = (uint16_t)o;
    __syncthreads();
    // pass k of the stage loop takes owned positions [k NTS, k NTS + m_k):
    // in order on even passes, reversed on odd ones
    const int NTS = a.nt_stage;
    uint16_t* perm = reinterpret_cast<uint16_t*>(a.TT + (size_t)blockIdx.x * TT_INTS + TT_PERM);
    for (int o = threadIdx.x; o < n_own; o += NT) {
        const int k = o / NTS, tk = o - k * NTS, mk = min(NTS, n_own - k * NTS);
        perm[o] = osort[(k & 1) ? k * NTS + (mk - 1 - tk) : o];
    }
    if (threadIdx.x == 0) {
        a.TT[(size_t)blockIdx.x * TT_INTS] = 3;
        a.TILES[4 * (size_t)blockIdx.x + 3] = TT_PERM + (n_own + 1) / 2;
    }
}

// The inputs of one owned bead that come from global memory. The head
// (position, (orig, chain), list length and first group, chain links) is
// what the pair loop needs and loads in the prologue; the rest (the chain
// neighbours' positions, and in the lean variant F0 or the noise and the
// re-sort position) loads when the bead's turn comes, in flight during its
// pair loop.
struct BeadHead {
    int s;
    double x[2];
    int2 me;
    int len;
    uint2 g0;  // first list group
    int4 lk;   // chain links (lean variant)
};

template <bool GEN>
__device__ __forceinline__ void load_head(const PoolArgs& a, const RepParams& P, const double* Xa, int s,
                                          BeadHead& b)
{
    b.s = s;
    load_vec<2>(Xa, s, b.x);
    b.me = *reinterpret_cast<const int2*>(a.META + 2 * s);
    b.len = -1;
    b.g0 = make_uint2(0u, 0u);
    if (P.flags & GCR_F_WCA) {
        b.len = a.NLEN[s];
        b.g0 = reinterpret_cast<const uint2*>(a.NL)[s];
    }
    b.lk = (!GEN && (P.flags & GCR_F_SPRINGS)) ? *reinterpret_cast<const int4*>(a.LINK + 4 * s)
                                               : make_int4(-1, -1, -1, -1);
}

template <int STAGE, bool GEN>
struct BeadRest {
    double pre[2];  // lean stage 1: F0 (SRK2 accumulates into it); lean stage 0: XI
    double xr[2];   // lean stage 0: position at the last re-sort
    ChainNb<2> cn;  // lean: chain neighbours
};

// F0 or the noise and the re-sort position (lean variant).
template <int STAGE, bool GEN>
__device__ __forceinline__ void load_rest_state(const PoolArgs& a, int s, BeadRest<STAGE, GEN>& b)
{
    if (GEN) return;
    if (STAGE == 1) load_vec<2>(a.F0, s, b.pre);
    if (STAGE == 0) {
        load_vec<2>(a.XI, s, b.pre);
        load_vec<2>(a.XR, s, b.xr);
    }
}

// The chain neighbours' positions (lean variant).
template <int STAGE, bool GEN>
__device__ __forceinline__ void load_rest(const PoolArgs& a, const double* Xa, const BeadHead& h,
                                          BeadRest<STAGE, GEN>& b)
{
    if (GEN) return;
    const int4 lk = h.lk;
    b.cn.lk = lk;
    if (lk.x >= 0) load_vec<2>(Xa, lk.x, b.cn.p2);
    if (lk.y >= 0) load_vec<2>(Xa, lk.y, b.cn.p1);
    if (lk.z >= 0) load_vec<2>(Xa, lk.z, b.cn.n1);
    if (lk.w >= 0) load_vec<2>(Xa, lk.w, b.cn.n2);
}

// Owned bead o of a tile from its layout table: (slot, LDS index), or s = -1.
__device__ __forceinline__ void table_owned(const int* tt, int o, int& s, int& li)
{
    s = -1;
    li = 0;
    int before = 0;
    const int own_y = tt[3];
    for (int ry = 0; ry < own_y; ++ry) {
        const int* R = tt + TT_OWN + 3 * ry;
        if (o < before + R[1]) {
            s = R[0] + (o - before);
            li = R[2] + (o - before);
            return;
        }
        before += R[1];
    }
}

// STAGE 0 / 1: the SRK force stages (STAGE 1 for SRK stage a.sub); STAGE 2: force
// evaluation of replica `j` only at the positions X (forces into FT,
// tangents into Tn; no noise, no update). NT threads per tile (256, 512 or
// 1024): large tiles keep more waves in flight. GEN = false is the lean
// variant for pools without SRK4, walls and FENE springs: their code stays
// out of the kernel, and with it the registers it would hold. Every variant
// keeps four waves per SIMD (at most 128 VGPRs): with tiles of 256 threads
// that holds four tiles per CU, and the 400k-bead system's 1024 tiles run
// at once.
// Prologue: with the tile's layout table (written at the last re-sort) each
// thread finds its first two owned beads and its share of the halo straight
// from the table and issues all their loads at once; without one (a tile
// whose table is not filled yet, or that the table does not describe) the
// layout is computed here (tile_layout) and the halo staged after it.
template <typename FT, int STAGE, int NT, bool GEN>
__global__ void __launch_bounds__(NT, 4) k_stage_tile2d(PoolArgs a, int j)
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
    // The replica's parameters and the predictor's dt live in LDS for the
    // whole kernel: read from global memory after the kernel's first store,
    // which could alias them, each would be a vector load waiting on memory.
    constexpr int PW = (int)(sizeof(RepParams) / 4);
    static_assert(sizeof(RepParams) % 4 == 0, "parameters copied by words");
    __shared__ RepParams Ps;
    __shared__ double dt_s;
    __shared__ TileShared sh;
    __shared__ LPos<FT> lp[TILE_MAX_BEADS];
    __shared__ uint16_t lcell[TILE_MAX_BEADS];
    __shared__ int tt[TT_INTS];
    // the table's head and rows, and the bead order where the list builder
    // wrote one (tile.w ints in all)
    const int tlen = tile.w > TT_PERM ? (tile.w < TT_INTS ? tile.w : TT_INTS) : TT_PERM;
    for (int i = threadIdx.x; i <= tlen + PW; i += NT) {
        if (i < tlen)
            tt[i] = a.TT[(size_t)blockIdx.x * TT_INTS + i];
        else if (i < tlen + PW)
            reinterpret_cast<uint32_t*>(&Ps)[i - tlen] = reinterpret_cast<const uint32_t*>(a.P + r)[i - tlen];
        else
            dt_s = S->dt;
    }
    __syncthreads();
    const RepParams& P = Ps;
    const double* Xa = STAGE == 1 ? a.X1 : a.X;
    const bool springs = (P.flags & GCR_F_SPRINGS) != 0;
    const bool fast = (tt[0] & 1) != 0;
    // the owned bead a position of the bead loop takes (the list builder's
    // order when the table holds one)
    const uint16_t* perm = ((tt[0] & 2) && tlen > TT_PERM) ? reinterpret_cast<const uint16_t*>(tt + TT_PERM) : nullptr;
    constexpr bool early = sizeof(FT) == 4;  // the FP64 kernels have no registers to spare
    TileLayout t;
    int n_own;
    BeadHead h0, h1;
    int li0 = 0, li1 = 0, c0_unused = 0;
    h0.s = h1.s = -1;
    if (fast) {
        t.in_lds = true;
        t.W = P.tile_x + 2;
        n_own = tt[2];
        const int total = tt[1];
        if (early && (int)threadIdx.x < n_own) {
            int s0, s1;
            const int o0 = (int)threadIdx.x, o1 = o0 + NT;
            table_owned(tt, perm ? (int)perm[o0] : o0, s0, li0);
            load_head<GEN>(a, P, Xa, s0, h0);
            if (o1 < n_own) {
                table_owned(tt, perm ? (int)perm[o1] : o1, s1, li1);
                load_head<GEN>(a, P, Xa, s1, h1);
            }
        }
        // the halo: staged bead i lies in the last halo row whose LDS offset
        // is <= i, in the first run of that row or after its x wrap
        constexpr int KS = (TILE_MAX_BEADS + NT - 1) / NT;
        const int rows = P.tile_y + 2;
        double2 xv[KS];
        int cv[KS], hv[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            hv[k] = -1;
            if (i < total) {
                int hy = 0;
                for (int rr = 1; rr < rows; ++rr) hy += tt[TT_HALO + 4 * rr + 3] <= i ? 1 : 0;
                const int* R = tt + TT_HALO + 4 * hy;
                const int off = i - R[3];
                const int g = off < R[1] ? R[0] + off : R[2] + (off - R[1]);
                xv[k] = *reinterpret_cast<const double2*>(Xa + 2 * g);
                cv[k] = a.CELL[g];
                hv[k] = hy;
            }
        }
        const int nc = P.nc;
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            if (hv[k] < 0) continue;
            const int i = (int)threadIdx.x + k * NT;
            const int gy = (tile.z - 1 + hv[k] + nc) % nc;
            const int gx = cv[k] - gy * nc;  // the cell the bead was binned into
            // halo column: x - (cx0 - 1), folded into [0, W) across the wrap
            // (left: the column at nc - 1; right: the columns from 0)
            int hx = gx - (tile.y - 1);
            hx += hx < 0 ? nc : (hx >= nc ? -nc : 0);
            lp[i] = make_lpos<FT>(cell_rel(xv[k].x, gx, P), cell_rel(xv[k].y, gy, P), hx, hv[k]);
        }
    } else {
        t = tile_layout<NT>(a, P, tile, sh);
        n_own = sh.own_pref[t.own_y];
        // the first two owned beads of this thread: slots from the layout (the
        // staged cell table lcell is not filled yet), their heads in flight
        // while the halo is staged
        int ry = 0;
        if (early && (int)threadIdx.x < n_own) {
            int s0;
            owned_bead_scan(sh, t, (int)threadIdx.x, ry, li0, c0_unused, s0);
            load_head<GEN>(a, P, Xa, s0, h0);
            if ((int)threadIdx.x + NT < n_own) {
                int ry1 = ry, s1;
                owned_bead_scan(sh, t, (int)threadIdx.x + NT, ry1, li1, c0_unused, s1);
                load_head<GEN>(a, P, Xa, s1, h1);
            }
        }
        if (t.in_lds) stage_positions<FT, NT>(a, P, Xa, sh, t, lp, nullptr, lcell);
    }
    __syncthreads();

    const PairConsts<FT> kc = pair_consts<FT>(P);
    const uint2* NL = reinterpret_cast<const uint2*>(a.NL);
    double fsq = 0.0;
    int ry = 0;
    for (int o = threadIdx.x; o < n_own; o += NT) {
        // each bead reads the parameters it needs from LDS again rather than
        // holding them in registers across the loop
        asm volatile("" ::: "memory");
        int li, s;
        BeadHead bd;
        if (early && o == (int)threadIdx.x) {  // heads loaded in the prologue
            li = li0;
            bd = h0;
        } else if (early && o == (int)threadIdx.x + NT) {
            li = li1;
            bd = h1;
        } else {
            int sx, c0;
            if (fast)
                table_owned(tt, perm ? (int)perm[o] : o, sx, li);
            else if (t.in_lds)
                owned_bead(sh, t, lcell, o, ry, li, c0, sx);
            else
                owned_bead_scan(sh, t, o, ry, li, c0, sx);
            load_head<GEN>(a, P, Xa, sx, bd);
        }
        s = bd.s;
        BeadRest<STAGE, GEN> br;
        load_rest<STAGE, GEN>(a, Xa, bd, br);
        // FP32 positions: in flight during the pair loop; FP64: after it (the
        // registers)
        if (early) load_rest_state<STAGE, GEN>(a, s, br);
        const double* xs = bd.x;
        const int2 me = bd.me;

        double acc[2], tang[2];
        double pa0 = 0.0, pa1 = 0.0;
        if (P.flags & GCR_F_WCA) {
            if (bd.len >= 0) {
                const LPos<FT> mp = lp[li];
                const FT2<FT> f = list_pair_sum<FT>(lp, NL + s, a.nl_stride, bd.len, bd.g0, mp, kc);
                pa0 = (double)f.x;
                pa1 = (double)f.y;
            } else {
                FT pg[2] = {(FT)0, (FT)0};
                pair_sum<2, FT>(Xa, a.META, a.CSTART, xs, me.x, me.y, a.CELL[s], P, pg);
                pa0 = (double)pg[0];
                pa1 = (double)pg[1];
            }
        }
        if (!early) load_rest_state<STAGE, GEN>(a, s, br);
        if (STAGE == 1 && !(GEN && (P.flags & GCR_F_SRK4))) {
            // SRK2 accumulates into F0
            if (GEN) {
                load_vec<2>(a.F0, s, acc);
            } else {
                acc[0] = br.pre[0];
                acc[1] = br.pre[1];
            }
        } else {
            acc[0] = 0.0;
            acc[1] = 0.0;
        }
        // bonded terms first, then the pair sum, then propulsion (SRK.py:228-253)
        bool tang_set = false;
        if (springs) {
            // the general variant loads the bonded neighbours after the pair
            // loop (its registers)
            if (GEN) load_chain_nb<2>(Xa, a.LINK, s, br.cn);
            bonded<2, FT, GEN>(br.cn, xs, a.CHAIN_A[me.y], P, acc, tang, tang_set);
        }
        acc[0] += pa0;
        acc[1] += pa1;
        if (STAGE == 2) {
            if (!tang_set) load_vec<2>(a.T, s, tang);
            if (P.flags & GCR_F_PROPEL) propel<2>(a, me.y, tang, acc);
            if (GEN && (P.flags & GCR_F_WALLS)) add_walls<2>(P, xs, acc);
            store_vec<2>(a.FT, s, acc);
            store_vec<2>(a.Tn, s, tang);
        } else {
            if (P.flags & GCR_F_PROPEL) {
                if (!tang_set) load_vec<2>(a.T, s, tang);
                propel<2>(a, me.y, tang, acc);
            }
            if (GEN && (P.flags & GCR_F_WALLS)) {  // at the step's start position in every stage
                double x0[2];
                load_vec<2>(a.X, s, x0);
                add_walls<2>(P, x0, acc);
            }
            if (GEN && (P.flags & (GCR_F_SLIDE | GCR_F_GROUP))) {  // tangents at this stage, for k_group_stage
                if (!tang_set) load_vec<2>(a.T, s, tang);
                store_vec<2>(a.Tn, s, tang);
            }
            if (!GEN && STAGE == 0) {  // stage_store's lean stage 0 with the preloaded noise
                store_vec<2>(a.F0, s, acc);
                virtual_update_in<2>(a, P, S, s, xs, acc, 1.0, 0, br.pre, br.xr, dt_s);
            } else {
                const double q = stage_store<2, STAGE, GEN>(a, P, S, s, me.x, xs, acc);
                fsq = q > fsq ? q : fsq;
            }
        }
    }
    if (STAGE == 1 && a.sub == 3) {
        const double m = block_max<NT>(fsq);
        if (threadIdx.x == 0) a.BMAX[blockIdx.x] = m;
    }
}

}  // namespace gcr
