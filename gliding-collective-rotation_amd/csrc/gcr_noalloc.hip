// gcr_noalloc.hip -- a force evaluation that allocates no device memory
// (gcr_eval_forces_noalloc), for callers that cannot call hipMalloc: code
// under HIP graph capture, serving loops that forbid allocation.
//
// It runs the kernels of gcr_eval_forces -- the cell re-sort (k_bin, the
// three scan phases, k_scatter, k_gather, k_links) and the per-bead force
// evaluation -- on a one-replica layout whose arrays are all carved out of
// one file-scope device array, `scratch`, of fixed size. Each device holds
// its own copy of the array until the process ends; its address is taken
// with hipGetSymbolAddress on the first call on each device and kept per
// device (the array's name in host code is not a device address). Every
// call on a device shares that device's array and overwrites it: calls that
// overlap on one device race.
//
// Included by gcr_capi.hip (it uses that file's fail(), set_physics() and
// grid_side()).
#include <mutex>

namespace gcr {

constexpr int NA_MAX_BEADS = GCR_NOALLOC_MAX_BEADS;
constexpr int NA_MAX_CELLS = GCR_NOALLOC_MAX_CELLS;
constexpr int NA_MAX_CHAINS = NA_MAX_BEADS;
constexpr int NA_MAX_BLK = (NA_MAX_BEADS + BS - 1) / BS;
constexpr int NA_MAX_CHK = (NA_MAX_CELLS + 1 + SCAN_CHUNK - 1) / SCAN_CHUNK;

// Byte offsets of the arrays inside `scratch` (every array 256-byte aligned,
// sized for 3D, the larger layout).
struct NaLayout {
    size_t P, S, blk_rep, blk_slot0, chk_rep, chk_cell0, rep_chk_off;
    size_t X, X1, T, Xn, Tn, X1n, FT, XR, XI, XIn;
    size_t META, METAn, KEYOFF, PAIRS, LINK, CELL, SLOT, NLEN;
    size_t CHAIN_A, CHAIN_CS, CHAIN_FIRST, CHAIN_LEN, HEADS, ENDS;
    size_t CCOUNT, CSTART, CHKSUM;
    size_t total;
};

constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr NaLayout na_layout()
{
    NaLayout o{};
    size_t at = 0;
    auto take = [&at](size_t bytes) {
        const size_t here = at;
        at = na_align(at + bytes);
        return here;
    };
    const size_t nb = NA_MAX_BEADS, vec = nb * 3 * sizeof(double);
    o.P = take(sizeof(RepParams));
    o.S = take(sizeof(RepState));
    o.blk_rep = take(NA_MAX_BLK * 4);
    o.blk_slot0 = take(NA_MAX_BLK * 4);
    o.chk_rep = take(NA_MAX_CHK * 4);
    o.chk_cell0 = take(NA_MAX_CHK * 4);
    o.rep_chk_off = take(2 * 4);
    o.X = take(vec);
    o.X1 = take(vec);
    o.T = take(vec);
    o.Xn = take(vec);
    o.Tn = take(vec);
    o.X1n = take(vec);
    o.FT = take(vec);
    o.XR = take(vec);
    o.XI = take(2 * vec);   // translational then rotational noise
    o.XIn = take(2 * vec);
    o.META = take(nb * 8);
    o.METAn = take(nb * 8);
    o.KEYOFF = take(nb * 8);
    o.PAIRS = take(nb * 8);
    o.LINK = take(nb * 16);
    o.CELL = take(nb * 4);
    o.SLOT = take(nb * 4);
    o.NLEN = take(nb * 4);
    o.CHAIN_A = take(NA_MAX_CHAINS * 8);
    o.CHAIN_CS = take(NA_MAX_CHAINS * 24);
    o.CHAIN_FIRST = take(NA_MAX_CHAINS * 4);
    o.CHAIN_LEN = take(NA_MAX_CHAINS * 4);
    o.HEADS = take(NA_MAX_CHAINS * 4);
    o.ENDS = take(NA_MAX_CHAINS * 4);
    o.CCOUNT = take((NA_MAX_CELLS + 1) * 4);
    o.CSTART = take((NA_MAX_CELLS + 1) * 4);
    o.CHKSUM = take(NA_MAX_CHK * 4);
    o.total = at;
    return o;
}

constexpr NaLayout NA = na_layout();
static_assert(NA.total <= GCR_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the stated size");

}  // namespace gcr

// The scratch array of gcr_eval_forces_noalloc (one per device).
static __device__ __align__(256) unsigned char scratch[GCR_NOALLOC_SCRATCH_BYTES];

namespace gcr {

// Pointers of the one-replica layout, writable (the PoolArgs tables are const).
struct NaBufs {
    int32_t *blk_rep, *blk_slot0, *chk_rep, *chk_cell0, *rep_chk_off;
    int32_t *CHAIN_FIRST, *CHAIN_LEN, *HEADS, *ENDS;
    double *CHAIN_A, *CHAIN_CS;
};

// Parameters and state of the replica, the block and chunk tables, and the
// counters that must start at zero (cell counts, runs per chain id).
__global__ void __launch_bounds__(BS) k_na_setup(PoolArgs a, NaBufs b, RepParams P, int nblk, int nchk)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i == 0) {
        *const_cast<RepParams*>(a.P) = P;
        RepState S{};
        S.max_steps = 0;
        S.mdu = -1.0;
        *a.S = S;  // frozen = 0: the re-sort kernels run (j = -1)
        b.rep_chk_off[0] = 0;
        b.rep_chk_off[1] = nchk;
    }
    if (i < nblk) {
        b.blk_rep[i] = 0;
        b.blk_slot0[i] = i * BS;
    }
    if (i < nchk) {
        b.chk_rep[i] = 0;
        b.chk_cell0[i] = i * SCAN_CHUNK;
    }
    if (i <= P.ncell) a.CCOUNT[i] = 0;
    if (i < P.n_chains) b.HEADS[i] = 0;
}

// The input beads in slot order = bead order: positions (also as X1),
// propulsion directions, (bead, chain), no list; the first and one-past-last
// bead of each run of equal chain_id. Input gcr_replica_add would reject (a
// position outside [0, L], a chain id outside [0, n_chains)) freezes the
// replica: nothing else runs and the outputs are NaN.
template <int DIM>
__global__ void __launch_bounds__(BS) k_na_beads(PoolArgs a, NaBufs b, const double* pos, const double* tang,
                                                 const int64_t* chain_id)
{
    const RepParams& P = *a.P;
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= P.n) return;
    double x[DIM], t[DIM];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        x[k] = pos[(size_t)i * DIM + k];
        bad = bad || !(x[k] >= 0.0 && x[k] <= P.L);
        t[k] = tang ? tang[(size_t)i * DIM + k] : 0.0;
    }
    store_vec<DIM>(a.X, i, x);
    store_vec<DIM>(a.X1, i, x);
    store_vec<DIM>(a.T, i, t);
    const int64_t cid = chain_id[i];
    const bool cid_ok = cid >= 0 && cid < P.n_chains;
    const int c = cid_ok ? (int)cid : 0;
    *reinterpret_cast<int2*>(a.META + 2 * i) = make_int2(i, c);
    a.SLOT[i] = i;
    a.NLEN[i] = -1;  // pairs by the cell scan
    if (cid_ok) {
        if (i == 0 || chain_id[i - 1] != cid) {
            b.CHAIN_FIRST[c] = i;
            atomicAdd(&b.HEADS[c], 1);
        }
        if (i + 1 == P.n || chain_id[i + 1] != cid) b.ENDS[c] = i + 1;
    }
    if (bad || !cid_ok) a.S->frozen = 1;
}

// Per chain: length, rest length and (v0, cos, sin) of the chiral angle, as
// gcr_replica_add sets them. A chain id without beads or split into several
// runs freezes the replica.
__global__ void __launch_bounds__(BS) k_na_chains(PoolArgs a, NaBufs b, const double* rest, const double* vtan,
                                                  double v0, int dim)
{
    const RepParams& P = *a.P;
    const int c = blockIdx.x * BS + threadIdx.x;
    if (c >= P.n_chains) return;
    if (b.HEADS[c] != 1) {
        a.S->frozen = 1;
        b.CHAIN_LEN[c] = 0;
        return;
    }
    b.CHAIN_LEN[c] = b.ENDS[c] - b.CHAIN_FIRST[c];
    b.CHAIN_A[c] = rest ? rest[c] : 0.0;
    const double ang = vtan ? vtan[c] : 0.0;
    double* cs = b.CHAIN_CS + 3 * c;
    cs[0] = v0;
    if (ang == 0.0 || dim != 2) {  // forces.py:214-217: plain v0 * t
        cs[1] = 1.0;
        cs[2] = 0.0;
    } else {
        cs[1] = cos(ang);
        cs[2] = sin(ang);
    }
}

// The force evaluation of gcr_eval_forces (eval_bead) when the input passed.
template <int DIM, typename FT>
__global__ void __launch_bounds__(BS) k_na_eval(PoolArgs a)
{
    if (a.S->frozen) return;
    const RepParams& P = *a.P;
    const int s = blockIdx.x * BS + threadIdx.x;
    if (s < P.n) eval_bead<DIM, FT>(a, P, s, 0);
}

// Forces and tangents from slots back into bead order; NaN for input that
// froze the replica.
template <int DIM>
__global__ void __launch_bounds__(BS) k_na_out(PoolArgs a, double* forces, double* tang_out)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= a.P->n) return;
    const bool bad = a.S->frozen != 0;
    const int s = bad ? 0 : a.SLOT[i];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        forces[(size_t)i * DIM + k] = bad ? __builtin_nan("") : a.FT[(size_t)s * DIM + k];
        if (tang_out) tang_out[(size_t)i * DIM + k] = bad ? __builtin_nan("") : a.Tn[(size_t)s * DIM + k];
    }
}

}  // namespace gcr

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
void* g_na_scratch[NA_MAX_DEVICES] = {};

// Device address of the current device's scratch array, taken once per
// device with that device current.
int na_scratch(unsigned char** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(GCR_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GCR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_scratch[dev] = ptr;
    }
    *out = static_cast<unsigned char*>(g_na_scratch[dev]);
    return GCR_OK;
}

template <typename T>
T* na_at(unsigned char* base, size_t off)
{
    return reinterpret_cast<T*>(base + off);
}

template <int DIM>
int na_launch(const PoolArgs& a, const NaBufs& b, const RepParams& P, int precision, const double* pos,
              const double* tang, const int64_t* chain_id, const double* rest, const double* vtan, double v0,
              double* forces, double* tang_out, hipStream_t st)
{
    const int nblk = (P.n + BS - 1) / BS;
    const int nchk = (P.ncell + SCAN_CHUNK - 1) / SCAN_CHUNK;
    const int nset = std::max(std::max(P.ncell + 1, P.n_chains), std::max(nblk, nchk));
    k_na_setup<<<(nset + BS - 1) / BS, BS, 0, st>>>(a, b, P, nblk, nchk);
    k_na_beads<DIM><<<nblk, BS, 0, st>>>(a, b, pos, tang, chain_id);
    k_na_chains<<<(P.n_chains + BS - 1) / BS, BS, 0, st>>>(a, b, rest, vtan, v0, DIM);
    // the cell re-sort of gcr_eval_forces; its kernels skip a frozen replica
    k_bin<DIM><<<nblk, BS, 0, st>>>(a, -1);
    k_scan_chunks<<<nchk, BS, 0, st>>>(a, -1);
    k_scan_reps<<<1, BS, 0, st>>>(a, -1);
    k_scan_final<<<nchk, BS, 0, st>>>(a, -1);
    k_scatter<<<nblk, BS, 0, st>>>(a, -1);
    k_gather<DIM><<<nblk, BS, 0, st>>>(a, 1, -1);
    k_links<DIM><<<nblk, BS, 0, st>>>(a, 1, -1);
    if (precision == GCR_MIXED)
        k_na_eval<DIM, float><<<nblk, BS, 0, st>>>(a);
    else
        k_na_eval<DIM, double><<<nblk, BS, 0, st>>>(a);
    k_na_out<DIM><<<nblk, BS, 0, st>>>(a, forces, tang_out);
    HIPCHK(hipGetLastError());
    return GCR_OK;
}

}  // namespace

extern "C" int gcr_eval_forces_noalloc(const gcr_replica_desc* d, int32_t dim, int32_t precision,
                                       const double* pos, const double* tangent, const int64_t* chain_id,
                                       const double* rest_lengths, const double* v_tan_angle,
                                       double* forces, double* tangent_out, void* stream)
{
    if (!d || !pos || !chain_id || !forces) return fail(GCR_ERR_ARG, "null argument");
    if (dim != 2 && dim != 3) return fail(GCR_ERR_ARG, "dim must be 2 or 3");
    if (precision != GCR_FP64 && precision != GCR_MIXED) return fail(GCR_ERR_ARG, "bad precision");
    const int64_t n = d->n_beads, nch = d->n_chains;
    if (n < 1 || nch < 1 || nch > n) return fail(GCR_ERR_ARG, "need 1 <= n_chains <= n_beads");
    if (!(d->L > 0) || !(d->radius > 0)) return fail(GCR_ERR_ARG, "L and radius must be positive");
    if (d->has_wca && !(d->wca[3] > 0 && 2 * d->wca[3] < d->L))
        return fail(GCR_ERR_ARG, "WCA cutoff must be in (0, L/2)");
    if (n > NA_MAX_BEADS)
        return fail(GCR_ERR_CAPACITY, "gcr_eval_forces_noalloc: more than GCR_NOALLOC_MAX_BEADS beads");
    const int64_t nc = grid_side(dim, d);
    const int64_t ncell = dim == 3 ? nc * nc * nc : nc * nc;
    if (ncell > NA_MAX_CELLS)
        return fail(GCR_ERR_CAPACITY, "gcr_eval_forces_noalloc: the cell grid exceeds GCR_NOALLOC_MAX_CELLS");
    unsigned char* base = nullptr;
    int rc = na_scratch(&base);
    if (rc) return rc;

    // the replica's parameters, as gcr_replica_add sets them for a pool's
    // first replica
    RepParams P{};
    P.n = (int32_t)n;
    P.n_chains = (int32_t)nch;
    P.nc = (int32_t)nc;
    P.ncell = (int32_t)ncell;
    P.L = d->L;
    P.halfL = d->L / 2.0;
    P.h = d->L / (double)nc;
    P.inv_h = (double)nc / d->L;
    P.rl2 = P.h * P.h * (1.0 + 1e-5);
    P.seed = d->seed;
    set_physics(P, d, 0);  // the longest chain only matters to the rigid rod (no update here)
    P.nblk = (int32_t)((n + BS - 1) / BS);

    PoolArgs a{};
    const size_t nb = NA_MAX_BEADS;
    a.P = na_at<RepParams>(base, NA.P);
    a.S = na_at<RepState>(base, NA.S);
    a.blk_rep = na_at<int32_t>(base, NA.blk_rep);
    a.blk_slot0 = na_at<int32_t>(base, NA.blk_slot0);
    a.chk_rep = na_at<int32_t>(base, NA.chk_rep);
    a.chk_cell0 = na_at<int32_t>(base, NA.chk_cell0);
    a.rep_chk_off = na_at<int32_t>(base, NA.rep_chk_off);
    a.nblk = P.nblk;
    a.nchk = (int32_t)((ncell + SCAN_CHUNK - 1) / SCAN_CHUNK);
    a.nrep = 1;
    a.X = na_at<double>(base, NA.X);
    a.X1 = na_at<double>(base, NA.X1);
    a.T = na_at<double>(base, NA.T);
    a.Xn = na_at<double>(base, NA.Xn);
    a.Tn = na_at<double>(base, NA.Tn);
    a.X1n = na_at<double>(base, NA.X1n);
    a.FT = na_at<double>(base, NA.FT);
    a.XR = na_at<double>(base, NA.XR);
    a.XI = na_at<double>(base, NA.XI);
    a.XIR = a.XI + nb * dim;
    a.XIn = na_at<double>(base, NA.XIn);
    a.XIRn = a.XIn + nb * dim;
    a.META = na_at<int32_t>(base, NA.META);
    a.METAn = na_at<int32_t>(base, NA.METAn);
    a.KEYOFF = na_at<int32_t>(base, NA.KEYOFF);
    a.PAIRS = na_at<int32_t>(base, NA.PAIRS);
    a.LINK = na_at<int32_t>(base, NA.LINK);
    a.CELL = na_at<int32_t>(base, NA.CELL);
    a.SLOT = na_at<int32_t>(base, NA.SLOT);
    a.NLEN = na_at<int32_t>(base, NA.NLEN);
    a.nl_stride = (int64_t)nb;
    a.CHAIN_A = na_at<double>(base, NA.CHAIN_A);
    a.CHAIN_CS = na_at<double>(base, NA.CHAIN_CS);
    a.CHAIN_FIRST = na_at<int32_t>(base, NA.CHAIN_FIRST);
    a.CHAIN_LEN = na_at<int32_t>(base, NA.CHAIN_LEN);
    a.CCOUNT = na_at<int32_t>(base, NA.CCOUNT);
    a.CSTART = na_at<int32_t>(base, NA.CSTART);
    a.CHKSUM = na_at<int32_t>(base, NA.CHKSUM);
    NaBufs b{};
    b.blk_rep = na_at<int32_t>(base, NA.blk_rep);
    b.blk_slot0 = na_at<int32_t>(base, NA.blk_slot0);
    b.chk_rep = na_at<int32_t>(base, NA.chk_rep);
    b.chk_cell0 = na_at<int32_t>(base, NA.chk_cell0);
    b.rep_chk_off = na_at<int32_t>(base, NA.rep_chk_off);
    b.CHAIN_FIRST = na_at<int32_t>(base, NA.CHAIN_FIRST);
    b.CHAIN_LEN = na_at<int32_t>(base, NA.CHAIN_LEN);
    b.HEADS = na_at<int32_t>(base, NA.HEADS);
    b.ENDS = na_at<int32_t>(base, NA.ENDS);
    b.CHAIN_A = na_at<double>(base, NA.CHAIN_A);
    b.CHAIN_CS = na_at<double>(base, NA.CHAIN_CS);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dim == 2)
        return na_launch<2>(a, b, P, precision, pos, tangent, chain_id, rest_lengths, v_tan_angle, d->v0,
                            forces, tangent_out, st);
    return na_launch<3>(a, b, P, precision, pos, tangent, chain_id, rest_lengths, v_tan_angle, d->v0,
                        forces, tangent_out, st);
}
