// hsgen_create_hs_noalloc: hsgen_create_hs in FP64 with no device allocation,
// for callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it).
//
// Every workspace of the call (the G list when it is host memory, the grouped
// GEMM problem and tile lists, the operands Y and W, or Z, Zs and W for a
// system that does not fold ||u_dot||^2 into Y) lives in `scratch`,
// one file-scope device array of HSGEN_NOALLOC_SCRATCH_BYTES that each GPU the
// library runs on holds until the process ends.  In host code the array's name
// is only a host-side stand-in: the address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device when a context is created on it
// (noalloc_setup, called by hsgen_init with the device current), and only
// that address is handed to kernels and copies.
//
// The tile lists are written on the GPU by k_noalloc_plan, so the call copies
// nothing from host memory except a host G list.  Every byte of scratch that
// a kernel reads is written earlier in the same call (the generator writes
// every row and padding column of Y (or Z and Zs), the T application every column of
// W and k_clear_rows its rows past 2K), so nothing needs clearing between
// calls.  Calls on one GPU share its array and the last call's contents stay
// there: calls that overlap on one GPU, from any contexts, race.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "hsgen_host.h"

namespace hsgen {

constexpr size_t kScratchBytes = HSGEN_NOALLOC_SCRATCH_BYTES;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};   // per device: the address of that GPU's `scratch`

// What k_noalloc_plan writes: the grouped-GEMM problems (T application per
// atom, then S and H) and their tile lists, as hsgen_create_hs builds them on
// the host.
struct NoallocPlan {
    const AtomDev* atoms;
    const MLoc* mloc;
    const double2* M;
    const double2* Z;
    const double2* Zs;
    double2* W;
    double2* H;
    double2* S;
    int64_t ldz, ldH, ldS;
    int32_t natoms, ng, gp, kpad;
    int32_t tbm, tbn;        // T-application tile
    int32_t cbm, cbn;        // contraction tile
    int32_t do_h, do_s;
    GemmProblem* probs;      // [natoms + 2]: W_a = M_a Z_a per atom, then S, then H
    TileWork* work;          // T-application tiles, then the S and H tiles
    int32_t* pre;            // [natoms + 1]: first T-application tile of each atom
};

// Tiles of tile row bi of a lower-triangular ng x ng problem: the bj with
// bj * cbn < (bi + 1) * cbm (the host's add_hs_problems rule).
__host__ __device__ inline int lower_row_tiles(int bi, int ntc, int cbm, int cbn)
{
    return std::min(ntc, (int)(((int64_t)(bi + 1) * cbm + cbn - 1) / cbn));
}

// One workgroup of 256 threads.  Writes probs[0 .. natoms + 1],
// pre[0 .. natoms] and work[0 .. nwork - 1] with nwork as the host counts it.
__global__ __launch_bounds__(256) void k_noalloc_plan(NoallocPlan p)
{
    const int tid = threadIdx.x;
    const int ntn = (p.gp + p.tbn - 1) / p.tbn;
    if (p.do_h)
        for (int a = tid; a < p.natoms; a += blockDim.x) {
            const AtomDev at = p.atoms[a];
            const MLoc ml = p.mloc[a];
            GemmProblem q;
            q.L = p.M + ml.off;
            q.R = p.Z + 2 * (int64_t)at.off;
            q.C = p.W + 2 * (int64_t)at.off;
            q.ldl = ml.ld;
            q.ldr = p.ldz;
            q.ldc = p.ldz;
            q.m = 2 * at.nlm;
            q.n = p.gp;
            q.kdepth = ml.ld;
            q.lower = 0;
            p.probs[a] = q;
        }
    if (tid == 0) {
        int acc = 0;
        for (int a = 0; a < p.natoms; ++a) {
            p.pre[a] = acc;
            if (p.do_h) acc += (2 * p.atoms[a].nlm + p.tbm - 1) / p.tbm * ntn;
        }
        p.pre[p.natoms] = acc;
        for (int w = 0; w < 2; ++w) {
            GemmProblem q;
            q.L = p.Z;
            q.R = w == 0 ? p.Zs : p.W;
            q.C = w == 0 ? p.S : p.H;
            q.ldl = p.ldz;
            q.ldr = p.ldz;
            q.ldc = w == 0 ? p.ldS : p.ldH;
            q.m = p.ng;
            q.n = p.ng;
            q.kdepth = p.kpad;
            q.lower = 1;
            p.probs[p.natoms + w] = q;
        }
    }
    __syncthreads();
    if (p.do_h)
        for (int a = tid; a < p.natoms; a += blockDim.x) {
            int w = p.pre[a];
            const int ntm = (2 * p.atoms[a].nlm + p.tbm - 1) / p.tbm;
            for (int bi = 0; bi < ntm; ++bi)
                for (int bj = 0; bj < ntn; ++bj) p.work[w++] = TileWork{a, bi, bj, 0};
        }
    // S and H tiles row by row, the S and H tile of one (bi, bj) side by side
    const int base = p.pre[p.natoms];
    const int ntr = (p.ng + p.cbm - 1) / p.cbm;
    const int ntc = (p.ng + p.cbn - 1) / p.cbn;
    const int nprob = p.do_s + p.do_h;
    for (int bi = tid; bi < ntr; bi += blockDim.x) {
        int before = 0;
        for (int r = 0; r < bi; ++r) before += lower_row_tiles(r, ntc, p.cbm, p.cbn);
        int w = base + before * nprob;
        const int nrow = lower_row_tiles(bi, ntc, p.cbm, p.cbn);
        for (int bj = 0; bj < nrow; ++bj) {
            if (p.do_s) p.work[w++] = TileWork{p.natoms, bi, bj, 0};
            if (p.do_h) p.work[w++] = TileWork{p.natoms + 1, bi, bj, 0};
        }
    }
}

void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kScratchDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

// Once per device, with the device current: the address of its `scratch`.
hsgen_status noalloc_setup(hsgen_ctx* ctx)
{
    if (ctx->device < 0 || ctx->device >= kScratchDevices) return HSGEN_OK;   // entry point reports it
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return HSGEN_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(ctx, HSGEN_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                    hipGetErrorString(e));
    }
    g_scratch_ptr[ctx->device] = ptr;
    return HSGEN_OK;
}

}  // namespace hsgen

using namespace hsgen;

extern "C" hsgen_status hsgen_create_hs_noalloc(hsgen_ctx* ctx, const hsgen_sys* s, const hsgen_kpoint* kp,
                                                const hsgen_opts* opts, double* H, double* S, int64_t ld,
                                                hsgen_report* rep)
{
    if (!ctx || !s) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_status st = check_kpoint(ctx, kp);
    if (st) return st;
    hsgen_opts o{};
    if (opts) o = *opts;
    if (o.precision != HSGEN_PREC_FP64)
        return fail(ctx, HSGEN_ERR_UNSUPPORTED, "the allocation-free entry point computes in FP64 only");
    const int ng = kp->ng;
    if (ld < std::max(ng, 1)) return fail(ctx, HSGEN_ERR_ARG, "ld %lld < ng %d", (long long)ld, ng);
    const bool do_h = !(o.flags & HSGEN_SKIP_H), do_s = !(o.flags & HSGEN_SKIP_S);
    if ((do_h && !H) || (do_s && !S)) return fail(ctx, HSGEN_ERR_ARG, "null output matrix");
    if ((do_h && !is_device_ptr(H)) || (do_s && !is_device_ptr(S)))
        return fail(ctx, HSGEN_ERR_ARG, "the allocation-free entry point needs H and S in device memory");
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (ng == 0) return HSGEN_OK;
    HIPCHK(hipSetDevice(ctx->device));
    char* base = (char*)scratch_of(ctx->device);
    if (!base) return fail(ctx, HSGEN_ERR_UNSUPPORTED, "no scratch array on device %d", ctx->device);
    o.tile = pick_tile(ctx, o, ng, s->kdim);
    Geometry geo = geometry(o.tile, s->kdim, ng);
    geo.fold = s->fold;
    const int tc = tapp_cfg(geo);
    const ZgemmLaunchCfg tcc = zgemm_cfg(tc);
    if (!s->d_M || !s->d_mloc || s->m_bk != geo.t.bk || s->m_bm != geo.t.bm)
        return fail(ctx, HSGEN_ERR_UNSUPPORTED,
                    "the system's T blocks are padded for tiles with BK %d, BM %d, tile %d needs %d, %d (a "
                    "hsgen_create_hs call with this tile pads them once)",
                    s->m_bk, s->m_bm, geo.cfg, geo.t.bk, geo.t.bm);
    // tile counts (k_noalloc_plan writes exactly these)
    const int na = s->natoms;
    const int ntn = (int)((geo.gp + tcc.bn - 1) / tcc.bn);
    int64_t n_tmul = 0;
    if (do_h)
        for (int a = 0; a < na; ++a) n_tmul += (int64_t)((2 * s->atoms[a].nlm + tcc.bm - 1) / tcc.bm) * ntn;
    const int ntr = (ng + geo.t.bm - 1) / geo.t.bm, ntc = (ng + geo.t.bn - 1) / geo.t.bn;
    int64_t n_hs = 0;
    for (int bi = 0; bi < ntr; ++bi) n_hs += lower_row_tiles(bi, ntc, geo.t.bm, geo.t.bn);
    n_hs *= (int)do_s + (int)do_h;
    // scratch layout, 256-byte aligned pieces
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (size_t)roundup((int64_t)(off + bytes), 256);
        return at;
    };
    const bool gdev = is_device_ptr(kp->g);
    const size_t o_g = gdev ? 0 : take(sizeof(int32_t) * 3 * (size_t)ng);
    const size_t o_probs = take(sizeof(GemmProblem) * (size_t)(na + 2));
    const size_t o_work = take(sizeof(TileWork) * (size_t)(n_tmul + n_hs));
    const size_t o_pre = take(sizeof(int32_t) * (size_t)(na + 1));
    const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * (size_t)geo.gp;
    const size_t o_z = take(zbytes);
    const size_t o_zs = geo.fold ? o_z : take(zbytes);
    const size_t o_w = do_h ? take(zbytes) : 0;
    if (off > kScratchBytes)
        return fail(ctx, HSGEN_ERR_ARG,
                    "N_G %d with K %d needs %zu bytes of scratch; the allocation-free entry point has %zu "
                    "(use hsgen_create_hs)",
                    ng, s->kdim, off, kScratchBytes);
    if (n_tmul + n_hs > INT32_MAX) return fail(ctx, HSGEN_ERR_ARG, "too many tiles");
    const int32_t* dg = kp->g;
    if (!gdev) {
        HIPCHK(hipMemcpyAsync(base + o_g, kp->g, sizeof(int32_t) * 3 * (size_t)ng, hipMemcpyHostToDevice,
                              ctx->stream));
        dg = (const int32_t*)(base + o_g);
    }
    double2* Z = (double2*)(base + o_z);
    double2* Zs = (double2*)(base + o_zs);
    double2* W = do_h ? (double2*)(base + o_w) : nullptr;
    GemmProblem* probs = (GemmProblem*)(base + o_probs);
    TileWork* work = (TileWork*)(base + o_work);
    NoallocPlan pl;
    std::memset(&pl, 0, sizeof pl);
    pl.atoms = s->d_atoms;
    pl.mloc = s->d_mloc;
    pl.M = s->d_M;
    pl.Z = Z;
    pl.Zs = Zs;
    pl.W = W;
    pl.H = (double2*)H;
    pl.S = (double2*)S;
    pl.ldz = geo.ldz;
    pl.ldH = ld;
    pl.ldS = ld;
    pl.natoms = na;
    pl.ng = ng;
    pl.gp = (int32_t)geo.gp;
    pl.kpad = (int32_t)geo.kpad;
    pl.tbm = tcc.bm;
    pl.tbn = tcc.bn;
    pl.cbm = geo.t.bm;
    pl.cbn = geo.t.bn;
    pl.do_h = do_h;
    pl.do_s = do_s;
    pl.probs = probs;
    pl.work = work;
    pl.pre = (int32_t*)(base + o_pre);
    hipLaunchKernelGGL(k_noalloc_plan, dim3(1), dim3(256), 0, ctx->stream, pl);
    HIPCHK(hipGetLastError());
    AbgenParams p = abgen_params(s, kp, dg);
    p.ncols = (int32_t)geo.gp;
    p.Z = Z;
    p.Zs = geo.fold ? nullptr : Zs;
    p.ldz = geo.ldz;
    p.fold = geo.fold;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(launch_abgen(p, true, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    if (do_h) {
        HIPCHK(launch_zgemm(tc, probs, work, (int)n_tmul, ctx->stream));
        HIPCHK(launch_clear_rows(W, geo.ldz, 2 * (int64_t)s->kdim, geo.ldz, (int)geo.gp, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    HIPCHK(launch_zgemm(geo.cfg, probs, work + n_tmul, (int)n_hs, ctx->stream));
    if (o.flags & HSGEN_FILL_UPPER) {
        if (do_h) HIPCHK(launch_fill_upper((double2*)H, ld, ng, ctx->stream));
        if (do_s) HIPCHK(launch_fill_upper((double2*)S, ld, ng, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev[4]));
    if (rep) {
        float t0 = 0, t1 = 0, t2 = 0;
        HIPCHK(hipEventElapsedTime(&t0, ctx->ev[0], ctx->ev[1]));
        HIPCHK(hipEventElapsedTime(&t1, ctx->ev[1], ctx->ev[3]));
        HIPCHK(hipEventElapsedTime(&t2, ctx->ev[3], ctx->ev[4]));
        rep->ms_abgen = t0;
        rep->ms_tmul = t1;
        rep->ms_contract = t2;
        rep->ms_total = t0 + t1 + t2;
        rep->ng = ng;
        rep->kdim = s->kdim;
        rep->flops_alg = 16.0 * (double)ng * ng * s->kdim + 32.0 * (double)ng * (double)s->sum_nlm2;
        rep->bytes_abgen = (geo.fold ? 1.0 : 2.0) * 2.0 * 16.0 * (double)ng * s->kdim;
        rep->tile_used = geo.cfg;
    }
    return HSGEN_OK;
}
