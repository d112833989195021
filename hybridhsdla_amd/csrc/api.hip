// C-ABI implementation of hsgen (include/hsgen.h): contexts, device-resident
// systems, and the per-k-point pipeline
//   abgen (A/B generator) -> tmul (W_a = M_a Z_a) -> lower-triangle S and H.
//
// Operand layout in HBM, per k-point (complex128, column-major, ld = ldz):
//   Z  = [A_1; B_1; A_2; B_2; ...]   rows 2*off_a .. 2*off_a + 2*nlm_a - 1
//   Zs = [A_1; N_1 B_1; ...]         N_a = diag(||u_dot_l||^2)
//   W  = [M_1 Z_1; M_2 Z_2; ...]     M_a = [[T_aa, T_ab], [T_ab^H, T_bb]]
// so that   S = Z^H Zs   and   H = Z^H W   (only i >= j stored).
// ldz = roundup(2K, BK) + BK, columns padded to the tile size; every padding
// element is zero so the kernels need no bounds checks along k.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/hsgen.h"
#include "hsgen_internal.h"

using namespace hsgen;

static const int kDefaultTile = 11;
static const int kFallbackTile = 8;   // register-staged, 64-bit addressing
static const int kIgemmCfg = 51;     // default INT8 kernel configuration (256 x 192 tiles)
static const int kSampleTile = 12;   // FP64 tile of the split-precision sample: 64 x 64, so 64 columns

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct hsgen_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    // workspace
    DevBuf Z, Zs, W, Htmp, Stmp, A, B, g, probs, work, unorm2, atoms_tmp;
    // split precision
    DevBuf oz_e, oz_res, oz_out, oz_probs, oz_work, oz_smp, oz_fprob;
    hipEvent_t oz_ev[2 * kMaxModuli] = {};   // INT8 GEMM start/stop per residue group
    int n_cu = 0;                             // compute units (persistent grids)
    // batched k-points: one captured graph per batch, replayed while the
    // call signature is unchanged
    struct BatchGraph {
        std::vector<int64_t> sig;
        hipGraphExec_t exec = nullptr;
        DevBuf g, probs, work;
    };
    std::vector<BatchGraph> bgraphs;
    DevBuf btmp;                              // device H, S of host-output k-points
    // autotuned tile configurations: (precision, size buckets) -> config id
    std::map<std::vector<int>, int> tuned;
    // generalized eigensolver
    rocblas_handle blas = nullptr;
    DevBuf eig_a, eig_b, eig_d;
    // pinned staging for device -> host copies of H and S
    void* stage[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
};

struct hsgen_sys {
    uint64_t serial = 0;               // unique per created system
    int natoms = 0, max_lmax = 0, nl = 0, L = 0, kdim = 0;
    double omega = 0.0, vol_prefac = 0.0;
    double bmat[9];
    std::vector<AtomDev> atoms;        // host copy
    std::vector<int64_t> m_off;        // per atom: offset of M_a in d_M (complex units)
    std::vector<int32_t> m_ld;         // per atom: leading dim of M_a
    std::vector<int32_t> m_cols;       // per atom: padded column count
    int m_bk = 0, m_bm = 0;            // tile geometry M was padded for
    int64_t sum_nlm2 = 0;              // sum_a nlm_a^2
    AtomDev* d_atoms = nullptr;
    RadDev* d_rad = nullptr;
    double* d_ynorm = nullptr;
    double2* d_M = nullptr;
    // host copies of the T blocks, so that M can be rebuilt for another tile
    // geometry
    std::vector<std::vector<std::complex<double>>> h_taa, h_tab, h_tbb;
    std::vector<int> h_owner;          // per atom: the atom whose copies hold its T
    // first atom of each rotation class and of each rmt class (AtomDev::cls)
    std::vector<int32_t> rot_rep, rmt_rep;
};

namespace {

hsgen_status fail(hsgen_ctx* ctx, hsgen_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return s;
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ctx, HSGEN_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                        hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

hsgen_status ensure(hsgen_ctx* ctx, DevBuf& b, size_t bytes, bool zero_new)
{
    if (b.bytes >= bytes && b.p) return HSGEN_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return fail(ctx, HSGEN_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, HSGEN_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes,
                    hipGetErrorString(e));
    }
    b.bytes = bytes;
    if (zero_new) {
        e = hipMemsetAsync(b.p, 0, bytes, ctx->stream);
        if (e != hipSuccess) return fail(ctx, HSGEN_ERR_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    }
    return HSGEN_OK;
}

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

bool is_device_ptr(const void* p)
{
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

inline int64_t roundup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// ylm4 normalisation table (ylm4.f:236-262), index l(l+1)+m.
std::vector<double> ynorm_table(int lmax)
{
    const int L = (lmax + 1) * (lmax + 1);
    std::vector<double> yn(L);
    const double fpi = 4.0 * kPi;
    for (int l = 0; l <= lmax; ++l) {
        const int lm0 = l * (l + 1);
        double cd = 1.0;
        const double a = std::sqrt((2 * l + 1) / fpi);
        yn[lm0] = a;
        for (int m = 1; m <= l; ++m) {
            cd = cd / ((double)(l + 1 - m) * (double)(l + m));
            yn[lm0 + m] = a * std::sqrt(cd);
            yn[lm0 - m] = ((m & 1) ? -1.0 : 1.0) * yn[lm0 + m];
        }
    }
    return yn;
}

struct Geometry {
    int cfg;
    ZgemmLaunchCfg t;
    int64_t kpad;   // kdepth of the S/H contraction = roundup(2K, BK)
    int64_t ldz;    // kpad + BK
    int64_t gp;     // padded column count
};

// column padding for the split-precision path: a multiple of both sides of
// the INT8 tile (the same operand supplies rows and columns of S and H)
int64_t int8_col_align(int tile)
{
    const IgemmLaunchCfg c = igemm_cfg((tile > 0 && igemm_cfg(tile).bm > 0) ? tile : kIgemmCfg);
    return std::lcm<int64_t>(c.bm, c.bn);
}

Geometry geometry(int tile, int kdim, int ng, int64_t col_align = 1)
{
    Geometry g;
    g.cfg = tile > 0 ? tile : kDefaultTile;
    g.t = zgemm_cfg(g.cfg);
    if (g.t.bm == 0) {
        g.cfg = kDefaultTile;
        g.t = zgemm_cfg(g.cfg);
    }
    g.kpad = roundup(2 * (int64_t)kdim, g.t.bk);
    g.ldz = g.kpad + g.t.bk;
    if (zgemm_offs32(g.cfg) && (int64_t)std::max(g.t.bm, g.t.bn) * g.ldz * 16 >= (int64_t(1) << 32)) {
        // operand panel beyond 32-bit offsets: the register-staged kernel
        g.cfg = kFallbackTile;
        g.t = zgemm_cfg(g.cfg);
        g.kpad = roundup(2 * (int64_t)kdim, g.t.bk);
        g.ldz = g.kpad + g.t.bk;
    }
    const int64_t tm = std::max<int64_t>({(int64_t)g.t.bm, (int64_t)g.t.bn, col_align});
    g.gp = roundup(std::max(ng, 1), tm);
    return g;
}

// Build the per-atom M_a = [[T_aa, T_ab], [T_ab^H, T_bb]] (T_aa, T_bb taken
// from their lower triangles) padded for tile geometry (bk, bm), deduplicated
// by T pointer triple, and upload.
hsgen_status upload_M(hsgen_ctx* ctx, hsgen_sys* s, const int32_t* nlm, const double* const* taa,
                      const double* const* tab, const double* const* tbb, int bk, int bm)
{
    if (s->d_M) {
        (void)hipFree(s->d_M);
        s->d_M = nullptr;
    }
    const int na = s->natoms;
    s->m_off.assign(na, 0);
    s->m_ld.assign(na, 0);
    s->m_cols.assign(na, 0);
    std::map<std::vector<const void*>, int> seen;
    std::vector<int> owner(na, -1);
    int64_t total = 0;
    for (int a = 0; a < na; ++a) {
        std::vector<const void*> key{taa[a], tab[a], tbb[a], (const void*)(intptr_t)nlm[a]};
        auto it = seen.find(key);
        const int n2 = 2 * nlm[a];
        s->m_ld[a] = (int32_t)roundup(n2, bk);
        s->m_cols[a] = (int32_t)roundup(n2, bm);
        if (it != seen.end()) {
            owner[a] = it->second;
            s->m_off[a] = s->m_off[it->second];
            continue;
        }
        seen[key] = a;
        owner[a] = a;
        s->m_off[a] = total;
        total += (int64_t)s->m_ld[a] * s->m_cols[a];
    }
    std::vector<std::complex<double>> hM((size_t)std::max<int64_t>(total, 1), 0.0);
    for (int a = 0; a < na; ++a) {
        if (owner[a] != a) continue;
        const int n = nlm[a];
        const int64_t ld = s->m_ld[a];
        std::complex<double>* M = hM.data() + s->m_off[a];
        const std::complex<double>* Taa = (const std::complex<double>*)taa[a];
        const std::complex<double>* Tab = (const std::complex<double>*)tab[a];
        const std::complex<double>* Tbb = (const std::complex<double>*)tbb[a];
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                // Hermitian expansion from the lower triangle
                std::complex<double> aa = r >= c ? Taa[r + (int64_t)n * c] : std::conj(Taa[c + (int64_t)n * r]);
                std::complex<double> bb = r >= c ? Tbb[r + (int64_t)n * c] : std::conj(Tbb[c + (int64_t)n * r]);
                if (r == c) { aa = aa.real(); bb = bb.real(); }
                M[r + ld * c] = aa;
                M[(n + r) + ld * (n + c)] = bb;
                M[r + ld * (n + c)] = Tab[r + (int64_t)n * c];              // T_ab block
                M[(n + r) + ld * c] = std::conj(Tab[c + (int64_t)n * r]);   // T_ab^H block
            }
    }
    HIPCHK(hipMalloc(&s->d_M, sizeof(double2) * hM.size()));
    HIPCHK(hipMemcpy(s->d_M, hM.data(), sizeof(double2) * hM.size(), hipMemcpyHostToDevice));
    s->m_bk = bk;
    s->m_bm = bm;
    return HSGEN_OK;
}

hsgen_status ensure_M(hsgen_ctx* ctx, hsgen_sys* s, int bk, int bm)
{
    if (s->d_M && s->m_bk == bk && s->m_bm == bm) return HSGEN_OK;
    const int na = s->natoms;
    std::vector<const double*> pa(na), pb(na), pc(na);
    std::vector<int32_t> nlm(na);
    for (int a = 0; a < na; ++a) {
        const int o = s->h_owner[a];   // atoms sharing T share the pointers (upload_M deduplicates)
        pa[a] = (const double*)s->h_taa[o].data();
        pb[a] = (const double*)s->h_tab[o].data();
        pc[a] = (const double*)s->h_tbb[o].data();
        nlm[a] = s->atoms[a].nlm;
    }
    return upload_M(ctx, s, nlm.data(), pa.data(), pb.data(), pc.data(), bk, bm);
}

// Append the S and H lower-tile work of one k-point.
void add_hs_problems(std::vector<GemmProblem>& probs, std::vector<TileWork>& work,
                     const Geometry& geo, int ng, const double2* Z, const double2* Zs,
                     const double2* W, double2* H, double2* S, int64_t ld, int flags)
{
    const int ntr = (int)((ng + geo.t.bm - 1) / geo.t.bm);   // tile rows
    const int ntc = (int)((ng + geo.t.bn - 1) / geo.t.bn);   // tile columns
    int ids[2] = {-1, -1};
    for (int which = 0; which < 2; ++which) {
        const bool isS = which == 0;
        if (isS && (flags & HSGEN_SKIP_S)) continue;
        if (!isS && (flags & HSGEN_SKIP_H)) continue;
        GemmProblem p;
        p.L = Z;
        p.R = isS ? Zs : W;
        p.C = isS ? S : H;
        p.ldl = p.ldr = 0;
        p.ldc = ld;
        p.m = p.n = ng;
        p.kdepth = 0;
        p.lower = 1;
        ids[which] = (int)probs.size();
        probs.push_back(p);
    }
    // Logical order: patches of kPatchR x kPatchC tiles, S and H tiles of a
    // patch next to each other (they share the Z row panel).  Tiles exist
    // where the tile's first row index >= its first column index.
    const int kPatchR = 4, kPatchC = 4;
    for (int pi = 0; pi < ntr; pi += kPatchR)
        for (int pj = 0; pj < ntc; pj += kPatchC)
            for (int which = 0; which < 2; ++which) {
                if (ids[which] < 0) continue;
                for (int bi = pi; bi < std::min(ntr, pi + kPatchR); ++bi)
                    for (int bj = pj; bj < std::min(ntc, pj + kPatchC); ++bj)
                        if ((int64_t)bj * geo.t.bn < (int64_t)(bi + 1) * geo.t.bm)
                            work.push_back(TileWork{ids[which], bi, bj, 0});
            }
}

// Permute a logical tile list so that each XCD runs consecutive logical tiles.
// Workgroups are dispatched round-robin over the 8 XCDs in blockIdx order, so
// in a window of 8*32 blocks XCD x runs blocks x, x+8, ...; giving those the
// logical tiles [32*(8w+x), 32*(8w+x)+32) keeps one patch per XCD L2.
void xcd_permute(std::vector<TileWork>& work)
{
    const char* env = std::getenv("HSGEN_XCD_REMAP");
    if (env && env[0] == '0') return;
    const size_t kXcd = 8, kPer = 32, kWin = kXcd * kPer;
    const size_t n = work.size();
    const size_t full = n / kWin * kWin;
    std::vector<TileWork> out(work);
    for (size_t b = 0; b < full; ++b) {
        const size_t w = b / kWin, x = b % kXcd, slot = (b % kWin) / kXcd;
        out[b] = work[w * kWin + x * kPer + slot];
    }
    work.swap(out);
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* hsgen_status_string(hsgen_status s)
{
    switch (s) {
    case HSGEN_OK: return "ok";
    case HSGEN_ERR_ARG: return "invalid argument";
    case HSGEN_ERR_HIP: return "HIP runtime error";
    case HSGEN_ERR_NOMEM: return "out of memory";
    case HSGEN_ERR_NODEV: return "no usable gfx950 device";
    case HSGEN_ERR_UNSUPPORTED: return "unsupported request";
    case HSGEN_ERR_ACCURACY: return "accuracy target unreachable";
    }
    return "unknown status";
}

const char* hsgen_last_error(const hsgen_ctx* ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

hsgen_status hsgen_init(hsgen_ctx** out, int32_t device)
{
    hsgen_ctx* ctx = nullptr;
    if (!out) return HSGEN_ERR_ARG;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return HSGEN_ERR_NODEV;
    }
    if (device < 0 || device >= n) return HSGEN_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HSGEN_ERR_NODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HSGEN_ERR_NODEV;
    ctx = new hsgen_ctx();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto& ev : ctx->ev) HIPCHK(hipEventCreate(&ev));
    for (auto& ev : ctx->oz_ev) HIPCHK(hipEventCreate(&ev));
    *out = ctx;
    return HSGEN_OK;
}

hsgen_status hsgen_release_workspace(hsgen_ctx* ctx)
{
    if (!ctx) return HSGEN_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (DevBuf* b : {&ctx->Z, &ctx->Zs, &ctx->W, &ctx->Htmp, &ctx->Stmp, &ctx->A, &ctx->B, &ctx->g,
                      &ctx->probs, &ctx->work, &ctx->unorm2, &ctx->atoms_tmp, &ctx->oz_e, &ctx->oz_res,
                      &ctx->oz_out, &ctx->oz_fprob, &ctx->oz_probs, &ctx->oz_work, &ctx->oz_smp, &ctx->btmp, &ctx->eig_a,
                      &ctx->eig_b, &ctx->eig_d})
        release(*b);
    for (auto& bg : ctx->bgraphs) {
        if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
        release(bg.g);
        release(bg.probs);
        release(bg.work);
    }
    ctx->bgraphs.clear();
    return HSGEN_OK;
}

hsgen_status hsgen_finalize(hsgen_ctx* ctx)
{
    if (!ctx) return HSGEN_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    hsgen_release_workspace(ctx);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : ctx->oz_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    for (void*& p : ctx->stage)
        if (p) (void)hipHostFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HSGEN_OK;
}

hsgen_status hsgen_system_create(hsgen_ctx* ctx, const hsgen_system* d, hsgen_sys** out)
{
    if (!ctx || !d || !out) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    *out = nullptr;
    if (d->natoms <= 0) return fail(ctx, HSGEN_ERR_ARG, "natoms must be > 0");
    if (d->max_lmax < 1 || d->max_lmax > kMaxLmax)
        return fail(ctx, HSGEN_ERR_ARG, "max_lmax %d outside [1, %d] (sphbes needs lmax >= 1)",
                    d->max_lmax, kMaxLmax);
    if (!d->lmax || !d->rmt || !d->pos || !d->rotmat || !d->match || !d->bmat || !d->t_aa ||
        !d->t_ab || !d->t_bb)
        return fail(ctx, HSGEN_ERR_ARG, "null array in system description");
    if (!(d->omega > 0.0)) return fail(ctx, HSGEN_ERR_ARG, "omega must be > 0");
    HIPCHK(hipSetDevice(ctx->device));
    static std::atomic<uint64_t> next_serial{0};
    hsgen_sys* s = new hsgen_sys();
    s->serial = ++next_serial;
    s->natoms = d->natoms;
    s->max_lmax = d->max_lmax;
    s->nl = d->max_lmax + 1;
    s->L = s->nl * s->nl;
    s->omega = d->omega;
    s->vol_prefac = 4.0 * kPi / std::sqrt(d->omega);
    std::memcpy(s->bmat, d->bmat, sizeof s->bmat);
    int lmx = -1;
    int off = 0;
    s->atoms.resize(d->natoms);
    std::vector<RadDev> rad((size_t)d->natoms * s->nl);
    for (int a = 0; a < d->natoms; ++a) {
        const int lm = d->lmax[a];
        if (lm < 0 || lm > d->max_lmax) {
            delete s;
            return fail(ctx, HSGEN_ERR_ARG, "lmax[%d] = %d outside [0, max_lmax]", a, lm);
        }
        if (!(d->rmt[a] > 0.0)) {
            delete s;
            return fail(ctx, HSGEN_ERR_ARG, "rmt[%d] must be > 0", a);
        }
        if (!d->t_aa[a] || !d->t_ab[a] || !d->t_bb[a]) {
            delete s;
            return fail(ctx, HSGEN_ERR_ARG, "null T block for atom %d", a);
        }
        lmx = std::max(lmx, lm);
        AtomDev& at = s->atoms[a];
        at.lmax = lm;
        at.nlm = (lm + 1) * (lm + 1);
        at.off = off;
        at.rmt = d->rmt[a];
        for (int i = 0; i < 3; ++i) at.pos[i] = d->pos[3 * a + i];
        for (int i = 0; i < 9; ++i) at.rot[i] = d->rotmat[9 * a + i];
        // classes of bitwise-equal rotations and radii (they share Y_lm and
        // the Bessel values in the generator)
        int rc = 0, mc = 0;
        while (rc < (int)s->rot_rep.size() && std::memcmp(s->atoms[s->rot_rep[rc]].rot, at.rot, sizeof at.rot)) ++rc;
        if (rc == (int)s->rot_rep.size()) s->rot_rep.push_back(a);
        while (mc < (int)s->rmt_rep.size() && std::memcmp(&s->atoms[s->rmt_rep[mc]].rmt, &at.rmt, sizeof at.rmt)) ++mc;
        if (mc == (int)s->rmt_rep.size()) s->rmt_rep.push_back(a);
        at.cls = (rc & 0xffff) | ((mc & 0xffff) << 16);
        off += at.nlm;
        s->sum_nlm2 += (int64_t)at.nlm * at.nlm;
        for (int l = 0; l < s->nl; ++l) {
            RadDev& r = rad[(size_t)a * s->nl + l];
            if (l <= lm) {
                const double* m = d->match + ((size_t)a * s->nl + l) * 5;
                r.u = m[0];
                r.u_dr = m[1];
                r.dE_u = m[2];
                r.dE_u_dr = m[3];
                r.dE_u_norm = m[4];
                r.inv_wronk = 1.0 / (r.dE_u * r.u_dr - r.u * r.dE_u_dr);   // AtomData.cpp:63-66
            } else {
                r = RadDev{0, 0, 0, 0, 0, 0};
            }
        }
    }
    s->kdim = off;
    auto yn = ynorm_table(s->max_lmax);
    hsgen_status st = HSGEN_OK;
    hipError_t e = hipSuccess;
    e = hipMalloc(&s->d_atoms, sizeof(AtomDev) * s->atoms.size());
    if (e == hipSuccess) e = hipMalloc(&s->d_rad, sizeof(RadDev) * rad.size());
    if (e == hipSuccess) e = hipMalloc(&s->d_ynorm, sizeof(double) * yn.size());
    if (e == hipSuccess)
        e = hipMemcpy(s->d_atoms, s->atoms.data(), sizeof(AtomDev) * s->atoms.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_rad, rad.data(), sizeof(RadDev) * rad.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_ynorm, yn.data(), sizeof(double) * yn.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hsgen_system_destroy(ctx, s);
        return fail(ctx, HSGEN_ERR_HIP, "system upload: %s", hipGetErrorString(e));
    }
    // keep host copies of T (deduplicated later in upload_M)
    auto& xtaa = s->h_taa;
    auto& xtab = s->h_tab;
    auto& xtbb = s->h_tbb;
    xtaa.resize(d->natoms);
    xtab.resize(d->natoms);
    xtbb.resize(d->natoms);
    s->h_owner.assign(d->natoms, 0);
    std::map<const double*, int> seen;
    for (int a = 0; a < d->natoms; ++a) {
        const size_t n2 = (size_t)s->atoms[a].nlm * s->atoms[a].nlm;
        auto cp = [&](std::vector<std::complex<double>>& v, const double* p) {
            v.assign((const std::complex<double>*)p, (const std::complex<double>*)p + n2);
        };
        auto it = seen.find(d->t_aa[a]);
        if (it != seen.end() && d->t_ab[a] == d->t_ab[it->second] && d->t_bb[a] == d->t_bb[it->second] &&
            s->atoms[a].nlm == s->atoms[it->second].nlm) {
            // same blocks as an earlier atom (same type): keep one copy
            s->h_owner[a] = it->second;
            continue;
        }
        seen[d->t_aa[a]] = a;
        s->h_owner[a] = a;
        cp(xtaa[a], d->t_aa[a]);
        cp(xtab[a], d->t_ab[a]);
        cp(xtbb[a], d->t_bb[a]);
    }
    Geometry geo = geometry(0, s->kdim, 1);
    st = ensure_M(ctx, s, geo.t.bk, geo.t.bm);
    if (st != HSGEN_OK) {
        hsgen_system_destroy(ctx, s);
        return st;
    }
    *out = s;
    return HSGEN_OK;
}

hsgen_status hsgen_system_destroy(hsgen_ctx* ctx, hsgen_sys* s)
{
    if (!s) return HSGEN_ERR_ARG;
    if (ctx) (void)hipSetDevice(ctx->device);
    if (s->d_atoms) (void)hipFree(s->d_atoms);
    if (s->d_rad) (void)hipFree(s->d_rad);
    if (s->d_ynorm) (void)hipFree(s->d_ynorm);
    if (s->d_M) (void)hipFree(s->d_M);
    delete s;
    return HSGEN_OK;
}

}  // extern "C"

namespace {

hsgen_status check_kpoint(hsgen_ctx* ctx, const hsgen_kpoint* kp)
{
    if (!kp) return fail(ctx, HSGEN_ERR_ARG, "null k-point");
    if (kp->ng < 0) return fail(ctx, HSGEN_ERR_ARG, "ng < 0");
    if (kp->ng > 0 && !kp->g) return fail(ctx, HSGEN_ERR_ARG, "null g list");
    return HSGEN_OK;
}

hsgen_status upload_g(hsgen_ctx* ctx, const hsgen_kpoint* kp)
{
    const size_t bytes = sizeof(int32_t) * 3 * (size_t)std::max(kp->ng, 1);
    hsgen_status st = ensure(ctx, ctx->g, bytes, false);
    if (st) return st;
    if (kp->ng > 0)
        HIPCHK(hipMemcpyAsync(ctx->g.p, kp->g, sizeof(int32_t) * 3 * (size_t)kp->ng,
                              hipMemcpyHostToDevice, ctx->stream));
    return HSGEN_OK;
}

AbgenParams abgen_params(const hsgen_sys* s, const hsgen_kpoint* kp, const int32_t* dg)
{
    AbgenParams p;
    std::memset(&p, 0, sizeof p);
    p.atoms = s->d_atoms;
    p.rad = s->d_rad;
    p.ynorm = s->d_ynorm;
    p.g = dg;
    for (int i = 0; i < 3; ++i) p.kpt[i] = kp->k[i];
    for (int i = 0; i < 9; ++i) p.bmat[i] = s->bmat[i];
    p.vol_prefac = s->vol_prefac;
    p.natoms = s->natoms;
    p.max_lmax = s->max_lmax;
    p.nl = s->nl;
    p.L = s->L;
    p.ng = kp->ng;
    p.kdim = s->kdim;
    // atoms per LDS chunk: keep the block's LDS at <= 48 KiB.  With few
    // rotation and rmt classes the class-shared generator runs (it needs no
    // per-atom Y_lm in LDS, so its chunks are larger).
    const int nrot = (int)s->rot_rep.size(), nrmt = (int)s->rmt_rep.size();
    if (nrot <= kAbgenMaxCls && nrmt <= kAbgenMaxCls &&
        abgen_cls_lds_bytes(1, s->nl, s->L, nrot, nrmt) <= 48 * 1024) {
        int chunk = 32;
        while (chunk > 1 && abgen_cls_lds_bytes(chunk, s->nl, s->L, nrot, nrmt) > 48 * 1024) chunk /= 2;
        p.chunk = chunk;
        p.nrot = nrot;
        p.nrmt = nrmt;
        for (int c = 0; c < nrot; ++c) p.rot_rep[c] = s->rot_rep[c];
        for (int c = 0; c < nrmt; ++c) p.rmt_rep[c] = s->rmt_rep[c];
        return p;
    }
    int chunk = 16;
    while (chunk > 1 && abgen_lds_bytes(chunk, s->nl, s->L) > 48 * 1024) chunk /= 2;
    p.chunk = chunk;
    p.nrot = 0;   // general generator
    return p;
}

// Tile configuration of the T application: HSGEN_TAPP_TILE (tuning) if it
// names a configuration whose tiles divide the contraction's (M is padded
// for those), else the contraction's own.
int tapp_cfg(const Geometry& geo)
{
    if (const char* e = std::getenv("HSGEN_TAPP_TILE")) {
        const int id = std::atoi(e);
        const ZgemmLaunchCfg c = zgemm_cfg(id);
        if (c.bm > 0 && geo.t.bk % c.bk == 0 && geo.t.bm % c.bm == 0 &&
            !(zgemm_offs32(id) && (int64_t)std::max(c.bm, c.bn) * geo.ldz * 16 >= (int64_t(1) << 32)))
            return id;
    }
    return geo.cfg;
}

// W_a = M_a Z_a for all atoms over ncols columns (one grouped launch), in
// tiles of the T-application configuration tc.
void add_tmul_problems(std::vector<GemmProblem>& probs, std::vector<TileWork>& work, const hsgen_sys* s,
                       const Geometry& geo, const ZgemmLaunchCfg& tc, int64_t ncols, const double2* Z, double2* W)
{
    const int ntn = (int)((ncols + tc.bn - 1) / tc.bn);
    for (int a = 0; a < s->natoms; ++a) {
        const AtomDev& at = s->atoms[a];
        GemmProblem p;
        p.L = s->d_M + s->m_off[a];
        p.ldl = s->m_ld[a];
        p.R = Z + 2 * (int64_t)at.off;
        p.ldr = geo.ldz;
        p.C = W + 2 * (int64_t)at.off;
        p.ldc = geo.ldz;
        p.m = 2 * at.nlm;
        p.n = (int32_t)ncols;
        p.kdepth = s->m_ld[a];
        p.lower = 0;
        const int id = (int)probs.size();
        probs.push_back(p);
        const int ntm = (p.m + tc.bm - 1) / tc.bm;
        for (int bi = 0; bi < ntm; ++bi)
            for (int bj = 0; bj < ntn; ++bj) work.push_back(TileWork{id, bi, bj, 0});
    }
}

hsgen_status run_grouped(hsgen_ctx* ctx, int cfg, std::vector<GemmProblem>& probs,
                         std::vector<TileWork>& work)
{
    if (work.empty()) return HSGEN_OK;
    hsgen_status st = ensure(ctx, ctx->probs, sizeof(GemmProblem) * probs.size(), false);
    if (st) return st;
    st = ensure(ctx, ctx->work, sizeof(TileWork) * work.size(), false);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(ctx->probs.p, probs.data(), sizeof(GemmProblem) * probs.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->work.p, work.data(), sizeof(TileWork) * work.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_zgemm(cfg, (const GemmProblem*)ctx->probs.p, (const TileWork*)ctx->work.p,
                        (int)work.size(), ctx->stream));
    // the host vectors must outlive the async copies
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return HSGEN_OK;
}

// Copy the lower triangle (and optionally everything) of a device matrix with
// ld = n into a host matrix with leading dimension ld: blocks of columns go
// through two pinned staging buffers (the DMA of one block overlaps the
// host-side copy of the previous one, which several threads share), so that
// only the requested triangle of the caller's matrix is written.
hsgen_status copy_lower_to_host(hsgen_ctx* ctx, const double2* dsrc, int n, double* H, int64_t ld,
                                bool full)
{
    const size_t kStage = (size_t)64 << 20;
    if (!ctx->stage[0]) {
        for (int b = 0; b < 2; ++b) {
            hipError_t e = hipHostMalloc(&ctx->stage[b], kStage, hipHostMallocDefault);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                ctx->stage[b] = nullptr;
                return fail(ctx, HSGEN_ERR_NOMEM, "hipHostMalloc(%zu bytes) failed: %s", kStage, hipGetErrorString(e));
            }
        }
        ctx->stage_bytes = kStage;
    }
    const size_t colb = sizeof(double2) * (size_t)n;
    const int cpb = (int)std::max<size_t>(1, ctx->stage_bytes / colb);     // columns per block
    if ((size_t)cpb * colb > ctx->stage_bytes) {
        // one column exceeds a staging buffer: copy column pieces directly
        std::vector<double2> tmp((size_t)n);
        for (int j = 0; j < n; ++j) {
            const int i0 = full ? 0 : j;
            HIPCHK(hipMemcpy(tmp.data() + i0, dsrc + (int64_t)j * n + i0, sizeof(double2) * (n - i0),
                             hipMemcpyDeviceToHost));
            std::memcpy((double2*)H + (int64_t)j * ld + i0, tmp.data() + i0, sizeof(double2) * (n - i0));
        }
        return HSGEN_OK;
    }
    const int nblk = (n + cpb - 1) / cpb;
    auto issue = [&](int b) -> hipError_t {
        const int j0 = b * cpb, j1 = std::min(n, j0 + cpb);
        hipError_t e = hipMemcpyAsync(ctx->stage[b & 1], dsrc + (int64_t)j0 * n, colb * (size_t)(j1 - j0),
                                      hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev[5 + (b & 1)], ctx->stream);
        return e;
    };
    const int nthr = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    HIPCHK(issue(0));
    for (int b = 0; b < nblk; ++b) {
        HIPCHK(hipEventSynchronize(ctx->ev[5 + (b & 1)]));
        if (b + 1 < nblk) HIPCHK(issue(b + 1));
        const int j0 = b * cpb, j1 = std::min(n, j0 + cpb);
        const double2* src = (const double2*)ctx->stage[b & 1];
        auto work = [&](int t) {
            for (int j = j0 + t; j < j1; j += nthr) {
                const int i0 = full ? 0 : j;
                std::memcpy((double2*)H + (int64_t)j * ld + i0, src + (size_t)(j - j0) * n + i0,
                            sizeof(double2) * (size_t)(n - i0));
            }
        };
        if (nthr > 1 && (size_t)(j1 - j0) * colb >= ((size_t)4 << 20)) {
            std::vector<std::thread> th;
            for (int t2 = 1; t2 < nthr; ++t2) th.emplace_back(work, t2);
            work(0);
            for (auto& x : th) x.join();
        } else {
            for (int t2 = 0; t2 < nthr; ++t2) work(t2);
        }
    }
    return HSGEN_OK;
}


// ---------------------------------------------------------------------------
// Split precision (Ozaki scheme II on INT8 MFMA)
// ---------------------------------------------------------------------------
// Pairwise-coprime moduli <= 255 (symmetric residues fit INT8), largest first.
const int kModuli[kMaxModuli] = {255, 254, 253, 251, 247, 241, 239, 233, 229, 227,
                                 223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

struct OzPlan {
    int nmod = 0;
    int beta = 0;
    double log2P = 0.0;
    double Phi = 0.0, Plo = 0.0;
    double whi[kMaxModuli], wlo[kMaxModuli];
};

int inv_mod(int a, int m)
{
    a %= m;
    if (a < 0) a += m;
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}

// beta for n moduli: the largest beta with P >= 2^(2 beta + 1.01).  The
// integer products are bounded by Cauchy-Schwarz, |C'_ij| <= ||L'_i|| ||R'_j||
// < 2^(2 beta) (1 + 2^-19), whatever the depth; beta <= 52 keeps every scaled
// element an exact double.
int oz_beta(int n)
{
    double l2p = 0.0;
    for (int m = 0; m < n; ++m) l2p += std::log2((double)kModuli[m]);
    return std::min(52, (int)std::floor((l2p - 1.01) / 2.0));
}

// Smallest beta for which rounding the elements of a column of depth kd
// (complex) keeps its norm below 2^beta (1 + 2^-20): the rounding adds at most
// sqrt(2 kd) / 2 to a norm below 2^beta.
int oz_beta_min(int64_t kd)
{
    return (int)std::ceil(19.0 + 0.5 * std::log2(2.0 * (double)kd));
}

OzPlan oz_plan(int nmod)
{
    OzPlan pl;
    pl.nmod = nmod;
    pl.beta = oz_beta(nmod);
    // P as double-double (exact while the product fits 106 bits)
    double ph = 1.0, plo = 0.0;
    for (int m = 0; m < nmod; ++m) {
        const double p = kModuli[m];
        const double h = ph * p;
        const double e = std::fma(ph, p, -h);
        const double l = plo * p + e;
        ph = h + l;
        plo = l - (ph - h);
        pl.log2P += std::log2(p);
    }
    pl.Phi = ph;
    pl.Plo = plo;
    for (int m = 0; m < nmod; ++m) {
        int u = 1;
        for (int k = 0; k < nmod; ++k)
            if (k != m) u = (u * inv_mod(kModuli[k], kModuli[m])) % kModuli[m];
        const double p = kModuli[m];
        const double hi = (double)u / p;
        const double r = std::fma(-hi, p, (double)u);   // exact remainder
        pl.whi[m] = hi;
        pl.wlo[m] = r / p;
    }
    return pl;
}

// Modelled relative Frobenius error of tril(C), C = L^H R, at beta.  Rounding
// an element of column i to beta bits of the column's norm exponent e_i leaves
// an independent error uniform within +-2^(e_i - beta - 1) in each real part,
// so E|dC_ij|^2 = 2^-2beta / 6 (2^(2 eL_i) ||R_j||^2 + 2^(2 eR_j) ||L_i||^2);
// the lower triangle holds about half of all (i, j).
double oz_model(const int32_t* el, const double* sl, const int32_t* er, const double* sr, int ng, int beta,
                double tril_norm2)
{
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    for (int i = 0; i < ng; ++i) {
        a += std::ldexp(1.0, 2 * (el[i] - beta));
        b += sr[i];
        c += std::ldexp(1.0, 2 * (er[i] - beta));
        d += sl[i];
    }
    return tril_norm2 > 0.0 ? std::sqrt((a * b + c * d) / 12.0 / tril_norm2) : 0.0;
}

struct OzResult {
    int nmod = 0;
    double est_h = 0.0, est_s = 0.0;     // modelled error, exact norms
    double smp_h = 0.0, smp_s = 0.0;     // measured error on the FP64 sample block
    double ms_igemm = 0.0;
};

// One k-point of a split-precision run: columns [col0, col0 + ng) of the
// resident Z, Zs, W (each k-point's columns padded to a multiple of both INT8
// tile sides) and its device outputs.
struct OzK {
    int ng = 0;
    int64_t col0 = 0;
    int64_t ncols = 0;      // padded column count (ng rounded up to the INT8 tile multiple)
    double2* dH = nullptr;
    int64_t ldH = 0;
    double2* dS = nullptr;
    int64_t ldS = 0;
};

// One split-precision evaluation of every k-point with nmod moduli: residues,
// INT8 GEMMs that store C mod p (one launch per modulus over the tiles of all
// k-points), final recombination into the outputs.
hsgen_status oz_execute(hsgen_ctx* ctx, const Geometry& geo, const std::vector<OzK>& ks, int flags, int icfg,
                        int nmod, const int32_t* eZ, const int32_t* eZs, const int32_t* eW, double* ms_igemm)
{
    const IgemmLaunchCfg ic = igemm_cfg(icfg);
    const int64_t kd = roundup(geo.kpad, ic.bk);           // INT8 contraction depth (complex)
    const int64_t gp = geo.gp;                             // columns of all k-points
    const bool do_s = !(flags & HSGEN_SKIP_S), do_h = !(flags & HSGEN_SKIP_H);
    const OzPlan pl = oz_plan(nmod);
    const double2* Z = (const double2*)ctx->Z.p;
    const double2* Zs = (const double2*)ctx->Zs.p;
    const double2* W = (const double2*)ctx->W.p;
    // tile list (shared by every modulus; per k-point the FP64 path's order,
    // problem ids S (if any) then H of each k-point in turn); a tile's slot is
    // its place in the list
    std::vector<IgemmProblem> probs;
    std::vector<TileWork> work;
    {
        std::vector<GemmProblem> gp_dummy;
        Geometry g2 = geo;
        g2.t.bm = ic.bm;
        g2.t.bn = ic.bn;
        for (const OzK& k : ks)
            add_hs_problems(gp_dummy, work, g2, k.ng, nullptr, nullptr, nullptr, nullptr, nullptr, 0, flags);
        xcd_permute(work);
        for (size_t w = 0; w < work.size(); ++w) work[w].slot = (int32_t)w;
    }
    // residue outputs of every modulus: C mod p_m as 2 bytes per element of
    // every tile (the recombination reads them all at the end)
    const size_t mod_stride = 2 * (size_t)ic.bm * ic.bn * work.size();
    hsgen_status st = ensure(ctx, ctx->oz_out, std::max<size_t>(mod_stride * nmod, 1), false);
    if (st) return st;
    uint8_t* rout = (uint8_t*)ctx->oz_out.p;
    // Residue arrays of one modulus ("slot"): Zr, Zs_r, W_r (+ rotated copies
    // the kernel reads), 2 bytes per complex.  As many moduli as fit in 60% of
    // the free memory share one residue pass, which reads the FP64 operands
    // once for the whole group.
    const int64_t ldo = 2 * kd;
    if ((uint64_t)ldo * (uint64_t)(ic.bm + ic.bn) >= (1ull << 32))
        return fail(ctx, HSGEN_ERR_UNSUPPORTED, "INT8 contraction depth %lld too large for 32-bit tile offsets",
                    (long long)kd);
    const size_t arr = (size_t)ldo * gp;
    const int narr = ic.needs_rot == 1 ? 4 : ic.needs_rot == 2 ? 5 : 3;
    const size_t slot_bytes = narr * arr;
    int group = 1;
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            fr += ctx->oz_res.bytes;   // the current residue buffer may be regrown
            group = (int)std::max<size_t>(1, std::min<size_t>((size_t)nmod, (size_t)(0.6 * fr) / slot_bytes));
        }
        group = std::min(group, 16);
        // the same number of passes with equal groups (less memory)
        const int npass = (nmod + group - 1) / group;
        group = (nmod + npass - 1) / npass;
    }
    st = ensure(ctx, ctx->oz_res, group * slot_bytes, false);
    if (st) return st;
    auto slot = [&](int i, int a) -> int8_t* { return (int8_t*)ctx->oz_res.p + i * slot_bytes + a * arr; };
    // Problem sets, one per (slot, depth chunk).  Kernels without an in-kernel
    // mod reduction get the depth in INT32-safe chunks, one launch each; the
    // chunks after the first add their residues mod p to the stored ones.
    const int64_t kc = ic.max_kchunk > 0 ? ic.max_kchunk : kd;
    const int nchunk = (int)((kd + kc - 1) / kc);
    size_t nprob = 0;
    for (int i = 0; i < group; ++i) {
        int8_t* Zr = slot(i, 0);
        int8_t* Zsr = slot(i, 1);
        int8_t* Wr = slot(i, 2);
        int8_t* Zrot = ic.needs_rot == 1 ? slot(i, 3) : nullptr;     // (-Li, Lr) of Z
        int8_t* Zsrot = ic.needs_rot == 2 ? slot(i, 3) : nullptr;    // (Ri, -Rr) of Zs
        int8_t* Wrot = ic.needs_rot == 2 ? slot(i, 4) : nullptr;     // (Ri, -Rr) of W
        std::vector<IgemmProblem> base;
        for (const OzK& k : ks) {
            const int64_t o = k.col0 * ldo;                           // bytes
            auto at = [&](int8_t* a) -> const int8_t* { return a ? a + o : nullptr; };
            if (do_s) base.push_back(IgemmProblem{at(Zr), at(Zrot), at(Zsr), at(Zsrot), ldo, ldo, k.ng, k.ng,
                                                  (int32_t)kd, 1});
            if (do_h) base.push_back(IgemmProblem{at(Zr), at(Zrot), at(Wr), at(Wrot), ldo, ldo, k.ng, k.ng,
                                                  (int32_t)kd, 1});
        }
        nprob = base.size();
        for (int c = 0; c < nchunk; ++c)
            for (IgemmProblem p : base) {
                const int64_t off = 2 * (int64_t)c * kc;   // bytes
                p.Lre += off;
                if (p.Lrot) p.Lrot += off;
                p.R += off;
                if (p.Rrot) p.Rrot += off;
                p.kdepth = (int32_t)std::min<int64_t>(kc, kd - (int64_t)c * kc);
                probs.push_back(p);
            }
    }
    // recombination targets, one per problem id
    std::vector<OzFinalProblem> fprobs;
    for (const OzK& k : ks) {
        if (do_s) fprobs.push_back(OzFinalProblem{k.dS, k.ldS, eZ + k.col0, eZs + k.col0, k.ng, k.ng});
        if (do_h) fprobs.push_back(OzFinalProblem{k.dH, k.ldH, eZ + k.col0, eW + k.col0, k.ng, k.ng});
    }
    st = ensure(ctx, ctx->oz_probs, sizeof(IgemmProblem) * std::max<size_t>(probs.size(), 1), false);
    if (st) return st;
    st = ensure(ctx, ctx->oz_work, sizeof(TileWork) * std::max<size_t>(work.size(), 1), false);
    if (st) return st;
    st = ensure(ctx, ctx->oz_fprob, sizeof(OzFinalProblem) * std::max<size_t>(fprobs.size(), 1), false);
    if (st) return st;
    if (!probs.empty())
        HIPCHK(hipMemcpyAsync(ctx->oz_probs.p, probs.data(), sizeof(IgemmProblem) * probs.size(),
                              hipMemcpyHostToDevice, ctx->stream));
    if (!work.empty())
        HIPCHK(hipMemcpyAsync(ctx->oz_work.p, work.data(), sizeof(TileWork) * work.size(),
                              hipMemcpyHostToDevice, ctx->stream));
    if (!fprobs.empty())
        HIPCHK(hipMemcpyAsync(ctx->oz_fprob.p, fprobs.data(), sizeof(OzFinalProblem) * fprobs.size(),
                              hipMemcpyHostToDevice, ctx->stream));
    const int rows = (int)std::min<int64_t>(geo.ldz, kd);
    // column ranges of the k-points (adjacent ones merged): the residue
    // passes convert only these columns
    std::vector<std::pair<int64_t, int64_t>> ranges;
    {
        std::vector<std::pair<int64_t, int64_t>> r;
        for (const OzK& k : ks) r.push_back({k.col0, k.col0 + k.ncols});
        std::sort(r.begin(), r.end());
        for (const auto& x : r)
            if (!ranges.empty() && ranges.back().second == x.first) ranges.back().second = x.second;
            else ranges.push_back(x);
    }
    int ngroups = 0;
    for (int m0 = 0; m0 < nmod; m0 += group, ++ngroups) {
        const int gn = std::min(group, nmod - m0);
        OzResParams rp;
        std::memset(&rp, 0, sizeof rp);
        rp.ldx = geo.ldz;
        rp.rows = (int32_t)kd;
        rp.rows_src = rows;
        rp.beta = pl.beta;
        rp.nmod = gn;
        rp.ldo = ldo;
        for (int i = 0; i < gn; ++i) {
            rp.p[i] = kModuli[m0 + i];
            rp.invp[i] = 1.0 / kModuli[m0 + i];
        }
        // residues of the k-points' columns of X into arrays a (and rotated arot)
        auto residues = [&](const double2* X, const int32_t* e, int a, int arot) -> hipError_t {
            for (const auto& rg : ranges) {
                const int64_t c0 = rg.first;
                rp.X = X + c0 * geo.ldz;
                rp.e = e + c0;
                for (int i = 0; i < gn; ++i) {
                    rp.out[i] = slot(i, a) + c0 * ldo;
                    rp.out_rot[i] = arot >= 0 ? slot(i, arot) + c0 * ldo : nullptr;
                }
                const hipError_t err = launch_residues(rp, (int)(rg.second - c0), ctx->stream);
                if (err != hipSuccess) return err;
            }
            return hipSuccess;
        };
        rp.rot_kind = 1;
        HIPCHK(residues(Z, eZ, 0, ic.needs_rot == 1 ? 3 : -1));
        rp.rot_kind = 2;
        if (do_s) HIPCHK(residues(Zs, eZs, 1, ic.needs_rot == 2 ? 3 : -1));
        if (do_h) HIPCHK(residues(W, eW, 2, ic.needs_rot == 2 ? 4 : -1));
        HIPCHK(hipEventRecord(ctx->oz_ev[2 * ngroups], ctx->stream));
        for (int i = 0; i < gn; ++i) {
            const int m = m0 + i;
            for (int c = 0; c < nchunk; ++c) {
                IgemmModulus md;
                md.p = kModuli[m];
                md.first_chunk = c == 0;
                md.chunk_stages = 2048 / (ic.bk / 32);    // 2048 MFMA k-steps <= 2080 (|acc| < 2^31)
                md.pad = 0;
                md.invp = 1.0 / kModuli[m];
                md.out = rout + (size_t)m * mod_stride;
                HIPCHK(launch_igemm(icfg, (const IgemmProblem*)ctx->oz_probs.p + (i * nchunk + c) * nprob,
                                    (const TileWork*)ctx->oz_work.p, (int)work.size(), md, ctx->stream));
            }
        }
        HIPCHK(hipEventRecord(ctx->oz_ev[2 * ngroups + 1], ctx->stream));
    }
    OzFinalParams fp;
    std::memset(&fp, 0, sizeof fp);
    fp.res = rout;
    fp.mod_stride = (int64_t)mod_stride;
    fp.prob = (const OzFinalProblem*)ctx->oz_fprob.p;
    fp.nmod = nmod;
    fp.beta = pl.beta;
    fp.bm = ic.bm;
    fp.bn = ic.bn;
    fp.wgm = ic.wgm;
    fp.wgn = ic.wgn;
    for (int m = 0; m < nmod; ++m) {
        fp.whi[m] = pl.whi[m];
        fp.wlo[m] = pl.wlo[m];
    }
    fp.Phi = pl.Phi;
    fp.Plo = pl.Plo;
    HIPCHK(launch_crt_final(fp, (const TileWork*)ctx->oz_work.p, (int)work.size(), ctx->stream));
    // the host vectors above must outlive the async copies
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *ms_igemm = 0.0;
    for (int g = 0; g < ngroups; ++g) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, ctx->oz_ev[2 * g], ctx->oz_ev[2 * g + 1]));
        *ms_igemm += t;
    }
    return HSGEN_OK;
}

// S = Z^H Zs and H = Z^H W of every k-point from the resident FP64 operands,
// on INT8 MFMA.  With nmod_req == 0 the modulus count is chosen for `target`:
// per k-point, 64 evenly spaced columns of S and H computed in FP64 (all rows)
// give |tril C|_F for the error model before any INT8 work, and the largest
// count any k-point needs is used for all; after the run the model is
// evaluated with the exact norms and the run repeated with more moduli if a
// k-point misses the target.  The same samples measure the actual errors.
hsgen_status run_ozaki(hsgen_ctx* ctx, const Geometry& geo, const std::vector<OzK>& ks, int flags, int icfg,
                       int nmod_req, double target, std::vector<OzResult>& res)
{
    const IgemmLaunchCfg ic = igemm_cfg(icfg);
    const int64_t kd = roundup(geo.kpad, ic.bk);           // INT8 contraction depth (complex)
    const int64_t gp = geo.gp;
    const bool do_s = !(flags & HSGEN_SKIP_S), do_h = !(flags & HSGEN_SKIP_H);
    const int nk = (int)ks.size();
    res.assign(nk, OzResult{});
    if (gp % ic.bm != 0 || gp % ic.bn != 0)
        return fail(ctx, HSGEN_ERR_ARG, "column padding %lld is not a multiple of the INT8 tile", (long long)gp);
    for (const OzK& k : ks)
        if (k.ng <= 0 || k.col0 % ic.bm != 0 || k.col0 % ic.bn != 0)
            return fail(ctx, HSGEN_ERR_ARG, "k-point with %d columns at offset %lld (INT8 tile multiple needed)",
                        k.ng, (long long)k.col0);
    if (!(target > 0.0)) target = 1e-12;
    const int bmin = oz_beta_min(kd);
    if (nmod_req > 0 && (nmod_req > kMaxModuli || oz_beta(nmod_req) < bmin))
        return fail(ctx, HSGEN_ERR_ACCURACY, "%d moduli give %d bits per element, depth %lld needs %d..52",
                    nmod_req, nmod_req > kMaxModuli ? 0 : oz_beta(nmod_req), (long long)kd, bmin);
    // per-column norm exponents and squared norms of Z, Zs, W; per k-point
    // sample statistics over as many columns as the sample tile is wide
    const ZgemmLaunchCfg sc = zgemm_cfg(kSampleTile);
    const int kNblk = 1024;
    std::vector<int> nb(nk);
    std::vector<size_t> soff(nk), smp_off(nk), gath_off(nk);
    size_t nbt = 0, smpt = 0, gatht = 0;
    for (int k = 0; k < nk; ++k) {
        nb[k] = std::min(ks[k].ng, sc.bn);
        soff[k] = nbt;
        nbt += nb[k];
        smp_off[k] = smpt;
        smpt += (size_t)ks[k].ng * nb[k];
        gath_off[k] = gatht;
        gatht += (size_t)geo.ldz * roundup(nb[k], sc.bn);
    }
    hsgen_status st = ensure(ctx, ctx->oz_e, (sizeof(int32_t) + sizeof(double)) * 3 * (size_t)gp +
                                                 sizeof(double) * (2 * (size_t)kNblk * nk + 2 * 3 * nbt) +
                                                 sizeof(int32_t) * std::max<size_t>(nbt, 1), false);
    if (st) return st;
    double* ssZ = (double*)ctx->oz_e.p;
    double* ssZs = ssZ + gp;
    double* ssW = ssZs + gp;
    double* part = ssW + gp;                       // per k-point 2 x kNblk partial sums of |tril C|^2
    double* spart = part + 2 * (size_t)kNblk * nk; // per k-point 2 x 3 nb sample statistics (S, H)
    int32_t* eZ = (int32_t*)(spart + 2 * 3 * nbt);
    int32_t* eZs = eZ + gp;
    int32_t* eW = eZs + gp;
    int32_t* dcols = eW + gp;                      // the sampled columns, relative to their k-point
    std::vector<int32_t> cols(std::max<size_t>(nbt, 1));
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nb[k]; ++j)
            cols[soff[k] + j] = (int32_t)(((2 * (int64_t)j + 1) * ks[k].ng) / (2 * (int64_t)nb[k]));
    if (nbt > 0)
        HIPCHK(hipMemcpyAsync(dcols, cols.data(), sizeof(int32_t) * nbt, hipMemcpyHostToDevice, ctx->stream));
    const double2* Z = (const double2*)ctx->Z.p;
    const double2* Zs = (const double2*)ctx->Zs.p;
    const double2* W = (const double2*)ctx->W.p;
    HIPCHK(launch_colexp(Z, geo.ldz, (int)geo.ldz, (int)gp, eZ, ssZ, ctx->stream));
    if (do_s) HIPCHK(launch_colexp(Zs, geo.ldz, (int)geo.ldz, (int)gp, eZs, ssZs, ctx->stream));
    if (do_h) HIPCHK(launch_colexp(W, geo.ldz, (int)geo.ldz, (int)gp, eW, ssW, ctx->stream));
    // FP64 samples: columns cols[] of S and H of each k-point, all rows, ld = ng
    // (the right operand's sampled columns are gathered first: the GEMM reads
    // contiguous columns, whole tiles of bn of them)
    st = ensure(ctx, ctx->oz_smp, sizeof(double2) * std::max<size_t>(2 * (smpt + gatht), 1), false);
    if (st) return st;
    double2* smpS = (double2*)ctx->oz_smp.p;
    double2* smpH = smpS + smpt;
    double2* gathS = smpH + smpt;
    double2* gathH = gathS + gatht;
    {
        std::vector<GemmProblem> probs;
        std::vector<TileWork> work;
        auto add = [&](const OzK& k, const double2* R, double2* C, int n) {
            GemmProblem p;
            p.L = Z + k.col0 * geo.ldz;
            p.R = R;
            p.C = C;
            p.ldl = p.ldr = geo.ldz;
            p.ldc = k.ng;
            p.m = k.ng;
            p.n = n;
            p.kdepth = (int32_t)geo.kpad;
            p.lower = 0;
            const int ntr = (k.ng + sc.bm - 1) / sc.bm;
            for (int bi = 0; bi < ntr; ++bi) work.push_back(TileWork{(int)probs.size(), bi, 0, 0});
            probs.push_back(p);
        };
        for (int k = 0; k < nk; ++k) {
            if (nb[k] == 0) continue;
            const OzK& q = ks[k];
            if (do_s) {
                HIPCHK(launch_gather_cols(Zs + q.col0 * geo.ldz, geo.ldz, dcols + soff[k], (int)geo.ldz, nb[k],
                                          gathS + gath_off[k], ctx->stream));
                add(q, gathS + gath_off[k], smpS + smp_off[k], nb[k]);
            }
            if (do_h) {
                HIPCHK(launch_gather_cols(W + q.col0 * geo.ldz, geo.ldz, dcols + soff[k], (int)geo.ldz, nb[k],
                                          gathH + gath_off[k], ctx->stream));
                add(q, gathH + gath_off[k], smpH + smp_off[k], nb[k]);
            }
        }
        st = run_grouped(ctx, kSampleTile, probs, work);
        if (st) return st;
    }
    for (int k = 0; k < nk; ++k) {
        if (nb[k] == 0) continue;
        const int ng = ks[k].ng;
        double* sp = spart + 6 * soff[k];
        if (do_s) HIPCHK(launch_block_cmp(nullptr, 0, smpS + smp_off[k], ng, ng, nb[k], dcols + soff[k], 0, sp,
                                          ctx->stream));
        if (do_h) HIPCHK(launch_block_cmp(nullptr, 0, smpH + smp_off[k], ng, ng, nb[k], dcols + soff[k], 0,
                                          sp + 3 * nb[k], ctx->stream));
    }
    std::vector<double> hss(3 * (size_t)gp + 2 * (size_t)kNblk * nk + 6 * nbt);
    std::vector<int32_t> he(3 * (size_t)gp);
    HIPCHK(hipMemcpyAsync(hss.data(), ssZ, sizeof(double) * hss.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(he.data(), eZ, sizeof(int32_t) * he.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double* hs_spart = hss.data() + 3 * (size_t)gp + 2 * (size_t)kNblk * nk;
    // |tril C|^2 from the sample: |C|^2 ~ (ng / nb) sum |C_:,cols|^2, the
    // diagonal likewise, and |tril C|^2 = (|C|^2 + |diag C|^2) / 2 for
    // Hermitian C
    std::vector<std::array<double, 2>> tril2_smp(nk, {0.0, 0.0});
    for (int k = 0; k < nk; ++k) {
        if (nb[k] == 0) continue;
        const double* sp = hs_spart + 6 * soff[k];
        for (int w = 0; w < 2; ++w) {
            double y = 0.0, dg = 0.0;
            for (int j = 0; j < nb[k]; ++j) {
                y += sp[3 * (w * nb[k] + j)];
                dg += sp[3 * (w * nb[k] + j) + 2];
            }
            tril2_smp[k][w] = 0.5 * (y + dg) * (double)ks[k].ng / (double)nb[k];
        }
    }
    auto model = [&](int k, int which, int beta, double tril2) {   // which 0: S (R = Zs), 1: H (R = W)
        const size_t o = (which == 0 ? 1 : 2) * (size_t)gp;
        const size_t c = (size_t)ks[k].col0;
        return oz_model(he.data() + c, hss.data() + c, he.data() + o + c, hss.data() + o + c, ks[k].ng, beta,
                        tril2);
    };
    auto worst_model = [&](int k, int beta, const std::array<double, 2>& tril2) {
        return std::max(do_s ? model(k, 0, beta, tril2[0]) : 0.0, do_h ? model(k, 1, beta, tril2[1]) : 0.0);
    };
    // modulus count per k-point: the smallest that meets the target on the
    // sample (k-points with equal counts share their INT8 launches)
    std::vector<int> nm(nk, nmod_req);
    if (nmod_req <= 0)
        for (int k = 0; k < nk; ++k) {
            int n = 2;
            while (n <= kMaxModuli && !(oz_beta(n) >= bmin && worst_model(k, oz_beta(n), tril2_smp[k]) <= 0.75 * target))
                ++n;
            if (n > kMaxModuli)
                return fail(ctx, HSGEN_ERR_ACCURACY, "no modulus count in [2, %d] reaches rel err %.3g", kMaxModuli,
                            target);
            nm[k] = n;
        }
    std::vector<char> todo(nk, 1);
    for (bool again = true; again;) {
        std::vector<int> counts;
        for (int k = 0; k < nk; ++k)
            if (todo[k]) counts.push_back(nm[k]);
        std::sort(counts.begin(), counts.end());
        counts.erase(std::unique(counts.begin(), counts.end()), counts.end());
        for (int n : counts) {
            std::vector<OzK> sub;
            for (int k = 0; k < nk; ++k)
                if (todo[k] && nm[k] == n) sub.push_back(ks[k]);
            double ms_ig = 0.0;
            st = oz_execute(ctx, geo, sub, flags, icfg, n, eZ, eZs, eW, &ms_ig);
            if (st) return st;
            // each k-point's share of the shared launches, by its work (ng^2)
            double w = 0.0;
            for (const OzK& q : sub) w += (double)q.ng * q.ng;
            for (int k = 0; k < nk; ++k)
                if (todo[k] && nm[k] == n) {
                    res[k].nmod = n;
                    res[k].ms_igemm = ms_ig * ((double)ks[k].ng * ks[k].ng / w);
                }
        }
        // modelled errors with the exact |tril C|
        for (int k = 0; k < nk; ++k) {
            if (!todo[k]) continue;
            double* pk = part + 2 * (size_t)kNblk * k;
            if (do_s) HIPCHK(launch_tril_sumsq(ks[k].dS, ks[k].ldS, ks[k].ng, pk, kNblk, ctx->stream));
            if (do_h) HIPCHK(launch_tril_sumsq(ks[k].dH, ks[k].ldH, ks[k].ng, pk + kNblk, kNblk, ctx->stream));
        }
        std::vector<double> hp(2 * (size_t)kNblk * nk);
        HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        again = false;
        for (int k = 0; k < nk; ++k) {
            if (!todo[k]) continue;
            todo[k] = 0;
            const int beta = oz_beta(nm[k]);
            double tril2[2] = {0.0, 0.0};
            for (int b = 0; b < kNblk; ++b) {
                tril2[0] += hp[2 * (size_t)kNblk * k + b];
                tril2[1] += hp[2 * (size_t)kNblk * k + kNblk + b];
            }
            OzResult& r = res[k];
            r.est_s = do_s ? model(k, 0, beta, tril2[0]) : 0.0;
            r.est_h = do_h ? model(k, 1, beta, tril2[1]) : 0.0;
            const double worst = std::max(r.est_s, r.est_h);
            if (nmod_req > 0 || worst <= target || nm[k] >= kMaxModuli) continue;
            // each modulus adds log2(p) / 2 ~ 3.9 bits to beta
            const int extra = (int)std::ceil((std::log2(worst / target) + 1.0) / 3.9);
            nm[k] = std::min(kMaxModuli, nm[k] + std::max(extra, 1));
            todo[k] = 1;
            again = true;
        }
    }
    // measured errors on the sample blocks (lower part, rows i >= column)
    for (int k = 0; k < nk; ++k) {
        if (nb[k] == 0) continue;
        const int ng = ks[k].ng;
        double* sp = spart + 6 * soff[k];
        if (do_s) HIPCHK(launch_block_cmp(ks[k].dS, ks[k].ldS, smpS + smp_off[k], ng, ng, nb[k], dcols + soff[k], 1,
                                          sp, ctx->stream));
        if (do_h) HIPCHK(launch_block_cmp(ks[k].dH, ks[k].ldH, smpH + smp_off[k], ng, ng, nb[k], dcols + soff[k], 1,
                                          sp + 3 * nb[k], ctx->stream));
    }
    std::vector<double> hsp(std::max<size_t>(6 * nbt, 1));
    if (nbt > 0)
        HIPCHK(hipMemcpyAsync(hsp.data(), spart, sizeof(double) * 6 * nbt, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nk; ++k) {
        if (nb[k] == 0) continue;
        const double* sp = hsp.data() + 6 * soff[k];
        for (int w = 0; w < 2; ++w) {
            double y = 0.0, d = 0.0;
            for (int j = 0; j < nb[k]; ++j) {
                y += sp[3 * (w * nb[k] + j)];
                d += sp[3 * (w * nb[k] + j) + 1];
            }
            const double e = y > 0.0 ? std::sqrt(d / y) : 0.0;
            if (w == 0) res[k].smp_s = do_s ? e : 0.0;
            else res[k].smp_h = do_h ? e : 0.0;
        }
    }
    return HSGEN_OK;
}

// Size buckets of the autotuning table: half-octaves of G and K.
std::vector<int> tune_key(int precision, int ng, int kdim)
{
    auto b = [](int x) { return (int)std::floor(2.0 * std::log2((double)std::max(x, 1))); };
    return {precision, b(ng), b(kdim)};
}

// Tile configuration for a call: the caller's, else the autotuned one for
// this size, else 0 (the built-in default).
int pick_tile(const hsgen_ctx* ctx, const hsgen_opts& o, int ng, int kdim)
{
    if (o.tile > 0) return o.tile;
    auto it = ctx->tuned.find(tune_key(o.precision, ng, kdim));
    return it == ctx->tuned.end() ? 0 : it->second;
}

// Common tail of create_hs / contract_hs: Z, Zs resident -> tmul -> S, H.
hsgen_status finish_hs(hsgen_ctx* ctx, hsgen_sys* s, const Geometry& geo, int ng, const hsgen_opts& o,
                       double* H, double* S, int64_t ld, hsgen_report* rep, double ms_front)
{
    const int flags = o.flags;
    hsgen_status st;
    double2* Z = (double2*)ctx->Z.p;
    double2* Zs = (double2*)ctx->Zs.p;
    const bool need_h = !(flags & HSGEN_SKIP_H);
    double2* W = nullptr;
    if (need_h) {
        st = ensure(ctx, ctx->W, sizeof(double2) * (size_t)geo.ldz * geo.gp, true);
        if (st) return st;
        W = (double2*)ctx->W.p;
    }
    const bool hdev = H ? is_device_ptr(H) : true;
    const bool sdev = S ? is_device_ptr(S) : true;
    double2* dH = (double2*)H;
    double2* dS = (double2*)S;
    int64_t ldH = ld, ldS = ld;
    if (need_h && !hdev) {
        st = ensure(ctx, ctx->Htmp, sizeof(double2) * (size_t)ng * ng, false);
        if (st) return st;
        dH = (double2*)ctx->Htmp.p;
        ldH = ng;
    }
    if (!(flags & HSGEN_SKIP_S) && !sdev) {
        st = ensure(ctx, ctx->Stmp, sizeof(double2) * (size_t)ng * ng, false);
        if (st) return st;
        dS = (double2*)ctx->Stmp.p;
        ldS = ng;
    }
    st = ensure_M(ctx, s, geo.t.bk, geo.t.bm);
    if (st) return st;

    std::vector<GemmProblem> probs;
    std::vector<TileWork> work;
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    if (need_h) {
        const int tc = tapp_cfg(geo);
        add_tmul_problems(probs, work, s, geo, zgemm_cfg(tc), ng, Z, W);
        st = run_grouped(ctx, tc, probs, work);
        if (st) return st;
        // rows 2K .. ldz-1 of W are not written by the T application; clear
        // them so that no earlier, larger system's values remain there (the
        // split-precision column scaling reads whole columns)
        HIPCHK(launch_clear_rows(W, geo.ldz, 2 * (int64_t)s->kdim, geo.ldz, (int)geo.gp, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    probs.clear();
    work.clear();
    int icfg = 0;
    OzResult oz;
    if (o.precision == HSGEN_PREC_OZAKI_INT8) {
        icfg = (o.tile > 0 && igemm_cfg(o.tile).bm > 0) ? o.tile : kIgemmCfg;
        std::vector<OzK> ks(1);
        ks[0].ng = ng;
        ks[0].ncols = geo.gp;
        ks[0].dH = dH;
        ks[0].ldH = ldH;
        ks[0].dS = dS;
        ks[0].ldS = ldS;
        std::vector<OzResult> ozr;
        st = run_ozaki(ctx, geo, ks, flags, icfg, o.slices, o.target_rel_err, ozr);
        if (st) return st;
        oz = ozr[0];
    } else {
        add_hs_problems(probs, work, geo, ng, Z, Zs, W, dH, dS, 0, flags);
        xcd_permute(work);
        for (auto& p : probs) {
            p.ldl = geo.ldz;
            p.ldr = geo.ldz;
            p.kdepth = (int32_t)geo.kpad;
            p.ldc = (p.C == dH) ? ldH : ldS;
        }
        st = run_grouped(ctx, geo.cfg, probs, work);
        if (st) return st;
    }
    if (flags & HSGEN_FILL_UPPER) {
        if (need_h) HIPCHK(launch_fill_upper(dH, ldH, ng, ctx->stream));
        if (!(flags & HSGEN_SKIP_S)) HIPCHK(launch_fill_upper(dS, ldS, ng, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev[4]));
    if (need_h && !hdev) {
        st = copy_lower_to_host(ctx, dH, ng, H, ld, (flags & HSGEN_FILL_UPPER) != 0);
        if (st) return st;
    }
    if (!(flags & HSGEN_SKIP_S) && !sdev) {
        st = copy_lower_to_host(ctx, dS, ng, S, ld, (flags & HSGEN_FILL_UPPER) != 0);
        if (st) return st;
    }
    if (rep) {
        float t1 = 0, t2 = 0;
        HIPCHK(hipEventElapsedTime(&t1, ctx->ev[2], ctx->ev[3]));
        HIPCHK(hipEventElapsedTime(&t2, ctx->ev[3], ctx->ev[4]));
        rep->ms_tmul = t1;
        rep->ms_contract = t2;
        rep->ms_total = ms_front + t1 + t2;
        rep->ng = ng;
        rep->kdim = s->kdim;
        rep->flops_alg = 16.0 * (double)ng * ng * s->kdim + 32.0 * (double)ng * (double)s->sum_nlm2;
        rep->slices_used = oz.nmod;
        rep->err_est_h = oz.est_h;
        rep->err_est_s = oz.est_s;
        rep->ms_igemm = oz.ms_igemm;
        rep->err_smp_h = oz.smp_h;
        rep->err_smp_s = oz.smp_s;
        rep->tile_used = o.precision == HSGEN_PREC_OZAKI_INT8 ? icfg : geo.cfg;
    }
    return HSGEN_OK;
}

// ---------------------------------------------------------------------------
// Batched k-points, FP64: per batch of k-points that fits in memory, A/B
// generation per k into one column-concatenated Z/Zs, ONE grouped launch for
// the T application of all their columns and ONE persistent grouped launch
// for the S and H tiles of all of them (largest k first).  The launch
// sequence is captured in a hipGraph and replayed while the call signature
// (system, k-points and their G lists, outputs, workspace) is unchanged, as
// in repeated SCF iterations over one k-set.
// ---------------------------------------------------------------------------
uint64_t hash_words(const void* data, size_t bytes, uint64_t h)
{
    const unsigned char* c = (const unsigned char*)data;
    for (size_t i = 0; i < bytes; ++i) h = (h ^ c[i]) * 1099511628211ull;   // FNV-1a
    return h;
}

hsgen_status batched_fp64(hsgen_ctx* ctx, hsgen_sys* s, int nk, const hsgen_kpoint* kps, const hsgen_opts& o,
                          double* const* H, double* const* S, const int64_t* ld, hsgen_report* rep)
{
    const int flags = o.flags;
    const bool do_s = !(flags & HSGEN_SKIP_S), do_h = !(flags & HSGEN_SKIP_H);
    int max_ng = 1;
    for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
    const Geometry geo = geometry(o.tile, s->kdim, max_ng);
    const int64_t tm = std::max(geo.t.bm, geo.t.bn);
    const int tc = tapp_cfg(geo);
    hsgen_status st = ensure_M(ctx, s, geo.t.bk, geo.t.bm);
    if (st) return st;
    // largest k first
    std::vector<int> order(nk);
    for (int i = 0; i < nk; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return kps[a].ng > kps[b].ng; });
    std::vector<bool> hdev(nk), sdev(nk);
    for (int i = 0; i < nk; ++i) {
        hdev[i] = !do_h || is_device_ptr(H[i]);
        sdev[i] = !do_s || is_device_ptr(S[i]);
    }
    // batches: operands (Z, Zs, W) plus device copies of host outputs within
    // ~70% of the memory the call may use
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    size_t budget = (size_t)(0.7 * (double)(fr + ctx->Z.bytes + ctx->Zs.bytes + ctx->W.bytes + ctx->btmp.bytes));
    if (const char* e = std::getenv("HSGEN_BATCH_BYTES"))   // cap (tuning and tests)
        if (std::strtoull(e, nullptr, 10) > 0) budget = std::min<size_t>(budget, std::strtoull(e, nullptr, 10));
    std::vector<std::vector<int>> batches;
    {
        size_t used = 0;
        std::vector<int> cur;
        for (int i : order) {
            const size_t gp = (size_t)roundup(std::max(kps[i].ng, 1), tm);
            const size_t need = 3 * sizeof(double2) * (size_t)geo.ldz * gp +
                                sizeof(double2) * (size_t)kps[i].ng * kps[i].ng * ((!hdev[i]) + (!sdev[i]));
            if (!cur.empty() && used + need > budget) {
                batches.push_back(cur);
                cur.clear();
                used = 0;
            }
            cur.push_back(i);
            used += need;
        }
        if (!cur.empty()) batches.push_back(cur);
    }
    if (ctx->bgraphs.size() < batches.size()) ctx->bgraphs.resize(batches.size());
    const int pgrid = ctx->n_cu > 0 ? ctx->n_cu : 256;   // one 8-wave workgroup per CU (LDS-limited)
    double ms_all = 0.0;
    for (size_t b = 0; b < batches.size(); ++b) {
        const std::vector<int>& ks = batches[b];
        std::vector<int64_t> col0(ks.size()), tmp0(ks.size());
        int64_t ncols = 0, ntmp = 0;
        for (size_t q = 0; q < ks.size(); ++q) {
            const int i = ks[q];
            col0[q] = ncols;
            ncols += roundup(std::max(kps[i].ng, 1), tm);
            tmp0[q] = ntmp;
            ntmp += (int64_t)kps[i].ng * kps[i].ng * ((!hdev[i]) + (!sdev[i]));
        }
        const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * ncols;
        st = ensure(ctx, ctx->Z, zbytes, false);
        if (st) return st;
        st = ensure(ctx, ctx->Zs, zbytes, false);
        if (st) return st;
        if (do_h) {
            st = ensure(ctx, ctx->W, zbytes, true);
            if (st) return st;
        }
        if (ntmp > 0) {
            st = ensure(ctx, ctx->btmp, sizeof(double2) * (size_t)ntmp, false);
            if (st) return st;
        }
        double2* Z = (double2*)ctx->Z.p;
        double2* Zs = (double2*)ctx->Zs.p;
        double2* W = (double2*)ctx->W.p;
        // device outputs of every k of the batch (the caller's, or scratch)
        std::vector<double2*> dH(ks.size(), nullptr), dS(ks.size(), nullptr);
        std::vector<int64_t> ldH(ks.size()), ldS(ks.size());
        for (size_t q = 0; q < ks.size(); ++q) {
            const int i = ks[q];
            double2* t0 = (double2*)ctx->btmp.p + tmp0[q];
            const bool th = do_h && !hdev[i];
            const bool ts = do_s && !sdev[i];
            dH[q] = do_h ? (th ? t0 : (double2*)H[i]) : nullptr;
            dS[q] = do_s ? (ts ? t0 + (th ? (int64_t)kps[i].ng * kps[i].ng : 0) : (double2*)S[i]) : nullptr;
            ldH[q] = th ? kps[i].ng : ld[i];
            ldS[q] = ts ? kps[i].ng : ld[i];
        }
        // call signature
        std::vector<int64_t> sig{(int64_t)s->serial, (int64_t)flags, geo.cfg, tc, (int64_t)(intptr_t)Z,
                                 (int64_t)(intptr_t)Zs, (int64_t)(intptr_t)W, (int64_t)(intptr_t)s->d_M,
                                 (int64_t)(intptr_t)ctx->btmp.p, (int64_t)ks.size()};
        for (size_t q = 0; q < ks.size(); ++q) {
            const int i = ks[q];
            int64_t kb[3];
            std::memcpy(kb, kps[i].k, sizeof kb);
            sig.insert(sig.end(), {(int64_t)i, (int64_t)kps[i].ng, kb[0], kb[1], kb[2], ldH[q], ldS[q],
                                   (int64_t)(intptr_t)dH[q], (int64_t)(intptr_t)dS[q],
                                   (int64_t)hash_words(kps[i].g, sizeof(int32_t) * 3 * (size_t)kps[i].ng,
                                                       1469598103934665603ull)});
        }
        hsgen_ctx::BatchGraph& bg = ctx->bgraphs[b];
        if (!(bg.exec && bg.sig == sig)) {
            if (bg.exec) {
                HIPCHK(hipGraphExecDestroy(bg.exec));
                bg.exec = nullptr;
            }
            // G lists, concatenated in batch order
            std::vector<int32_t> gall;
            std::vector<int64_t> goff(ks.size());
            for (size_t q = 0; q < ks.size(); ++q) {
                goff[q] = (int64_t)gall.size();
                gall.insert(gall.end(), kps[ks[q]].g, kps[ks[q]].g + 3 * (size_t)kps[ks[q]].ng);
            }
            st = ensure(ctx, bg.g, sizeof(int32_t) * std::max<size_t>(gall.size(), 3), false);
            if (st) return st;
            // problems: T application over all columns, then S and H per k
            std::vector<GemmProblem> probs;
            std::vector<TileWork> tw_t, tw_c;
            if (do_h) add_tmul_problems(probs, tw_t, s, geo, zgemm_cfg(tc), ncols, Z, W);
            for (size_t q = 0; q < ks.size(); ++q) {
                const int i = ks[q];
                const size_t p0 = probs.size();
                add_hs_problems(probs, tw_c, geo, kps[i].ng, Z + col0[q] * geo.ldz, Zs + col0[q] * geo.ldz,
                                W + col0[q] * geo.ldz, dH[q], dS[q], 0, flags);
                for (size_t pi = p0; pi < probs.size(); ++pi) {
                    probs[pi].ldl = geo.ldz;
                    probs[pi].ldr = geo.ldz;
                    probs[pi].kdepth = (int32_t)geo.kpad;
                    probs[pi].ldc = probs[pi].C == dH[q] ? ldH[q] : ldS[q];
                }
            }
            xcd_permute(tw_c);
            std::vector<TileWork> work(tw_t);
            work.insert(work.end(), tw_c.begin(), tw_c.end());
            st = ensure(ctx, bg.probs, sizeof(GemmProblem) * std::max<size_t>(probs.size(), 1), false);
            if (st) return st;
            st = ensure(ctx, bg.work, sizeof(TileWork) * std::max<size_t>(work.size(), 1), false);
            if (st) return st;
            if (!gall.empty())
                HIPCHK(hipMemcpyAsync(bg.g.p, gall.data(), sizeof(int32_t) * gall.size(), hipMemcpyHostToDevice,
                                      ctx->stream));
            HIPCHK(hipMemcpyAsync(bg.probs.p, probs.data(), sizeof(GemmProblem) * probs.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
            if (!work.empty())
                HIPCHK(hipMemcpyAsync(bg.work.p, work.data(), sizeof(TileWork) * work.size(),
                                      hipMemcpyHostToDevice, ctx->stream));
            // the host vectors must outlive the copies
            HIPCHK(hipStreamSynchronize(ctx->stream));
            // capture the launch sequence
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (size_t q = 0; q < ks.size() && ce == hipSuccess; ++q) {
                const int i = ks[q];
                AbgenParams p = abgen_params(s, &kps[i], (const int32_t*)bg.g.p + goff[q]);
                p.ncols = (int32_t)roundup(std::max(kps[i].ng, 1), tm);
                p.Z = Z + col0[q] * geo.ldz;
                p.Zs = Zs + col0[q] * geo.ldz;
                p.ldz = geo.ldz;
                ce = launch_abgen(p, true, ctx->stream);
            }
            const GemmProblem* dp = (const GemmProblem*)bg.probs.p;
            const TileWork* dw = (const TileWork*)bg.work.p;
            if (ce == hipSuccess && do_h) {
                ce = launch_zgemm(tc, dp, dw, (int)tw_t.size(), ctx->stream, tc == geo.cfg ? pgrid : 0);
                if (ce == hipSuccess)
                    ce = launch_clear_rows(W, geo.ldz, 2 * (int64_t)s->kdim, geo.ldz, (int)ncols, ctx->stream);
            }
            if (ce == hipSuccess)
                ce = launch_zgemm(geo.cfg, dp, dw + tw_t.size(), (int)tw_c.size(), ctx->stream, pgrid);
            for (size_t q = 0; q < ks.size() && ce == hipSuccess && (flags & HSGEN_FILL_UPPER); ++q) {
                if (do_h) ce = launch_fill_upper(dH[q], ldH[q], kps[ks[q]].ng, ctx->stream);
                if (ce == hipSuccess && do_s) ce = launch_fill_upper(dS[q], ldS[q], kps[ks[q]].ng, ctx->stream);
            }
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(ctx, HSGEN_ERR_HIP, "graph capture failed: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            const hipError_t ie = hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                bg.exec = nullptr;
                return fail(ctx, HSGEN_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
            }
            bg.sig = sig;
        }
        HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
        HIPCHK(hipGraphLaunch(bg.exec, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        HIPCHK(hipEventSynchronize(ctx->ev[1]));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ms_all += ms;
        for (size_t q = 0; q < ks.size(); ++q) {
            const int i = ks[q];
            const bool full = (flags & HSGEN_FILL_UPPER) != 0;
            if (do_h && !hdev[i]) {
                st = copy_lower_to_host(ctx, dH[q], kps[i].ng, H[i], ld[i], full);
                if (st) return st;
            }
            if (do_s && !sdev[i]) {
                st = copy_lower_to_host(ctx, dS[q], kps[i].ng, S[i], ld[i], full);
                if (st) return st;
            }
        }
    }
    if (rep) {
        for (int i = 0; i < nk; ++i) {
            hsgen_report& r = rep[i];
            std::memset(&r, 0, sizeof r);
            r.ng = kps[i].ng;
            r.kdim = s->kdim;
            r.flops_alg = 16.0 * (double)kps[i].ng * kps[i].ng * s->kdim + 32.0 * (double)kps[i].ng * (double)s->sum_nlm2;
            r.bytes_abgen = 2.0 * 2.0 * 16.0 * (double)kps[i].ng * s->kdim;
            r.tile_used = geo.cfg;
        }
        if (nk > 0) rep[0].ms_total = ms_all;
    }
    return HSGEN_OK;
}

// ---------------------------------------------------------------------------
// Batched k-points, split precision: per batch of k-points that fits in
// memory, A/B generation per k into one column-concatenated Z/Zs (each k's
// columns padded to a multiple of both INT8 tile sides), ONE grouped launch
// for the T application of all their columns, then one split-precision
// evaluation of the whole batch: per modulus count that its k-points need,
// one INT8 launch per modulus over the tiles of all of them.  Each k-point
// gets the modulus count its own error model asks for, so its H and S are
// those of a single-k call, bit for bit.
// ---------------------------------------------------------------------------
hsgen_status batched_split(hsgen_ctx* ctx, hsgen_sys* s, int nk, const hsgen_kpoint* kps, const hsgen_opts& o,
                           double* const* H, double* const* S, const int64_t* ld, hsgen_report* rep)
{
    const int flags = o.flags;
    const bool do_s = !(flags & HSGEN_SKIP_S), do_h = !(flags & HSGEN_SKIP_H);
    const int icfg = (o.tile > 0 && igemm_cfg(o.tile).bm > 0) ? o.tile : kIgemmCfg;
    const IgemmLaunchCfg ic = igemm_cfg(icfg);
    int max_ng = 1;
    for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
    const int64_t align = int8_col_align(icfg);
    const Geometry g0 = geometry(0, s->kdim, max_ng, align);
    const int64_t tm = std::max<int64_t>({(int64_t)g0.t.bm, (int64_t)g0.t.bn, align});
    const int64_t kd = roundup(g0.kpad, ic.bk);
    const int narr = ic.needs_rot == 1 ? 4 : ic.needs_rot == 2 ? 5 : 3;
    hsgen_status st = ensure_M(ctx, s, g0.t.bk, g0.t.bm);
    if (st) return st;
    if (rep) std::memset(rep, 0, sizeof(hsgen_report) * nk);
    // largest k first; k-points without G-vectors have nothing to compute
    std::vector<int> order;
    for (int i = 0; i < nk; ++i)
        if (kps[i].ng > 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return kps[a].ng > kps[b].ng; });
    std::vector<bool> hdev(nk), sdev(nk);
    for (int i = 0; i < nk; ++i) {
        hdev[i] = !do_h || is_device_ptr(H[i]);
        sdev[i] = !do_s || is_device_ptr(S[i]);
    }
    // batches within ~70% of the memory the call may use: operands Z, Zs, W,
    // device copies of host outputs, the stored residue outputs (2 bytes per
    // element of every S/H tile and modulus, counted for 14 moduli or the
    // requested count) and the residue arrays of as many moduli as one
    // residue pass converts when the moduli take two passes of at most 16
    // (HSGEN_RESIDUE_PASSES for tuning; each pass reads Z, Zs, W once, and
    // one pass per batch saves those reads but needs more, smaller batches,
    // which cost as much again at 64 k-points)
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    size_t held = 0;
    for (const DevBuf* b : {&ctx->Z, &ctx->Zs, &ctx->W, &ctx->btmp, &ctx->oz_res, &ctx->oz_out, &ctx->oz_smp})
        held += b->bytes;
    size_t budget = (size_t)(0.7 * (double)(fr + held));
    if (const char* e = std::getenv("HSGEN_BATCH_BYTES"))   // cap (tuning and tests)
        if (std::strtoull(e, nullptr, 10) > 0) budget = std::min<size_t>(budget, std::strtoull(e, nullptr, 10));
    const int nmod_est = o.slices > 0 ? o.slices : 14;
    int npass = 2;
    if (const char* e = std::getenv("HSGEN_RESIDUE_PASSES"))   // tuning
        npass = std::max(1, std::atoi(e));
    npass = std::max(npass, (nmod_est + 15) / 16);
    const int gslots = (nmod_est + npass - 1) / npass;
    auto ncols_of = [&](int i) { return roundup(std::max(kps[i].ng, 1), tm); };
    std::vector<size_t> need(nk, 0);
    size_t total = 0;
    for (int i : order) {
        const size_t gp = (size_t)ncols_of(i);
        const size_t ng = (size_t)kps[i].ng;
        const size_t ntr = (ng + ic.bm - 1) / ic.bm, ntc = (ng + ic.bn - 1) / ic.bn;
        size_t tiles = 0;
        for (size_t bi = 0; bi < ntr; ++bi)
            tiles += std::min(ntc, ((bi + 1) * ic.bm + ic.bn - 1) / ic.bn);
        need[i] = 3 * sizeof(double2) * (size_t)g0.ldz * gp + sizeof(double2) * ng * ng * ((!hdev[i]) + (!sdev[i])) +
                  (size_t)(do_s + do_h) * tiles * 2 * ic.bm * ic.bn * nmod_est + (size_t)gslots * narr * 2 * kd * gp;
        total += need[i];
    }
    // consecutive k-points (largest first) into batches of at most `cap`
    auto split = [&](size_t cap) {
        std::vector<std::vector<int>> bs;
        size_t used = 0;
        for (int i : order) {
            if (bs.empty() || used + need[i] > cap) {
                bs.emplace_back();
                used = 0;
            }
            bs.back().push_back(i);
            used += need[i];
        }
        return bs;
    };
    // as few batches as the budget allows, of even size (a small last batch
    // would leave its launches short of work)
    std::vector<std::vector<int>> batches = split(budget);
    const size_t nbatch = batches.size();
    for (int step = 0; step < 50 && nbatch > 1; ++step) {
        const size_t cap = (size_t)((double)total / nbatch * (1.0 + 0.02 * step));
        if (cap >= budget) break;
        std::vector<std::vector<int>> bs = split(cap);
        if (bs.size() <= nbatch) {
            batches.swap(bs);
            break;
        }
    }
    double ms_all = 0.0;
    for (const std::vector<int>& bk : batches) {
        std::vector<int64_t> col0(bk.size()), tmp0(bk.size());
        int64_t ncols = 0, ntmp = 0;
        for (size_t q = 0; q < bk.size(); ++q) {
            const int i = bk[q];
            col0[q] = ncols;
            ncols += ncols_of(i);
            tmp0[q] = ntmp;
            ntmp += (int64_t)kps[i].ng * kps[i].ng * ((!hdev[i]) + (!sdev[i]));
        }
        Geometry geo = g0;
        geo.gp = ncols;
        const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * ncols;
        st = ensure(ctx, ctx->Z, zbytes, false);
        if (st) return st;
        st = ensure(ctx, ctx->Zs, zbytes, false);
        if (st) return st;
        if (do_h) {
            st = ensure(ctx, ctx->W, zbytes, true);
            if (st) return st;
        }
        if (ntmp > 0) {
            st = ensure(ctx, ctx->btmp, sizeof(double2) * (size_t)ntmp, false);
            if (st) return st;
        }
        double2* Z = (double2*)ctx->Z.p;
        double2* Zs = (double2*)ctx->Zs.p;
        double2* W = (double2*)ctx->W.p;
        // G lists, concatenated in batch order
        std::vector<int32_t> gall;
        std::vector<int64_t> goff(bk.size());
        for (size_t q = 0; q < bk.size(); ++q) {
            goff[q] = (int64_t)gall.size();
            gall.insert(gall.end(), kps[bk[q]].g, kps[bk[q]].g + 3 * (size_t)kps[bk[q]].ng);
        }
        st = ensure(ctx, ctx->g, sizeof(int32_t) * gall.size(), false);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(ctx->g.p, gall.data(), sizeof(int32_t) * gall.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        // outputs of every k of the batch (the caller's device matrices, or scratch)
        std::vector<OzK> ks(bk.size());
        for (size_t q = 0; q < bk.size(); ++q) {
            const int i = bk[q];
            double2* t0 = (double2*)ctx->btmp.p + tmp0[q];
            const bool th = do_h && !hdev[i];
            const bool ts = do_s && !sdev[i];
            OzK& k = ks[q];
            k.ng = kps[i].ng;
            k.col0 = col0[q];
            k.ncols = ncols_of(i);
            k.dH = do_h ? (th ? t0 : (double2*)H[i]) : nullptr;
            k.dS = do_s ? (ts ? t0 + (th ? (int64_t)k.ng * k.ng : 0) : (double2*)S[i]) : nullptr;
            k.ldH = th ? k.ng : ld[i];
            k.ldS = ts ? k.ng : ld[i];
        }
        HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
        for (size_t q = 0; q < bk.size(); ++q) {
            const int i = bk[q];
            AbgenParams p = abgen_params(s, &kps[i], (const int32_t*)ctx->g.p + goff[q]);
            p.ncols = (int32_t)ks[q].ncols;
            p.Z = Z + col0[q] * geo.ldz;
            p.Zs = Zs + col0[q] * geo.ldz;
            p.ldz = geo.ldz;
            HIPCHK(launch_abgen(p, true, ctx->stream));
        }
        if (do_h) {
            // T application over every column of the batch in one grouped launch
            std::vector<GemmProblem> probs;
            std::vector<TileWork> work;
            const int tc = tapp_cfg(geo);
            add_tmul_problems(probs, work, s, geo, zgemm_cfg(tc), ncols, Z, W);
            st = run_grouped(ctx, tc, probs, work);
            if (st) return st;
            HIPCHK(launch_clear_rows(W, geo.ldz, 2 * (int64_t)s->kdim, geo.ldz, (int)ncols, ctx->stream));
        }
        std::vector<OzResult> ozr;
        st = run_ozaki(ctx, geo, ks, flags, icfg, o.slices, o.target_rel_err, ozr);
        if (st) return st;
        for (size_t q = 0; q < bk.size() && (flags & HSGEN_FILL_UPPER); ++q) {
            if (do_h) HIPCHK(launch_fill_upper(ks[q].dH, ks[q].ldH, ks[q].ng, ctx->stream));
            if (do_s) HIPCHK(launch_fill_upper(ks[q].dS, ks[q].ldS, ks[q].ng, ctx->stream));
        }
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        HIPCHK(hipEventSynchronize(ctx->ev[1]));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ms_all += ms;
        for (size_t q = 0; q < bk.size(); ++q) {
            const int i = bk[q];
            const bool full = (flags & HSGEN_FILL_UPPER) != 0;
            if (do_h && !hdev[i]) {
                st = copy_lower_to_host(ctx, ks[q].dH, kps[i].ng, H[i], ld[i], full);
                if (st) return st;
            }
            if (do_s && !sdev[i]) {
                st = copy_lower_to_host(ctx, ks[q].dS, kps[i].ng, S[i], ld[i], full);
                if (st) return st;
            }
            if (rep) {
                hsgen_report& r = rep[i];
                r.slices_used = ozr[q].nmod;
                r.err_est_h = ozr[q].est_h;
                r.err_est_s = ozr[q].est_s;
                r.ms_igemm = ozr[q].ms_igemm;
                r.err_smp_h = ozr[q].smp_h;
                r.err_smp_s = ozr[q].smp_s;
            }
        }
    }
    if (rep) {
        for (int i = 0; i < nk; ++i) {
            hsgen_report& r = rep[i];
            r.ng = kps[i].ng;
            r.kdim = s->kdim;
            r.flops_alg = 16.0 * (double)kps[i].ng * kps[i].ng * s->kdim + 32.0 * (double)kps[i].ng * (double)s->sum_nlm2;
            r.bytes_abgen = 2.0 * 2.0 * 16.0 * (double)kps[i].ng * s->kdim;
            r.tile_used = icfg;
        }
        if (nk > 0) rep[0].ms_total = ms_all;
    }
    return HSGEN_OK;
}

}  // namespace

extern "C" {

hsgen_status hsgen_construct_ab(hsgen_ctx* ctx, const hsgen_sys* s, const hsgen_kpoint* kp,
                                double* A, double* B)
{
    if (!ctx || !s || !A || !B) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_status st = check_kpoint(ctx, kp);
    if (st) return st;
    if (kp->ng == 0) return HSGEN_OK;
    HIPCHK(hipSetDevice(ctx->device));
    st = upload_g(ctx, kp);
    if (st) return st;
    const size_t n = (size_t)s->L * s->natoms * kp->ng;
    const bool adev = is_device_ptr(A), bdev = is_device_ptr(B);
    double2* dA = (double2*)A;
    double2* dB = (double2*)B;
    if (!adev) {
        st = ensure(ctx, ctx->A, sizeof(double2) * n, false);
        if (st) return st;
        dA = (double2*)ctx->A.p;
    }
    if (!bdev) {
        st = ensure(ctx, ctx->B, sizeof(double2) * n, false);
        if (st) return st;
        dB = (double2*)ctx->B.p;
    }
    AbgenParams p = abgen_params(s, kp, (const int32_t*)ctx->g.p);
    p.ncols = kp->ng;
    p.A = dA;
    p.B = dB;
    HIPCHK(launch_abgen(p, false, ctx->stream));
    if (!adev) HIPCHK(hipMemcpyAsync(A, dA, sizeof(double2) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (!bdev) HIPCHK(hipMemcpyAsync(B, dB, sizeof(double2) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return HSGEN_OK;
}

hsgen_status hsgen_create_hs(hsgen_ctx* ctx, const hsgen_sys* sc, const hsgen_kpoint* kp,
                             const hsgen_opts* opts, double* H, double* S, int64_t ld,
                             hsgen_report* rep)
{
    if (!ctx || !sc) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_sys* s = const_cast<hsgen_sys*>(sc);
    hsgen_status st = check_kpoint(ctx, kp);
    if (st) return st;
    hsgen_opts o{};
    if (opts) o = *opts;
    if (o.precision != HSGEN_PREC_FP64 && o.precision != HSGEN_PREC_OZAKI_INT8)
        return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
    if (o.slices < 0 || o.slices > kMaxModuli)
        return fail(ctx, HSGEN_ERR_ARG, "slices %d outside [0, %d]", o.slices, kMaxModuli);
    const int ng = kp->ng;
    if (ld < std::max(ng, 1)) return fail(ctx, HSGEN_ERR_ARG, "ld %lld < ng %d", (long long)ld, ng);
    o.tile = pick_tile(ctx, o, ng, s->kdim);
    if ((!H && !(o.flags & HSGEN_SKIP_H)) || (!S && !(o.flags & HSGEN_SKIP_S)))
        return fail(ctx, HSGEN_ERR_ARG, "null output matrix");
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (ng == 0) return HSGEN_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const Geometry geo = geometry(o.precision == HSGEN_PREC_FP64 ? o.tile : 0, s->kdim, ng,
                                  o.precision == HSGEN_PREC_FP64 ? 1 : int8_col_align(o.tile));
    st = upload_g(ctx, kp);
    if (st) return st;
    const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * geo.gp;
    st = ensure(ctx, ctx->Z, zbytes, false);
    if (st) return st;
    st = ensure(ctx, ctx->Zs, zbytes, false);
    if (st) return st;
    AbgenParams p = abgen_params(s, kp, (const int32_t*)ctx->g.p);
    p.ncols = (int32_t)geo.gp;
    p.Z = (double2*)ctx->Z.p;
    p.Zs = (double2*)ctx->Zs.p;
    p.ldz = geo.ldz;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(launch_abgen(p, true, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    st = finish_hs(ctx, s, geo, ng, o, H, S, ld, rep, 0.0);
    if (st) return st;
    if (rep) {
        float t0 = 0;
        HIPCHK(hipEventElapsedTime(&t0, ctx->ev[0], ctx->ev[1]));
        rep->ms_abgen = t0;
        rep->ms_total += t0;
        rep->bytes_abgen = 2.0 * 2.0 * 16.0 * (double)ng * s->kdim;   // Z and Zs, A and B rows
    }
    return HSGEN_OK;
}

hsgen_status hsgen_create_hs_batched(hsgen_ctx* ctx, const hsgen_sys* sys, int32_t nk,
                                     const hsgen_kpoint* kps, const hsgen_opts* opts,
                                     double* const* H, double* const* S, const int64_t* ld,
                                     hsgen_report* rep)
{
    if (!ctx || !sys || nk < 0 || (nk > 0 && (!kps || !H || !S || !ld)))
        return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_opts o{};
    if (opts) o = *opts;
    if (o.precision != HSGEN_PREC_FP64 && o.precision != HSGEN_PREC_OZAKI_INT8)
        return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
    for (int i = 0; i < nk; ++i) {
        hsgen_status st = check_kpoint(ctx, &kps[i]);
        if (st) return st;
        if (ld[i] < std::max(kps[i].ng, 1))
            return fail(ctx, HSGEN_ERR_ARG, "k-point %d: ld %lld < ng %d", i, (long long)ld[i], kps[i].ng);
        if ((!H[i] && !(o.flags & HSGEN_SKIP_H)) || (!S[i] && !(o.flags & HSGEN_SKIP_S)))
            return fail(ctx, HSGEN_ERR_ARG, "k-point %d: null output matrix", i);
    }
    if (nk == 0) return HSGEN_OK;
    if (o.precision == HSGEN_PREC_FP64) {
        HIPCHK(hipSetDevice(ctx->device));
        int max_ng = 1;
        for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
        o.tile = pick_tile(ctx, o, max_ng, sys->kdim);
        return batched_fp64(ctx, const_cast<hsgen_sys*>(sys), nk, kps, o, H, S, ld, rep);
    }
    if (o.slices < 0 || o.slices > kMaxModuli)
        return fail(ctx, HSGEN_ERR_ARG, "slices %d outside [0, %d]", o.slices, kMaxModuli);
    HIPCHK(hipSetDevice(ctx->device));
    int max_ng = 1;
    for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
    o.tile = pick_tile(ctx, o, max_ng, sys->kdim);
    return batched_split(ctx, const_cast<hsgen_sys*>(sys), nk, kps, o, H, S, ld, rep);
}

hsgen_status hsgen_autotune(hsgen_ctx* ctx, const hsgen_sys* sys, const hsgen_kpoint* kp, int32_t precision,
                            int32_t* best_tile, double* best_ms)
{
    if (!ctx || !sys) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_status st = check_kpoint(ctx, kp);
    if (st) return st;
    if (precision != HSGEN_PREC_FP64 && precision != HSGEN_PREC_OZAKI_INT8)
        return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", precision);
    const int ng = kp->ng;
    if (ng <= 0) return fail(ctx, HSGEN_ERR_ARG, "autotuning needs ng > 0");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<int> cands;
    if (precision == HSGEN_PREC_FP64) {
        for (int id = 1; id < 64; ++id)
            if (zgemm_cfg(id).bm > 0) cands.push_back(id);
    } else {
        for (int id = 1; id < 64; ++id)
            if (igemm_cfg(id).bm > 0) cands.push_back(id);
    }
    // device outputs owned by the call
    st = ensure(ctx, ctx->Htmp, sizeof(double2) * (size_t)ng * ng, false);
    if (st) return st;
    st = ensure(ctx, ctx->Stmp, sizeof(double2) * (size_t)ng * ng, false);
    if (st) return st;
    double* dH = (double*)ctx->Htmp.p;
    double* dS = (double*)ctx->Stmp.p;
    int best = 0;
    double bms = 0.0;
    for (int id : cands) {
        hsgen_opts o{};
        o.precision = precision;
        o.tile = id;
        // split precision: a fixed modulus count; the INT8 GEMM time per
        // modulus is what the configuration changes
        o.slices = precision == HSGEN_PREC_FP64 ? 0 : 10;
        hsgen_report r{};
        double ms = 0.0;
        for (int rep = 0; rep < 2; ++rep) {   // warm-up, then the timed run
            st = hsgen_create_hs(ctx, sys, kp, &o, dH, dS, ng, &r);
            if (st) return st;
            ms = precision == HSGEN_PREC_FP64 ? r.ms_contract : r.ms_igemm;
        }
        if (best == 0 || ms < bms) {
            best = id;
            bms = ms;
        }
    }
    ctx->tuned[tune_key(precision, ng, sys->kdim)] = best;
    if (best_tile) *best_tile = best;
    if (best_ms) *best_ms = bms;
    return HSGEN_OK;
}

hsgen_status hsgen_contract_hs(hsgen_ctx* ctx, int32_t L, int32_t a, int32_t G, const double* A,
                               const double* B, const int32_t* t_size, const double* const* t_aa,
                               const double* const* t_ab, const double* const* t_bb,
                               const int32_t* l_maxs, int32_t max_lmax, const double* sqrt_u_norms,
                               const hsgen_opts* opts, double* H, double* S, int64_t ld,
                               hsgen_report* rep)
{
    if (!ctx) return HSGEN_ERR_ARG;
    if (L <= 0 || a <= 0 || G < 0) return fail(ctx, HSGEN_ERR_ARG, "bad dimensions L=%d a=%d G=%d", L, a, G);
    if (!A || !B || !t_size || !t_aa || !t_ab || !t_bb || !l_maxs || !sqrt_u_norms)
        return fail(ctx, HSGEN_ERR_ARG, "null argument");
    if (max_lmax < 0 || (max_lmax + 1) * (max_lmax + 1) > L)
        return fail(ctx, HSGEN_ERR_ARG, "max_lmax %d inconsistent with L %d", max_lmax, L);
    hsgen_opts o{};
    if (opts) o = *opts;
    if (o.precision != HSGEN_PREC_FP64 && o.precision != HSGEN_PREC_OZAKI_INT8)
        return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
    if (o.slices < 0 || o.slices > kMaxModuli)
        return fail(ctx, HSGEN_ERR_ARG, "slices %d outside [0, %d]", o.slices, kMaxModuli);
    if (ld < std::max(G, 1)) return fail(ctx, HSGEN_ERR_ARG, "ld < G");
    if (rep) std::memset(rep, 0, sizeof *rep);
    HIPCHK(hipSetDevice(ctx->device));
    // a throw-away system holding only what the contraction needs
    hsgen_sys s;
    s.natoms = a;
    s.max_lmax = max_lmax;
    s.nl = max_lmax + 1;
    s.L = L;
    int off = 0;
    s.atoms.resize(a);
    std::vector<double> un2((size_t)a * s.nl, 0.0);
    for (int i = 0; i < a; ++i) {
        const int lm = l_maxs[i];
        if (lm < 0 || lm > max_lmax) return fail(ctx, HSGEN_ERR_ARG, "l_maxs[%d] out of range", i);
        const int nlm = (lm + 1) * (lm + 1);
        if (t_size[i] != nlm)
            return fail(ctx, HSGEN_ERR_ARG, "t_size[%d] = %d but (l_maxs+1)^2 = %d", i, t_size[i], nlm);
        AtomDev& at = s.atoms[i];
        std::memset(&at, 0, sizeof at);
        at.lmax = lm;
        at.nlm = nlm;
        at.off = off;
        off += nlm;
        s.sum_nlm2 += (int64_t)nlm * nlm;
        for (int l = 0; l <= lm; ++l) {
            const double v = sqrt_u_norms[(size_t)i * s.nl + l];
            un2[(size_t)i * s.nl + l] = v * v;
        }
    }
    s.kdim = off;
    if (G == 0) return HSGEN_OK;
    o.tile = pick_tile(ctx, o, G, s.kdim);
    const Geometry geo = geometry(o.precision == HSGEN_PREC_FP64 ? o.tile : 0, s.kdim, G,
                                  o.precision == HSGEN_PREC_FP64 ? 1 : int8_col_align(o.tile));
    hsgen_status st = upload_M(ctx, &s, t_size, t_aa, t_ab, t_bb, geo.t.bk, geo.t.bm);
    if (st) return st;
    struct Guard {
        hsgen_sys* s;
        ~Guard() { if (s->d_M) (void)hipFree(s->d_M); s->d_M = nullptr; }
    } guard{&s};
    // device copies of A, B if they are host memory
    const size_t nab = (size_t)L * a * G;
    const double2* dA = (const double2*)A;
    const double2* dB = (const double2*)B;
    if (!is_device_ptr(A)) {
        st = ensure(ctx, ctx->A, sizeof(double2) * nab, false);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(ctx->A.p, A, sizeof(double2) * nab, hipMemcpyHostToDevice, ctx->stream));
        dA = (const double2*)ctx->A.p;
    }
    if (!is_device_ptr(B)) {
        st = ensure(ctx, ctx->B, sizeof(double2) * nab, false);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(ctx->B.p, B, sizeof(double2) * nab, hipMemcpyHostToDevice, ctx->stream));
        dB = (const double2*)ctx->B.p;
    }
    st = ensure(ctx, ctx->atoms_tmp, sizeof(AtomDev) * a, false);
    if (st) return st;
    st = ensure(ctx, ctx->unorm2, sizeof(double) * un2.size(), false);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(ctx->atoms_tmp.p, s.atoms.data(), sizeof(AtomDev) * a, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->unorm2.p, un2.data(), sizeof(double) * un2.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * geo.gp;
    st = ensure(ctx, ctx->Z, zbytes, false);
    if (st) return st;
    st = ensure(ctx, ctx->Zs, zbytes, false);
    if (st) return st;
    PackParams pp;
    pp.A = dA;
    pp.B = dB;
    pp.atoms = (const AtomDev*)ctx->atoms_tmp.p;
    pp.unorm2 = (const double*)ctx->unorm2.p;
    pp.natoms = a;
    pp.L = L;
    pp.nl = s.nl;
    pp.ng = G;
    pp.kdim = s.kdim;
    pp.ldab = (int64_t)L * a;
    pp.Z = (double2*)ctx->Z.p;
    pp.Zs = (double2*)ctx->Zs.p;
    pp.ldz = geo.ldz;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(launch_pack_ref(pp, (int)geo.gp, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    st = finish_hs(ctx, &s, geo, G, o, H, S, ld, rep, 0.0);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (st) return st;
    if (rep) {
        float t0 = 0;
        HIPCHK(hipEventElapsedTime(&t0, ctx->ev[0], ctx->ev[1]));
        rep->ms_abgen = t0;
        rep->ms_total += t0;
    }
    return HSGEN_OK;
}


// The step after H/S generation: the generalized Hermitian eigenproblem
// H x = lambda S x (Fleur hands H and S to its own solvers, eigen.F:750,781;
// the reference stops at H and S).  A library call rather than a hand-written
// kernel: rocSOLVER's zhegvd (Cholesky of S, reduction to a standard problem,
// divide and conquer) on the context's stream, on copies of the lower
// triangles hsgen writes.
hsgen_status hsgen_eigensolve(hsgen_ctx* ctx, int32_t n, const double* H, const double* S, int64_t ld, double* w,
                              double* V, int64_t ldv, double* ms)
{
    if (!ctx) return HSGEN_ERR_ARG;
    if (n < 0 || !H || !S || !w) return fail(ctx, HSGEN_ERR_ARG, "null argument or n < 0");
    if (ld < std::max(n, 1) || (V && ldv < std::max(n, 1)))
        return fail(ctx, HSGEN_ERR_ARG, "leading dimension smaller than n = %d", n);
    if (n == 0) return HSGEN_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nn = (size_t)n * n;
    hsgen_status st = ensure(ctx, ctx->eig_a, sizeof(double2) * nn, false);
    if (st) return st;
    st = ensure(ctx, ctx->eig_b, sizeof(double2) * nn, false);
    if (st) return st;
    st = ensure(ctx, ctx->eig_d, sizeof(double) * 2 * (size_t)n + sizeof(rocblas_int), false);
    if (st) return st;
    double2* A = (double2*)ctx->eig_a.p;
    double2* B = (double2*)ctx->eig_b.p;
    double* D = (double*)ctx->eig_d.p;
    double* E = D + n;
    rocblas_int* info = (rocblas_int*)(E + n);
    if (!ctx->blas) {
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
            ctx->blas = nullptr;
            return fail(ctx, HSGEN_ERR_HIP, "rocblas_create_handle failed");
        }
    }
    if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
        return fail(ctx, HSGEN_ERR_HIP, "rocblas_set_stream failed");
    const size_t row = sizeof(double2) * (size_t)n;
    HIPCHK(hipMemcpy2DAsync(A, row, H, sizeof(double2) * ld, row, n, hipMemcpyDefault, ctx->stream));
    HIPCHK(hipMemcpy2DAsync(B, row, S, sizeof(double2) * ld, row, n, hipMemcpyDefault, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    const rocblas_status rs = rocsolver_zhegvd(ctx->blas, rocblas_eform_ax,
                                               V ? rocblas_evect_original : rocblas_evect_none, rocblas_fill_lower, n,
                                               (rocblas_double_complex*)A, n, (rocblas_double_complex*)B, n, D, E,
                                               info);
    if (rs != rocblas_status_success)
        return fail(ctx, HSGEN_ERR_HIP, "rocsolver_zhegvd: %s", rocblas_status_to_string(rs));
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    rocblas_int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(w, D, sizeof(double) * n, hipMemcpyDefault, ctx->stream));
    if (V) HIPCHK(hipMemcpy2DAsync(V, sizeof(double2) * ldv, A, row, row, n, hipMemcpyDefault, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hinfo > n) return fail(ctx, HSGEN_ERR_ARG, "S is not positive definite (leading minor %d)", (int)(hinfo - n));
    if (hinfo > 0) return fail(ctx, HSGEN_ERR_ACCURACY, "eigensolver did not converge (info %d)", (int)hinfo);
    if (ms) {
        float tm = 0.0f;
        HIPCHK(hipEventElapsedTime(&tm, ctx->ev[0], ctx->ev[1]));
        *ms = tm;
    }
    return HSGEN_OK;
}

}  // extern "C"
