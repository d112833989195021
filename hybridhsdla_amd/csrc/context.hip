// Contexts, device-resident systems and workspace of the hsgen C ABI
// (include/hsgen.h): hsgen_init / hsgen_finalize (replacing dev_init /
// dev_finalize, wrappers.h:6-7), hsgen_system_create (the state of AtomData
// after generate_matchdata and augment_tmats, AtomData.h:16-52), and the
// host-side helpers every path shares.
#include <algorithm>
#include <array>
#include <atomic>
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

#include "hsgen_host.h"

namespace hsgen {

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

namespace {

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

}  // namespace

bool fold_enabled()
{
    const char* e = std::getenv("HSGEN_FOLD");
    return !(e && e[0] == '0');
}

bool fold_ok(double n)
{
    // 1 / sqrt(N) and products of two of them stay far from overflow
    return n > 1e-150 && n < 1e150;
}

// Build the per-atom M_a = [[T_aa, T_ab], [T_ab^H, T_bb]] (T_aa, T_bb taken
// from their lower triangles) padded for tile geometry (bk, bm), deduplicated
// by T pointer triple (and B-row scales), and upload.  With rsc (folded
// operands, rsc[a * nl + l] = sqrt(N_l) of atom a) the block is
// M'_a = D^-1/2 M_a D^-1/2, D = diag(1 on A rows, N_l on B rows), so that
// Z^H M Z = Y^H M' Y with Y = D^1/2 Z.
hsgen_status upload_M(hsgen_ctx* ctx, hsgen_sys* s, const int32_t* nlm, const double* const* taa,
                      const double* const* tab, const double* const* tbb, const double* rsc, int bk, int bm)
{
    if (s->d_M) {
        (void)hipFree(s->d_M);
        s->d_M = nullptr;
        s->d_mloc = nullptr;
    }
    const int na = s->natoms;
    const int nl = s->nl;
    s->m_off.assign(na, 0);
    s->m_ld.assign(na, 0);
    s->m_cols.assign(na, 0);
    std::map<std::pair<std::vector<const void*>, std::vector<double>>, int> seen;
    std::vector<int> owner(na, -1);
    int64_t total = 0;
    for (int a = 0; a < na; ++a) {
        std::pair<std::vector<const void*>, std::vector<double>> key{
            {taa[a], tab[a], tbb[a], (const void*)(intptr_t)nlm[a]},
            rsc ? std::vector<double>(rsc + (size_t)a * nl, rsc + (size_t)(a + 1) * nl) : std::vector<double>()};
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
    std::vector<double> sq;
    for (int a = 0; a < na; ++a) {
        if (owner[a] != a) continue;
        const int n = nlm[a];
        const int64_t ld = s->m_ld[a];
        std::complex<double>* M = hM.data() + s->m_off[a];
        const std::complex<double>* Taa = (const std::complex<double>*)taa[a];
        const std::complex<double>* Tab = (const std::complex<double>*)tab[a];
        const std::complex<double>* Tbb = (const std::complex<double>*)tbb[a];
        // B-row scale per lm (1 without folding)
        sq.assign(n, 1.0);
        if (rsc)
            for (int lm = 0; lm < n; ++lm) {
                int l = (int)std::sqrt((double)lm + 0.5);
                while ((l + 1) * (l + 1) <= lm) ++l;
                while (l * l > lm) --l;
                sq[lm] = rsc[(size_t)a * nl + l];
            }
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                // Hermitian expansion from the lower triangle
                std::complex<double> aa = r >= c ? Taa[r + (int64_t)n * c] : std::conj(Taa[c + (int64_t)n * r]);
                std::complex<double> bb = r >= c ? Tbb[r + (int64_t)n * c] : std::conj(Tbb[c + (int64_t)n * r]);
                if (r == c) { aa = aa.real(); bb = bb.real(); }
                std::complex<double> ab = Tab[r + (int64_t)n * c];              // T_ab block
                std::complex<double> ba = std::conj(Tab[c + (int64_t)n * r]);   // T_ab^H block
                if (rsc) {
                    ab /= sq[c];
                    ba /= sq[r];
                    bb /= sq[r] * sq[c];
                }
                M[r + ld * c] = aa;
                M[(n + r) + ld * (n + c)] = bb;
                M[r + ld * (n + c)] = ab;
                M[(n + r) + ld * c] = ba;
            }
    }
    // the per-atom layout table (offset, leading dimension) follows the
    // blocks in the same allocation, for kernels that build their own work
    // lists
    const size_t mbytes = (size_t)roundup((int64_t)(sizeof(double2) * hM.size()), 256);
    std::vector<MLoc> loc(na);
    for (int a = 0; a < na; ++a) loc[a] = MLoc{s->m_off[a], s->m_ld[a], s->m_cols[a]};
    HIPCHK(hipMalloc(&s->d_M, mbytes + sizeof(MLoc) * na));
    HIPCHK(hipMemcpy(s->d_M, hM.data(), sizeof(double2) * hM.size(), hipMemcpyHostToDevice));
    s->d_mloc = (MLoc*)((char*)s->d_M + mbytes);
    HIPCHK(hipMemcpy(s->d_mloc, loc.data(), sizeof(MLoc) * na, hipMemcpyHostToDevice));
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
    return upload_M(ctx, s, nlm.data(), pa.data(), pb.data(), pc.data(), s->fold ? s->h_sqrtn.data() : nullptr,
                    bk, bm);
}

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

// Copy the lower triangle (and optionally everything) of a device matrix with
// ld = n into a host matrix with leading dimension ld: blocks of columns go
// through two pinned staging buffers (the DMA of one block overlaps the
// host-side copy of the previous one, which several threads share), so that
// only the requested triangle of the caller's matrix is written.
hsgen_status copy_lower_to_host(hsgen_ctx* ctx, const double2* dsrc, int n, double* H, int64_t ld,
                                bool full)
{
    if (n <= 0) return HSGEN_OK;
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

}  // namespace hsgen

using namespace hsgen;

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

namespace {

// The context's stream and timing events on its device.
hsgen_status init_stream(hsgen_ctx* ctx)
{
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto& ev : ctx->ev) HIPCHK(hipEventCreate(&ev));
    for (auto& ev : ctx->oz_ev) HIPCHK(hipEventCreate(&ev));
    return HSGEN_OK;
}

}  // namespace

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
    // a failed set-up releases what was created: the caller gets no context
    hsgen_status st = init_stream(ctx);
    if (st == HSGEN_OK) st = noalloc_setup(ctx);
    if (st != HSGEN_OK) {
        hsgen_finalize(ctx);
        return st;
    }
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
    for (auto& og : ctx->ozgraphs) destroy_oz_graph(og);
    ctx->ozgraphs.clear();
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
    s->h_sqrtn.assign((size_t)d->natoms * s->nl, 0.0);
    bool fold = fold_enabled();
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
                r.sqrt_un = std::sqrt(r.dE_u_norm);                        // h_and_s.cpp:710
                if (!fold_ok(r.dE_u_norm)) fold = false;
            } else {
                r = RadDev{0, 0, 0, 0, 0, 0, 0};
            }
            s->h_sqrtn[(size_t)a * s->nl + l] = r.sqrt_un;
        }
    }
    s->kdim = off;
    s->fold = fold;
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
