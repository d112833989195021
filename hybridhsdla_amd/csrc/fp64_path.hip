// FP64 path of the hsgen C ABI: tile geometry and grouped-GEMM work lists,
// one k-point (hsgen_create_hs, hsgen_construct_ab, hsgen_contract_hs) and many
// k-points in one captured launch sequence (hsgen_create_hs_batched).
//   abgen (A/B generator) -> tmul (W_a = M_a Z_a) -> lower-triangle S and H.
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

// column padding for the split-precision path: a multiple of both sides of
// the INT8 tile (the same operand supplies rows and columns of S and H)
int split_cfg(int precision, int tile)
{
    const int es = precision == HSGEN_PREC_OZAKI_BF16 ? 2 : 1;
    if (tile > 0) {
        const IgemmLaunchCfg c = igemm_cfg(tile);
        if (c.bm > 0 && c.esize == es) return tile;
    }
    return es == 2 ? kBf16Cfg : kIgemmCfg;
}

int64_t split_col_align(int precision, int tile)
{
    const IgemmLaunchCfg c = igemm_cfg(split_cfg(precision, tile));
    return std::lcm<int64_t>(c.bm, c.bn);
}

Geometry geometry(int tile, int kdim, int ng, int64_t col_align)
{
    Geometry g;
    g.fold = false;
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
    // HSGEN_PATCH (environment, tuning) sets another patch side.
    const char* pe = std::getenv("HSGEN_PATCH");
    const int kPatch = pe && std::atoi(pe) > 0 ? std::atoi(pe) : 4;
    const int kPatchR = kPatch, kPatchC = kPatch;
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

// Common tail of create_hs / contract_hs: Y (or Z, Zs) resident -> tmul -> S, H.
hsgen_status finish_hs(hsgen_ctx* ctx, hsgen_sys* s, const Geometry& geo, int ng, const hsgen_opts& o,
                       double* H, double* S, int64_t ld, hsgen_report* rep, double ms_front)
{
    const int flags = o.flags;
    hsgen_status st;
    double2* Z = (double2*)ctx->Z.p;
    double2* Zs = zs_ptr(ctx, geo);
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
    if (is_split(o.precision)) {
        icfg = split_cfg(o.precision, o.tile);
        std::vector<OzK> ks(1);
        ks[0].ng = ng;
        ks[0].ncols = geo.gp;
        ks[0].dH = dH;
        ks[0].ldH = ldH;
        ks[0].dS = dS;
        ks[0].ldS = ldS;
        std::vector<OzResult> ozr;
        st = run_ozaki(ctx, geo, ks, flags, icfg, o.slices, o.target_rel_err, ozr, false);
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
        rep->tile_used = is_split(o.precision) ? icfg : geo.cfg;
    }
    return HSGEN_OK;
}

namespace {

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
    Geometry geo = geometry(o.tile, s->kdim, max_ng);
    geo.fold = s->fold;
    const int64_t tm = std::max(geo.t.bm, geo.t.bn);
    const int tc = tapp_cfg(geo);
    hsgen_status st = ensure_M(ctx, s, geo.t.bk, geo.t.bm);
    if (st) return st;
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
    // batches: operands (Y, W or Z, Zs, W) plus device copies of host outputs within
    // 90% of the memory the call may use (the rest of the call's device memory,
    // the graph's problem and tile lists, is a few MB)
    if (geo.fold) release(ctx->Zs);   // folded operands need no Zs
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    size_t budget = (size_t)(0.9 * (double)(fr + ctx->Z.bytes + ctx->Zs.bytes + ctx->W.bytes + ctx->btmp.bytes));
    if (const char* e = std::getenv("HSGEN_BATCH_BYTES"))   // cap (tuning and tests)
        if (std::strtoull(e, nullptr, 10) > 0) budget = std::min<size_t>(budget, std::strtoull(e, nullptr, 10));
    std::vector<std::vector<int>> batches;
    {
        size_t used = 0;
        std::vector<int> cur;
        for (int i : order) {
            const size_t gp = (size_t)roundup(std::max(kps[i].ng, 1), tm);
            const size_t need = n_operands(geo) * sizeof(double2) * (size_t)geo.ldz * gp +
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
    // persistent grids: every workgroup slot of every CU (one for the default
    // tile, whose 128 KiB of LDS fill a CU)
    const int ncu = ctx->n_cu > 0 ? ctx->n_cu : 256;
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
        if (!geo.fold) {
            st = ensure(ctx, ctx->Zs, zbytes, false);
            if (st) return st;
        }
        if (do_h) {
            st = ensure(ctx, ctx->W, zbytes, true);
            if (st) return st;
        }
        if (ntmp > 0) {
            st = ensure(ctx, ctx->btmp, sizeof(double2) * (size_t)ntmp, false);
            if (st) return st;
        }
        double2* Z = (double2*)ctx->Z.p;
        double2* Zs = zs_ptr(ctx, geo);
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
            const int grid_t = tc == geo.cfg ? ncu * zgemm_blocks_per_cu(tc) : 0;
            const int grid_c = ncu * zgemm_blocks_per_cu(geo.cfg);
            // capture the launch sequence
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (size_t q = 0; q < ks.size() && ce == hipSuccess; ++q) {
                const int i = ks[q];
                AbgenParams p = abgen_params(s, &kps[i], (const int32_t*)bg.g.p + goff[q]);
                p.ncols = (int32_t)roundup(std::max(kps[i].ng, 1), tm);
                p.Z = Z + col0[q] * geo.ldz;
                p.Zs = geo.fold ? nullptr : Zs + col0[q] * geo.ldz;
                p.ldz = geo.ldz;
                p.fold = geo.fold;
                ce = launch_abgen(p, true, ctx->stream);
            }
            const GemmProblem* dp = (const GemmProblem*)bg.probs.p;
            const TileWork* dw = (const TileWork*)bg.work.p;
            if (ce == hipSuccess && do_h) {
                ce = launch_zgemm(tc, dp, dw, (int)tw_t.size(), ctx->stream, grid_t);
                if (ce == hipSuccess)
                    ce = launch_clear_rows(W, geo.ldz, 2 * (int64_t)s->kdim, geo.ldz, (int)ncols, ctx->stream);
            }
            if (ce == hipSuccess)
                ce = launch_zgemm(geo.cfg, dp, dw + tw_t.size(), (int)tw_c.size(), ctx->stream, grid_c);
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
            r.bytes_abgen = (geo.fold ? 1.0 : 2.0) * 2.0 * 16.0 * (double)kps[i].ng * s->kdim;
            r.tile_used = geo.cfg;
        }
        if (nk > 0) rep[0].ms_total = ms_all;
    }
    return HSGEN_OK;
}

}  // namespace

}  // namespace hsgen

using namespace hsgen;

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
    if (!valid_precision(o.precision)) return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
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
    Geometry geo = geometry(o.precision == HSGEN_PREC_FP64 ? o.tile : fp64_tile_by_size(ctx, ng), s->kdim, ng,
                            o.precision == HSGEN_PREC_FP64 ? 1 : split_col_align(o.precision, o.tile));
    geo.fold = s->fold;
    st = upload_g(ctx, kp);
    if (st) return st;
    const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * geo.gp;
    st = ensure(ctx, ctx->Z, zbytes, false);
    if (st) return st;
    if (!geo.fold) {
        st = ensure(ctx, ctx->Zs, zbytes, false);
        if (st) return st;
    }
    AbgenParams p = abgen_params(s, kp, (const int32_t*)ctx->g.p);
    p.ncols = (int32_t)geo.gp;
    p.Z = (double2*)ctx->Z.p;
    p.Zs = geo.fold ? nullptr : (double2*)ctx->Zs.p;
    p.ldz = geo.ldz;
    p.fold = geo.fold;
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
        rep->bytes_abgen = (geo.fold ? 1.0 : 2.0) * 2.0 * 16.0 * (double)ng * s->kdim;   // Y, or Z and Zs
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
    if (!valid_precision(o.precision)) return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
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
        o.tile = pick_tile(ctx, o, max_ng, sys->kdim, kps, nk);
        return batched_fp64(ctx, const_cast<hsgen_sys*>(sys), nk, kps, o, H, S, ld, rep);
    }
    if (o.slices < 0 || o.slices > kMaxModuli)
        return fail(ctx, HSGEN_ERR_ARG, "slices %d outside [0, %d]", o.slices, kMaxModuli);
    HIPCHK(hipSetDevice(ctx->device));
    int max_ng = 1;
    for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
    o.tile = pick_tile(ctx, o, max_ng, sys->kdim, kps, nk);
    return batched_split(ctx, const_cast<hsgen_sys*>(sys), nk, kps, o, H, S, ld, rep);
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
    if (!valid_precision(o.precision)) return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", o.precision);
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
    // B-row scales: sqrt_u_norms themselves (folded operand Y = [A; diag(sqrt_u_norms) B],
    // as the reference's W, h_and_s_stripped_v2.cpp:64-86) or their squares (Zs)
    std::vector<double> un2((size_t)a * s.nl, 0.0);
    s.h_sqrtn.assign((size_t)a * s.nl, 0.0);
    bool fold = fold_enabled();
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
            s.h_sqrtn[(size_t)i * s.nl + l] = v;
            if (!fold_ok(v * v)) fold = false;
        }
    }
    s.kdim = off;
    s.fold = fold;
    if (G == 0) return HSGEN_OK;
    o.tile = pick_tile(ctx, o, G, s.kdim);
    Geometry geo = geometry(o.precision == HSGEN_PREC_FP64 ? o.tile : fp64_tile_by_size(ctx, G), s.kdim, G,
                            o.precision == HSGEN_PREC_FP64 ? 1 : split_col_align(o.precision, o.tile));
    geo.fold = s.fold;
    hsgen_status st = upload_M(ctx, &s, t_size, t_aa, t_ab, t_bb, s.fold ? s.h_sqrtn.data() : nullptr, geo.t.bk,
                               geo.t.bm);
    if (st) return st;
    struct Guard {
        hsgen_sys* s;
        ~Guard() { if (s->d_M) (void)hipFree(s->d_M); s->d_M = nullptr; s->d_mloc = nullptr; }
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
    HIPCHK(hipMemcpyAsync(ctx->unorm2.p, s.fold ? s.h_sqrtn.data() : un2.data(), sizeof(double) * un2.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    const size_t zbytes = sizeof(double2) * (size_t)geo.ldz * geo.gp;
    st = ensure(ctx, ctx->Z, zbytes, false);
    if (st) return st;
    if (!geo.fold) {
        st = ensure(ctx, ctx->Zs, zbytes, false);
        if (st) return st;
    }
    PackParams pp;
    pp.A = dA;
    pp.B = dB;
    pp.atoms = (const AtomDev*)ctx->atoms_tmp.p;
    pp.scale = (const double*)ctx->unorm2.p;
    pp.natoms = a;
    pp.L = L;
    pp.nl = s.nl;
    pp.ng = G;
    pp.kdim = s.kdim;
    pp.ldab = (int64_t)L * a;
    pp.Z = (double2*)ctx->Z.p;
    pp.Zs = geo.fold ? nullptr : (double2*)ctx->Zs.p;
    pp.ldz = geo.ldz;
    pp.fold = geo.fold;
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

}  // extern "C"
