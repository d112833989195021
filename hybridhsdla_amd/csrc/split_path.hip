// Split precision of the hsgen C ABI: H and S on INT8 MFMA by Ozaki scheme
// II (operands as INT8 residues modulo pairwise-coprime moduli, exact INT32
// products, FP64 recombination by the Chinese remainder theorem), with the
// modulus count chosen from an error model; one k-point or a batch of them.
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

namespace {

// ---------------------------------------------------------------------------
// Pairwise-coprime moduli <= 255 (symmetric residues fit INT8), largest first.
const int kModuli[kMaxModuli] = {255, 254, 253, 251, 247, 241, 239, 233, 229, 227,
                                 223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

struct OzPlan {
    int nmod = 0;
    int beta = 0;
    double log2P = 0.0;
    double Phi = 0.0, Plo = 0.0;
    double w1[kMaxModuli], w2[kMaxModuli], w3[kMaxModuli];   // u_m / p_m in three parts
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
        // u / p = k1 2^-40 + k2 2^-80 + (r2 / p) 2^-80, by integer long division
        const int64_t p = kModuli[m];
        const int64_t a1 = (int64_t)u << 40;                // < 2^48
        const int64_t k1 = a1 / p, r1 = a1 - k1 * p;
        const int64_t a2 = r1 << 40;
        const int64_t k2 = a2 / p, r2 = a2 - k2 * p;
        pl.w1[m] = std::ldexp((double)k1, -40);             // k1, k2 < 2^40: exact
        pl.w2[m] = std::ldexp((double)k2, -80);
        pl.w3[m] = std::ldexp((double)r2 / (double)p, -80);
    }
    return pl;
}

// Modelled relative Frobenius error of tril(C), C = L^H R, at beta.  Rounding
// an element of column i to beta bits of the column's norm exponent e_i leaves
// an independent error uniform within +-2^(e_i - beta - 1) in each real part,
// so E|dC_ij|^2 = 2^-2beta / 6 (2^(2 eL_i) ||R_j||^2 + 2^(2 eR_j) ||L_i||^2);
// the lower triangle holds about half of all (i, j).  The column sums do not
// depend on beta: they are formed once per product (as multiples of
// 2^(2 max e), which scales every partial sum by the same power of two) and
// the model is then evaluated for any beta in O(1).
struct OzModelSums {
    int el_max = 0, er_max = 0;
    double a = 0.0;   // sum_i 2^(2 (eL_i - el_max))
    double b = 0.0;   // sum_j ||R_j||^2
    double c = 0.0;   // sum_j 2^(2 (eR_j - er_max))
    double d = 0.0;   // sum_i ||L_i||^2
};

// 4^-d for d = 0 .. 537 (below 2^-1074 a power of two rounds to zero)
const double* pow4_neg()
{
    static const std::array<double, 538> t = [] {
        std::array<double, 538> v{};
        for (int d = 0; d < 538; ++d) v[d] = std::ldexp(1.0, -2 * d);
        return v;
    }();
    return t.data();
}

OzModelSums oz_model_sums(const int32_t* el, const double* sl, const int32_t* er, const double* sr, int ng)
{
    OzModelSums m;
    if (ng <= 0) return m;
    const double* p4 = pow4_neg();
    m.el_max = *std::max_element(el, el + ng);
    m.er_max = *std::max_element(er, er + ng);
    for (int i = 0; i < ng; ++i) {
        const int dl = m.el_max - el[i], dr = m.er_max - er[i];
        m.a += dl < 538 ? p4[dl] : 0.0;
        m.b += sr[i];
        m.c += dr < 538 ? p4[dr] : 0.0;
        m.d += sl[i];
    }
    return m;
}

double oz_model(const OzModelSums& m, int beta, double tril_norm2)
{
    const double a = std::ldexp(m.a, 2 * (m.el_max - beta));
    const double c = std::ldexp(m.c, 2 * (m.er_max - beta));
    return tril_norm2 > 0.0 ? std::sqrt((a * m.b + c * m.d) / 12.0 / tril_norm2) : 0.0;
}

}  // namespace

void destroy_oz_graph(hsgen_ctx::OzGraph& g)
{
    for (hipGraphExec_t ex : g.seg)
        if (ex) (void)hipGraphExecDestroy(ex);
    g.seg.clear();
    g.key.clear();
    release(g.probs);
    release(g.work);
    release(g.fprob);
}

namespace {

// Residue arrays of one modulus ("slot") for INT8 configuration ic: those of
// Z (or Y), of Zs (none when folded: S reads Y's for both operands), of W,
// and the rotated copies the kernel reads; indices into the slot, -1 absent.
struct OzSlotLayout {
    int z, zs, w, zrot, zsrot, wrot, narr;
};

OzSlotLayout oz_slots(const IgemmLaunchCfg& ic, bool fold)
{
    OzSlotLayout l;
    int n = 0;
    l.z = n++;
    l.zs = fold ? l.z : n++;
    l.w = n++;
    l.zrot = ic.needs_rot == 1 ? n++ : -1;    // (-Li, Lr) of the left operand
    l.zsrot = ic.needs_rot == 2 ? n++ : -1;   // (Ri, -Rr) of S's right operand
    l.wrot = ic.needs_rot == 2 ? n++ : -1;    // (Ri, -Rr) of W
    l.narr = n;
    return l;
}

// One split-precision evaluation of every k-point with nmod moduli: residues,
// INT8 GEMMs that store C mod p (one launch per modulus over the tiles of all
// k-points), final recombination into the outputs.  With capture the launch
// segments (per residue group: its residue pass, then its INT8 launches; then
// the recombination) are captured into hipGraphs on the first call with a
// given key (modulus count, k-points, outputs, workspace addresses) and
// replayed on later ones, with their problem and tile lists in buffers of
// their own; the HIP events between the segments time the INT8 launches as
// in the eager run.
hsgen_status oz_execute(hsgen_ctx* ctx, const Geometry& geo, const std::vector<OzK>& ks, int flags, int icfg,
                        int nmod, const int32_t* eZ, const int32_t* eZs, const int32_t* eW, double* ms_igemm,
                        bool capture)
{
    const IgemmLaunchCfg ic = igemm_cfg(icfg);
    const int64_t kd = roundup(geo.kpad, ic.bk);           // INT8 contraction depth (complex)
    const int64_t gp = geo.gp;                             // columns of all k-points
    const bool do_s = !(flags & HSGEN_SKIP_S), do_h = !(flags & HSGEN_SKIP_H);
    const OzPlan pl = oz_plan(nmod);
    const double2* Z = (const double2*)ctx->Z.p;
    const double2* Zs = zs_ptr(ctx, geo);
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
    // Residue arrays of one modulus ("slot"): Zr, Zs_r, W_r (no Zs_r when
    // folded) and rotated copies the kernel reads, 2 bytes per complex.  As many moduli as fit in 60% of
    // the free memory share one residue pass, which reads the FP64 operands
    // once for the whole group.
    const int64_t ldo = 2 * kd * ic.esize;   // bytes per column of a residue array
    if ((uint64_t)ldo * (uint64_t)(ic.bm + ic.bn) >= (1ull << 32))
        return fail(ctx, HSGEN_ERR_UNSUPPORTED, "INT8 contraction depth %lld too large for 32-bit tile offsets",
                    (long long)kd);
    const size_t arr = (size_t)ldo * gp;
    const OzSlotLayout sl = oz_slots(ic, geo.fold);
    const size_t slot_bytes = sl.narr * arr;
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
    auto slot_or_null = [&](int i, int a) -> int8_t* { return a >= 0 ? slot(i, a) : nullptr; };
    for (int i = 0; i < group; ++i) {
        int8_t* Zr = slot(i, sl.z);
        int8_t* Zsr = slot(i, sl.zs);
        int8_t* Wr = slot(i, sl.w);
        int8_t* Zrot = slot_or_null(i, sl.zrot);
        int8_t* Zsrot = slot_or_null(i, sl.zsrot);
        int8_t* Wrot = slot_or_null(i, sl.wrot);
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
                const int64_t off = 2 * ic.esize * (int64_t)c * kc;   // bytes
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
    // the cached segments of this key (capture), else the context's shared lists
    hsgen_ctx::OzGraph* cg = nullptr;
    bool fresh = false;
    if (capture) {
        std::vector<int64_t> key{icfg, nmod, group, flags, (int64_t)geo.fold, geo.ldz, kd, gp,
                                 (int64_t)(intptr_t)Z, (int64_t)(intptr_t)Zs, (int64_t)(intptr_t)W,
                                 (int64_t)(intptr_t)ctx->oz_res.p, (int64_t)(intptr_t)rout, (int64_t)(intptr_t)eZ,
                                 (int64_t)(intptr_t)eZs, (int64_t)(intptr_t)eW};
        for (const OzK& k : ks)
            key.insert(key.end(), {k.col0, (int64_t)k.ng, k.ncols, (int64_t)(intptr_t)k.dH, k.ldH,
                                   (int64_t)(intptr_t)k.dS, k.ldS});
        for (auto& e : ctx->ozgraphs)
            if (e.key == key) cg = &e;
        if (!cg) {
            if (ctx->ozgraphs.size() >= kMaxOzGraphs) {
                destroy_oz_graph(ctx->ozgraphs.front());
                ctx->ozgraphs.pop_front();
            }
            ctx->ozgraphs.emplace_back();
            cg = &ctx->ozgraphs.back();
            cg->key = key;
            fresh = true;
        }
    }
    DevBuf& b_probs = cg ? cg->probs : ctx->oz_probs;
    DevBuf& b_work = cg ? cg->work : ctx->oz_work;
    DevBuf& b_fprob = cg ? cg->fprob : ctx->oz_fprob;
    if (!cg || fresh) {
        st = ensure(ctx, b_probs, sizeof(IgemmProblem) * std::max<size_t>(probs.size(), 1), false);
        if (st) return st;
        st = ensure(ctx, b_work, sizeof(TileWork) * std::max<size_t>(work.size(), 1), false);
        if (st) return st;
        st = ensure(ctx, b_fprob, sizeof(OzFinalProblem) * std::max<size_t>(fprobs.size(), 1), false);
        if (st) return st;
        if (!probs.empty())
            HIPCHK(hipMemcpyAsync(b_probs.p, probs.data(), sizeof(IgemmProblem) * probs.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
        if (!work.empty())
            HIPCHK(hipMemcpyAsync(b_work.p, work.data(), sizeof(TileWork) * work.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
        if (!fprobs.empty())
            HIPCHK(hipMemcpyAsync(b_fprob.p, fprobs.data(), sizeof(OzFinalProblem) * fprobs.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    // run one launch segment: eagerly, or as the cached graph (captured now
    // on the key's first call)
    size_t segi = 0;
    auto segment = [&](auto&& body) -> hsgen_status {
        if (!cg) {
            HIPCHK(body());
            return HSGEN_OK;
        }
        if (fresh) {
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t ce = body();
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
            hipGraphExec_t ex = nullptr;
            hipError_t ie = hipSuccess;
            if (ce == hipSuccess && ee == hipSuccess) ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (ce != hipSuccess || ee != hipSuccess || ie != hipSuccess) {
                cg->key.clear();   // never matches again
                return fail(ctx, HSGEN_ERR_HIP, "graph capture of the split-precision launches failed: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee != hipSuccess ? ee : ie));
            }
            cg->seg.push_back(ex);
        }
        if (segi >= cg->seg.size()) return fail(ctx, HSGEN_ERR_HIP, "cached launch segment %zu missing", segi);
        HIPCHK(hipGraphLaunch(cg->seg[segi], ctx->stream));
        ++segi;
        return HSGEN_OK;
    };
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
        rp.esize = ic.esize;
        for (int i = 0; i < gn; ++i) {
            rp.p[i] = kModuli[m0 + i];
            rp.invp[i] = 1.0 / kModuli[m0 + i];
            const int c = (1 << 24) % kModuli[m0 + i];
            rp.c24[i] = c > (kModuli[m0 + i] - 1) / 2 ? c - kModuli[m0 + i] : c;
        }
        // residues of the k-points' columns of X into arrays a (and rotated arot)
        auto residues = [&](const double2* X, const int32_t* e, int a, int arot) -> hipError_t {
            for (const auto& rg : ranges) {
                const int64_t c0 = rg.first;
                rp.X = X + c0 * geo.ldz;
                rp.e = e + c0;
                for (int i = 0; i < gn; ++i) {
                    rp.out[i] = a >= 0 ? slot(i, a) + c0 * ldo : nullptr;
                    rp.out_rot[i] = arot >= 0 ? slot(i, arot) + c0 * ldo : nullptr;
                }
                const hipError_t err = launch_residues(rp, (int)(rg.second - c0), ctx->stream);
                if (err != hipSuccess) return err;
            }
            return hipSuccess;
        };
        st = segment([&]() -> hipError_t {
            rp.rot_kind = 1;
            hipError_t e = residues(Z, eZ, sl.z, sl.zrot);
            rp.rot_kind = 2;
            // folded: S's right operand is Y, whose residues the left pass
            // wrote; only a rotated copy, if the kernel reads one, is missing
            if (e == hipSuccess && do_s && (!geo.fold || sl.zsrot >= 0))
                e = residues(Zs, eZs, geo.fold ? -1 : sl.zs, sl.zsrot);
            if (e == hipSuccess && do_h) e = residues(W, eW, sl.w, sl.wrot);
            return e;
        });
        if (st) return st;
        HIPCHK(hipEventRecord(ctx->oz_ev[2 * ngroups], ctx->stream));
        st = segment([&]() -> hipError_t {
            for (int i = 0; i < gn; ++i) {
                const int m = m0 + i;
                for (int c = 0; c < nchunk; ++c) {
                    IgemmModulus md;
                    md.p = kModuli[m];
                    md.first_chunk = c == 0;
                    // INT8: 2048 MFMA k-steps <= 2080 (|acc| < 2^31); BF16: 32 k-steps of
                    // 32 x 127^2 each, exact in FP32 (|acc| < 2^24); a stage holds bk / 32
                    // (INT8) or bk / 16 (BF16) k-steps
                    md.chunk_stages = ic.esize == 1 ? 2048 / (ic.bk / 32) : 32 / (ic.bk / 16);
                    md.pad = 0;
                    md.invp = 1.0 / kModuli[m];
                    md.out = rout + (size_t)m * mod_stride;
                    const hipError_t e =
                        launch_igemm(icfg, (const IgemmProblem*)b_probs.p + (i * nchunk + c) * nprob,
                                     (const TileWork*)b_work.p, (int)work.size(), md, ctx->stream);
                    if (e != hipSuccess) return e;
                }
            }
            return hipSuccess;
        });
        if (st) return st;
        HIPCHK(hipEventRecord(ctx->oz_ev[2 * ngroups + 1], ctx->stream));
    }
    OzFinalParams fp;
    std::memset(&fp, 0, sizeof fp);
    fp.res = rout;
    fp.mod_stride = (int64_t)mod_stride;
    fp.prob = (const OzFinalProblem*)b_fprob.p;
    fp.nmod = nmod;
    fp.beta = pl.beta;
    fp.bm = ic.bm;
    fp.bn = ic.bn;
    fp.wgm = ic.wgm;
    fp.wgn = ic.wgn;
    for (int m = 0; m < nmod; ++m) {
        fp.w1[m] = pl.w1[m];
        fp.w2[m] = pl.w2[m];
        fp.w3[m] = pl.w3[m];
    }
    fp.Phi = pl.Phi;
    fp.Plo = pl.Plo;
    st = segment([&]() { return launch_crt_final(fp, (const TileWork*)b_work.p, (int)work.size(), ctx->stream); });
    if (st) return st;
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

}  // namespace


// S = Z^H Zs and H = Z^H W of every k-point from the resident FP64 operands,
// on INT8 MFMA.  With nmod_req == 0 the modulus count is chosen for `target`:
// per k-point, 64 evenly spaced columns of S and H computed in FP64 (all rows)
// give |tril C|_F for the error model before any INT8 work, and the largest
// count any k-point needs is used for all; after the run the model is
// evaluated with the exact norms and the run repeated with more moduli if a
// k-point misses the target.  The same samples measure the actual errors.
hsgen_status run_ozaki(hsgen_ctx* ctx, const Geometry& geo, const std::vector<OzK>& ks, int flags, int icfg,
                       int nmod_req, double target, std::vector<OzResult>& res, bool capture)
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
    const double2* Zs = zs_ptr(ctx, geo);
    const double2* W = (const double2*)ctx->W.p;
    HIPCHK(launch_colexp(Z, geo.ldz, (int)geo.ldz, (int)gp, eZ, ssZ, ctx->stream));
    if (do_s && geo.fold) {
        // S's right operand is Y itself: the same exponents and norms
        HIPCHK(hipMemcpyAsync(eZs, eZ, sizeof(int32_t) * gp, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ssZs, ssZ, sizeof(double) * gp, hipMemcpyDeviceToDevice, ctx->stream));
    } else if (do_s) {
        HIPCHK(launch_colexp(Zs, geo.ldz, (int)geo.ldz, (int)gp, eZs, ssZs, ctx->stream));
    }
    if (do_h) HIPCHK(launch_colexp(W, geo.ldz, (int)geo.ldz, (int)gp, eW, ssW, ctx->stream));
    // FP64 samples: columns cols[] of S and H of each k-point, all rows, ld = ng
    // (the right operand's sampled columns are gathered first: the GEMM reads
    // contiguous columns, whole tiles of bn of them)
    // The sample GEMM of one k-point has only ng / 64 tiles per product: at
    // large depth it is split into nsplit depth chunks (one per 16384 rows, at
    // most 8), so that the launch fills the chip; the chunks' products are
    // summed in chunk order afterwards.  nsplit depends on the depth only, so
    // a k-point's sample is the same in single and batched calls.
    const int nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(8, geo.kpad / 16384));
    const int64_t kchunk = roundup((geo.kpad + nsplit - 1) / nsplit, sc.bk);
    st = ensure(ctx, ctx->oz_smp, sizeof(double2) * std::max<size_t>(2 * (smpt + gatht) + 2 * smpt * (nsplit - 1), 1),
                false);
    if (st) return st;
    double2* smpS = (double2*)ctx->oz_smp.p;
    double2* smpH = smpS + smpt;
    double2* gathS = smpH + smpt;
    double2* gathH = gathS + gatht;
    double2* smp_part = gathH + gatht;             // chunks 1 .. nsplit-1 of smpS and smpH, 2 smpt each
    {
        std::vector<GemmProblem> probs;
        std::vector<TileWork> work;
        auto add = [&](const OzK& k, const double2* R, double2* C, int n) {
            for (int c = 0; c < nsplit; ++c) {
                const int64_t k0 = c * kchunk;
                if (k0 >= geo.kpad) break;
                GemmProblem p;
                p.L = Z + k.col0 * geo.ldz + k0;
                p.R = R + k0;
                p.C = c == 0 ? C : smp_part + (size_t)(c - 1) * 2 * smpt + (C - smpS);
                p.ldl = p.ldr = geo.ldz;
                p.ldc = k.ng;
                p.m = k.ng;
                p.n = n;
                p.kdepth = (int32_t)std::min<int64_t>(kchunk, geo.kpad - k0);
                p.lower = 0;
                const int ntr = (k.ng + sc.bm - 1) / sc.bm;
                for (int bi = 0; bi < ntr; ++bi) work.push_back(TileWork{(int)probs.size(), bi, 0, 0});
                probs.push_back(p);
            }
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
        if (nsplit > 1) HIPCHK(launch_sum_parts(smpS, smp_part, 2 * (int64_t)smpt, nsplit - 1, ctx->stream));
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
    // the model's column sums per k-point and product (which 0: S, R = Zs; 1: H, R = W)
    std::vector<std::array<OzModelSums, 2>> msums(nk);
    for (int k = 0; k < nk; ++k)
        for (int which = 0; which < 2; ++which) {
            const size_t o = (which == 0 ? 1 : 2) * (size_t)gp;
            const size_t c = (size_t)ks[k].col0;
            msums[k][which] =
                oz_model_sums(he.data() + c, hss.data() + c, he.data() + o + c, hss.data() + o + c, ks[k].ng);
        }
    auto model = [&](int k, int which, int beta, double tril2) { return oz_model(msums[k][which], beta, tril2); };
    auto worst_model = [&](int k, int beta, const std::array<double, 2>& tril2) {
        return std::max(do_s ? model(k, 0, beta, tril2[0]) : 0.0, do_h ? model(k, 1, beta, tril2[1]) : 0.0);
    };
    // modulus count per k-point: the smallest whose modelled error is within
    // kMargin of the target (the model is an expectation: measured errors
    // fall within 10% of it), first on the sample, then with the exact norms
    // (k-points with equal counts share their INT8 launches)
    const double kMargin = 0.75;
    // past nuseful moduli beta stays at its cap of 52 bits: more add nothing
    int nuseful = 2;
    while (nuseful < kMaxModuli && oz_beta(nuseful) < 52) ++nuseful;
    std::vector<int> nm(nk, nmod_req);
    if (nmod_req <= 0)
        for (int k = 0; k < nk; ++k) {
            int n = 2;
            while (n <= nuseful &&
                   !(oz_beta(n) >= bmin && worst_model(k, oz_beta(n), tril2_smp[k]) <= kMargin * target))
                ++n;
            if (n > nuseful)
                return fail(ctx, HSGEN_ERR_ACCURACY, "no modulus count in [2, %d] reaches rel err %.3g", nuseful,
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
            st = oz_execute(ctx, geo, sub, flags, icfg, n, eZ, eZs, eW, &ms_ig, capture);
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
            if (nmod_req > 0 || worst <= kMargin * target || nm[k] >= nuseful) continue;
            // each modulus adds log2(p) / 2 ~ 3.9 bits to beta
            const int extra = (int)std::ceil((std::log2(worst / (kMargin * target)) + 1.0) / 3.9);
            nm[k] = std::min(nuseful, nm[k] + std::max(extra, 1));
            todo[k] = 1;
            again = true;
        }
    }
    // a k-point whose modelled error still exceeds the target at the largest
    // useful modulus count: its outputs hold the best result, but the call fails
    if (nmod_req <= 0)
        for (int k = 0; k < nk; ++k)
            if (std::max(res[k].est_s, res[k].est_h) > target)
                return fail(ctx, HSGEN_ERR_ACCURACY,
                            "k-point %d: modelled rel err %.3g (H) / %.3g (S) with %d moduli exceeds the target %.3g",
                            k, res[k].est_h, res[k].est_s, res[k].nmod, target);
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
    const int icfg = split_cfg(o.precision, o.tile);
    const IgemmLaunchCfg ic = igemm_cfg(icfg);
    int max_ng = 1;
    for (int i = 0; i < nk; ++i) max_ng = std::max(max_ng, kps[i].ng);
    const int64_t align = split_col_align(o.precision, icfg);
    Geometry g0 = geometry(fp64_tile_by_size(ctx, max_ng, kps, nk), s->kdim, max_ng, align);
    g0.fold = s->fold;
    const int64_t tm = std::max<int64_t>({(int64_t)g0.t.bm, (int64_t)g0.t.bn, align});
    const int64_t kd = roundup(g0.kpad, ic.bk);
    const int narr = oz_slots(ic, g0.fold).narr;
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
    if (g0.fold) release(ctx->Zs);   // folded operands need no Zs
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
        need[i] = n_operands(g0) * sizeof(double2) * (size_t)g0.ldz * gp + sizeof(double2) * ng * ng * ((!hdev[i]) + (!sdev[i])) +
                  (size_t)(do_s + do_h) * tiles * 2 * ic.bm * ic.bn * nmod_est + (size_t)gslots * narr * 2 * ic.esize * kd * gp;
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
            p.Zs = geo.fold ? nullptr : Zs + col0[q] * geo.ldz;
            p.ldz = geo.ldz;
            p.fold = geo.fold;
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
        st = run_ozaki(ctx, geo, ks, flags, icfg, o.slices, o.target_rel_err, ozr, true);
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
            r.bytes_abgen = (g0.fold ? 1.0 : 2.0) * 2.0 * 16.0 * (double)kps[i].ng * s->kdim;
            r.tile_used = icfg;
        }
        if (nk > 0) rep[0].ms_total = ms_all;
    }
    return HSGEN_OK;
}

}  // namespace hsgen
