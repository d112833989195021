// The tile autotuner (hsgen_autotune), the step after H/S generation, the
// generalized eigensolve (hsgen_eigensolve), and a diagnostic of the INT8
// matrix pipe (hsgen_int8_mfma_ceiling).
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
#include <rocsolver/rocsolver.h>

namespace hsgen {

namespace {

__device__ inline int mix32(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return (int)x;
}

typedef int i4v_t __attribute__((ext_vector_type(4)));

// Registers only: 12 independent accumulators per wave, operands that change
// from one MFMA to the next (random bytes, as residues are).  The sustained
// rate of v_mfma_i32_16x16x64_i8 with nothing else running sets the ceiling
// the split-precision GEMMs meet at the board's power limit.
constexpr int kCeilAcc = 12;
__global__ __launch_bounds__(512) void k_int8_mfma_ceiling(int iters, int* sink)
{
    const unsigned t = threadIdx.x + blockIdx.x * 512u;
    i4v_t a[6], b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a[k] = i4v_t{mix32(t * 97u + k), mix32(t * 89u + k + 7), mix32(t * 83u + k + 13), mix32(t * 79u + k + 29)};
        b[k] = i4v_t{mix32(t * 71u + k + 3), mix32(t * 67u + k + 5), mix32(t * 61u + k + 11), mix32(t * 59u + k + 17)};
    }
    i4v_t acc[kCeilAcc];
#pragma unroll
    for (int k = 0; k < kCeilAcc; ++k) acc[k] = i4v_t{k, 0, 0, 0};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kCeilAcc; ++k)
            acc[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[k % 6], b[(k + k / 6) % 6], acc[k], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < kCeilAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    if (s == 0x7fffffff) sink[0] = s;   // keeps the products live; almost never stores
}

}  // namespace

std::vector<int> tune_key(int precision, int ng, int kdim)
{
    auto b = [](int x) { return (int)std::floor(2.0 * std::log2((double)std::max(x, 1))); };
    return {precision, b(ng), b(kdim)};
}

namespace {

// S and H lower tiles of bm x bn of one k-point of ng: the tiles whose first
// row index is at least their first column index
int64_t lower_tiles(int ng, int bm, int bn)
{
    const int64_t ntr = (std::max(ng, 0) + bm - 1) / bm, ntc = (std::max(ng, 0) + bn - 1) / bn;
    int64_t n = 0;
    for (int64_t bi = 0; bi < ntr; ++bi) n += std::min<int64_t>(ntc, ((bi + 1) * bm + bn - 1) / bn);
    return 2 * n;
}

int64_t call_tiles(int ng, const hsgen_kpoint* kps, int nk, int bm, int bn)
{
    if (!kps) return lower_tiles(ng, bm, bn);
    int64_t n = 0;
    for (int i = 0; i < nk; ++i) n += lower_tiles(kps[i].ng, bm, bn);
    return n;
}

}  // namespace

int fp64_tile_by_size(const hsgen_ctx* ctx, int ng, const hsgen_kpoint* kps, int nk)
{
    const int64_t t = call_tiles(ng, kps, nk, 128, 128);
    return t < (int64_t)kSmallWaves * std::max(ctx->n_cu, 1) ? kSmallTile : 0;
}

int pick_tile(const hsgen_ctx* ctx, const hsgen_opts& o, int ng, int kdim, const hsgen_kpoint* kps, int nk)
{
    if (o.tile > 0) return o.tile;
    auto it = ctx->tuned.find(tune_key(o.precision, ng, kdim));
    if (it != ctx->tuned.end()) return it->second;
    if (o.precision == HSGEN_PREC_FP64) return fp64_tile_by_size(ctx, ng, kps, nk);
    const int64_t t = call_tiles(ng, kps, nk, 256, 192);
    if (o.precision == HSGEN_PREC_OZAKI_INT8)
        return t < kInt8FewTiles ? kInt8SmallCfg : t < kInt8SomeTiles ? kInt8MidCfg : 0;
    return t < kInt8SomeTiles ? kBf16SmallCfg : 0;
}

}  // namespace hsgen

using namespace hsgen;

extern "C" {

hsgen_status hsgen_autotune(hsgen_ctx* ctx, const hsgen_sys* sys, const hsgen_kpoint* kp, int32_t precision,
                            int32_t* best_tile, double* best_ms)
{
    if (!ctx || !sys) return fail(ctx, HSGEN_ERR_ARG, "null argument");
    hsgen_status st = check_kpoint(ctx, kp);
    if (st) return st;
    if (!valid_precision(precision)) return fail(ctx, HSGEN_ERR_ARG, "unknown precision %d", precision);
    const int ng = kp->ng;
    if (ng <= 0) return fail(ctx, HSGEN_ERR_ARG, "autotuning needs ng > 0");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<int> cands;
    if (precision == HSGEN_PREC_FP64) {
        for (int id = 1; id < 64; ++id)
            if (zgemm_cfg(id).bm > 0) cands.push_back(id);
    } else {
        const int es = precision == HSGEN_PREC_OZAKI_BF16 ? 2 : 1;   // the precision's MFMA type
        for (int id = 1; id < 64; ++id)
            if (igemm_cfg(id).bm > 0 && igemm_cfg(id).esize == es) cands.push_back(id);
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

// Diagnostic: the INT8 MFMA ceiling of this board (see k_int8_mfma_ceiling).
hsgen_status hsgen_int8_mfma_ceiling(hsgen_ctx* ctx, int32_t iters, double* tops, double* ms)
{
    if (!ctx) return HSGEN_ERR_ARG;
    if (iters <= 0) iters = 300000;
    HIPCHK(hipSetDevice(ctx->device));
    hsgen_status st = ensure(ctx, ctx->eig_d, std::max<size_t>(ctx->eig_d.bytes, sizeof(int)), false);
    if (st) return st;
    const int nblk = 8 * std::max(ctx->n_cu, 1);   // 8 workgroups of 8 waves per CU
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    hipLaunchKernelGGL(k_int8_mfma_ceiling, dim3(nblk), dim3(512), 0, ctx->stream, (int)iters, (int*)ctx->eig_d.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev[1]));
    float tm = 0.0f;
    HIPCHK(hipEventElapsedTime(&tm, ctx->ev[0], ctx->ev[1]));
    const double ops = 2.0 * nblk * 8.0 * (double)iters * kCeilAcc * 16384.0;   // 16 x 16 x 64 MACs per MFMA
    if (ms) *ms = tm;
    if (tops) *tops = tm > 0.0f ? ops / (tm * 1e-3) / 1e12 : 0.0;
    return HSGEN_OK;
}

}  // extern "C"
