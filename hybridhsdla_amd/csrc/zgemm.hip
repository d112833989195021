// Grouped complex128 GEMM  C = L^H R  on CDNA4 FP64 MFMA (v_mfma_f64_16x16x4).
//
// One launch serves a list of tiles drawn from several problems (the S and H
// contractions of one or many k-points, or the per-atom T application), so
// uneven problem sizes share one grid.  Lower-triangular problems contribute
// only tiles bi >= bj and store only i >= j, which is what the reference's
// zherk/zher2k/zgemmt calls with uplo='L' produce.
//
// Complex arithmetic on the real MFMA: with l = conj(L[k,i]) and r = R[k,j],
//   Re C += Lr.Rr + Li.Ri      Im C += Lr.Ri - Li.Rr
// i.e. four 16x16x4 real MFMAs per 16x16 complex block and k-step of 4, each
// lane feeding (Lr, Li) and (Rr, Ri) read as one 16-byte LDS word.
//
// MFMA operand layout (verified on gfx950): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, D[i][j] in lane j + 16(i % 4), register i / 4.
//
// Tiles: BM x BN complex per workgroup, WGM x WGN waves, BK complex of depth
// per LDS stage, double-buffered; global loads of stage s+1 are issued before
// the MFMAs of stage s.  LDS holds each operand column's BK complex values
// contiguously plus one complex of padding, which makes the 16-lane column
// reads bank-conflict free.
#include "hsgen_internal.h"

namespace hsgen {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));   // one complex128 (re, im)
typedef __attribute__((address_space(1))) const v2d* gptr;

template <int BM, int BN, int BK, int WGM, int WGN>
struct ZgemmCfg {
    static constexpr int kThreads = WGM * WGN * 64;
    static constexpr int kWM = BM / WGM;
    static constexpr int kWN = BN / WGN;
    static constexpr int kRB = kWM / 16;
    static constexpr int kCB = kWN / 16;
    static constexpr int kColStride = 2 * BK + 2;                 // doubles per column slot
    static constexpr int kStageDoubles = (BM + BN) * kColStride;
    static constexpr int kLoads = (BM + BN) * BK / kThreads;      // complex loads per thread per stage
    static_assert(kWM % 16 == 0 && kWN % 16 == 0, "wave tile must be a multiple of 16");
    static_assert(BK % 4 == 0, "BK must be a multiple of 4");
    static_assert((BM + BN) * BK % kThreads == 0, "stage must split evenly over threads");
};

// One output tile (work item wid); lds holds two stages.
template <int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void zgemm_tile(const GemmProblem* __restrict__ probs,
                                           const TileWork* __restrict__ work, int wid, double* lds)
{
    using Cfg = ZgemmCfg<BM, BN, BK, WGM, WGN>;
    constexpr int NT = Cfg::kThreads;
    constexpr int RB = Cfg::kRB, CB = Cfg::kCB;
    constexpr int CS = Cfg::kColStride;
    constexpr int NL = Cfg::kLoads;

    const TileWork tw = work[wid];
    const GemmProblem pb = probs[tw.prob];
    const int I0 = tw.bi * BM;
    const int J0 = tw.bj * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WGM;
    const int wn = wave / WGM;

    // per-thread global source pointers (complex units) for each of its
    // loads, in the global address space so that the loads are global_load
    // (counted by vmcnt only) rather than flat loads, which would also hold
    // up every wait on the LDS reads
    gptr src[NL];
    int ldsoff[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        const int e = tid + NT * r;
        const int c = e / BK;
        const int kk = e - c * BK;
        const double2* p = (c < BM) ? pb.L + (int64_t)(I0 + c) * pb.ldl + kk
                                    : pb.R + (int64_t)(J0 + c - BM) * pb.ldr + kk;
        src[r] = (gptr)(const void*)p;
        ldsoff[r] = c * CS + 2 * kk;
    }

    d4 cr[RB][CB], ci[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            cr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            ci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        }

    const int nst = pb.kdepth / BK;
    v2d stage[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) stage[r] = src[r][0];
#pragma unroll
    for (int r = 0; r < NL; ++r) *(v2d*)&lds[ldsoff[r]] = stage[r];
    __syncthreads();

    // lane's LDS read offsets inside a stage
    const int rowA = (wm * Cfg::kWM + (lane & 15)) * CS + 2 * (lane >> 4);
    const int rowB = (BM + wn * Cfg::kWN + (lane & 15)) * CS + 2 * (lane >> 4);

    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        // Loads of the next stage go out before this stage's MFMAs and are
        // only waited for when they are written to LDS afterwards.  The last
        // stage re-reads itself (in bounds, result unused) to keep the loop
        // free of branches.
        const int64_t knext = (int64_t)min(s + 1, nst - 1) * BK;
#pragma unroll
        for (int r = 0; r < NL; ++r) stage[r] = src[r][knext];
        // keep the scheduler from sinking the loads down to their LDS store
        __builtin_amdgcn_sched_barrier(0);
        const double* lb = lds + buf * Cfg::kStageDoubles;
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            v2d a[RB], b[CB];
#pragma unroll
            for (int i = 0; i < RB; ++i) a[i] = *(const v2d*)&lb[rowA + i * 16 * CS + 8 * ks];
#pragma unroll
            for (int j = 0; j < CB; ++j) b[j] = *(const v2d*)&lb[rowB + j * 16 * CS + 8 * ks];
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, cr[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, ci[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, cr[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const double nai = -a[i].y;
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[j].x, ci[i][j], 0, 0, 0);
            }
        }
        // ... and from hoisting the wait for them above the MFMAs
        __builtin_amdgcn_sched_barrier(0);
        double* nb = lds + (buf ^ 1) * Cfg::kStageDoubles;
#pragma unroll
        for (int r = 0; r < NL; ++r) *(v2d*)&nb[ldsoff[r]] = stage[r];
        __syncthreads();
    }

    // epilogue: lane holds D[(lane>>4) + 4 q][lane & 15] in register q
    const bool lower = pb.lower != 0;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int gj = J0 + wn * Cfg::kWN + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gi = I0 + wm * Cfg::kWM + i * 16 + (lane >> 4) + 4 * q;
                if (gi < pb.m && gj < pb.n && (!lower || gi >= gj)) {
                    double im = ci[i][j][q];
                    if (lower && gi == gj) im = 0.0;
                    pb.C[(int64_t)gj * pb.ldc + gi] = make_double2(cr[i][j][q], im);
                }
            }
        }
}

// ---------------------------------------------------------------------------
// Pipelined variant: the operands go straight from global memory into a
// double-buffered LDS ring (global_load_lds_dwordx4 writes lane l's 16 bytes
// at LDS base + 16 l), so no registers stage them, and the fragments of the
// next k-step are read before the MFMAs of the current one, across the stage
// barrier too.  LDS stage layout: column c (L columns first, then R) holds
// its BK complex values as BK chunks of 16 bytes; logical chunk q sits at
// position q ^ zswz<BK>(c), which makes the 16-lane fragment reads (16
// consecutive columns, one chunk) hit 16 distinct bank groups.  Same MFMA
// sequence per accumulator as zgemm_tile, so the same bits.
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WGM, int WGN>
struct Zgemm2Cfg {
    static constexpr int kThreads = WGM * WGN * 64;
    static constexpr int kWaves = WGM * WGN;
    static constexpr int kWM = BM / WGM;
    static constexpr int kWN = BN / WGN;
    static constexpr int kRB = kWM / 16;
    static constexpr int kCB = kWN / 16;
    static constexpr int kColBytes = 16 * BK;
    static constexpr int kStageBytes = (BM + BN) * kColBytes;
    static constexpr int kWaveLoads = (BM + BN) * BK / 64;       // 1 KiB wave loads per stage
    static constexpr int kLoadsPerWave = kWaveLoads / kWaves;
    static constexpr int kLdsBytes = 2 * kStageBytes;
    static_assert(BK == 16 || BK == 8, "the swizzle spreads 16 consecutive columns over 16 bank groups");
    static_assert(kWM % 16 == 0 && kWN % 16 == 0, "wave tile must be a multiple of 16");
    static_assert(kWaveLoads % kWaves == 0, "stage must split evenly over waves");
    static_assert(BM % (64 / BK) == 0, "a wave load must stay in one operand");
};

// Chunk swizzle of column c: with 16 chunks per column (256 bytes) c % 16
// picks the chunk; with 8 (128 bytes, two columns per 256-byte bank row) the
// column's parity picks the half row and (c / 2) % 8 the chunk.
template <int BK>
__device__ __forceinline__ int zswz(int c)
{
    return BK == 16 ? (c & 15) : ((c >> 1) & 7);
}

// s_waitcnt immediates: vmcnt = 0 and nothing else; vmcnt = 0 and lgkmcnt = 0
// (vmcnt bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8])
constexpr int kVmcnt0 = (0 & 15) | (7 << 4) | (15 << 8);
constexpr int kVmcnt0Lgkm0 = (0 & 15) | (7 << 4) | (0 << 8);

template <int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void zgemm2_tile(const GemmProblem* __restrict__ probs,
                                            const TileWork* __restrict__ work, int wid, char* lds)
{
    using Cfg = Zgemm2Cfg<BM, BN, BK, WGM, WGN>;
    constexpr int RB = Cfg::kRB, CB = Cfg::kCB;
    constexpr int LPW = Cfg::kLoadsPerWave;
    constexpr int SB = Cfg::kStageBytes;
    constexpr int CBY = Cfg::kColBytes;

    const TileWork tw = work[wid];
    const GemmProblem pb = probs[tw.prob];
    const int I0 = tw.bi * BM;
    const int J0 = tw.bj * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WGM;
    const int wn = wave / WGM;

    // this lane's source for each of the wave's loads of a stage: a wave load
    // fills 64 chunks = 64 / BK whole columns of one operand, so its base is
    // wave-uniform and each lane adds a 32-bit byte offset (a tile's column
    // panel of an operand spans (BM or BN) x ld complex values, far below 4 GiB
    // at every size the library accepts)
    const char* base[LPW];
    uint32_t off[LPW];
#pragma unroll
    for (int r = 0; r < LPW; ++r) {
        const int g = wave * LPW + r;
        const int c0 = g * (64 / BK);
        const int slot = g * 64 + lane;
        const int c = slot / BK;
        const int q = (slot % BK) ^ zswz<BK>(c);
        if (c0 < BM) {
            base[r] = (const char*)(pb.L + (int64_t)I0 * pb.ldl);
            off[r] = (uint32_t)(((int64_t)c * pb.ldl + q) * 16);
        } else {
            base[r] = (const char*)(pb.R + (int64_t)J0 * pb.ldr);
            off[r] = (uint32_t)(((int64_t)(c - BM) * pb.ldr + q) * 16);
        }
    }
    typedef __attribute__((address_space(3))) char lchar;
    lchar* lbase = (lchar*)lds;
    auto issue = [&](int stage_k, int buf) {
#pragma unroll
        for (int r = 0; r < LPW; ++r) {
            const __attribute__((address_space(1))) char* gb =
                (const __attribute__((address_space(1))) char*)(base[r] + (int64_t)stage_k * (16 * BK));
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(gb + off[r]),
                                             (__attribute__((address_space(3))) void*)(lbase + buf * SB +
                                                                                       (wave * LPW + r) * 1024),
                                             16, 0, 0);
        }
    };

    d4 cr[RB][CB], ci[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            cr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            ci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        }

    // fragment read positions: column and lane's chunk within the k-step
    int colA[RB], colB[CB];
#pragma unroll
    for (int i = 0; i < RB; ++i) colA[i] = wm * Cfg::kWM + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < CB; ++j) colB[j] = BM + wn * Cfg::kWN + j * 16 + (lane & 15);
    const int q0 = lane >> 4;
    auto read_frags = [&](int buf, int ks, v2d* a, v2d* b) {
        const char* lb = lds + buf * SB;
#pragma unroll
        for (int i = 0; i < RB; ++i)
            a[i] = *(const v2d*)&lb[colA[i] * CBY + 16 * ((q0 + 4 * ks) ^ zswz<BK>(colA[i]))];
#pragma unroll
        for (int j = 0; j < CB; ++j)
            b[j] = *(const v2d*)&lb[colB[j] * CBY + 16 * ((q0 + 4 * ks) ^ zswz<BK>(colB[j]))];
    };
    auto mfmas = [&](const v2d* a, const v2d* b) {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j)
                cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j)
                ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, ci[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j)
                cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const double nai = -a[i].y;
#pragma unroll
            for (int j = 0; j < CB; ++j)
                ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[j].x, ci[i][j], 0, 0, 0);
        }
    };

    const int nst = pb.kdepth / BK;
    constexpr int NKS = BK / 4;
    if (nst > 0) {
        // stage 0 in, published; stage 1 in flight
        issue(0, 0);
        __builtin_amdgcn_s_waitcnt(kVmcnt0);
        asm volatile("s_barrier" ::: "memory");
        issue(min(1, nst - 1), 1);
        v2d fa[RB], fb[CB];
        read_frags(0, 0, fa, fb);
        for (int s = 0; s < nst; ++s) {
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                v2d a[RB], b[CB];
#pragma unroll
                for (int i = 0; i < RB; ++i) a[i] = fa[i];
#pragma unroll
                for (int j = 0; j < CB; ++j) b[j] = fb[j];
                if (ks + 1 < NKS) {
                    read_frags(s & 1, ks + 1, fa, fb);
                } else if (s + 1 < nst) {
                    // stage s+1 landed (this wave's loads; the barrier publishes
                    // everyone's) and this wave's reads of stage s are complete;
                    // after the barrier no wave reads stage s, whose buffer takes
                    // stage s+2 (the last stage again at the end: in bounds, unused)
                    __builtin_amdgcn_s_waitcnt(kVmcnt0Lgkm0);
                    asm volatile("s_barrier" ::: "memory");
                    issue(min(s + 2, nst - 1), s & 1);
                    read_frags((s + 1) & 1, 0, fa, fb);
                }
                mfmas(a, b);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
    }   // a zero-depth problem stores zeros

    // epilogue: lane holds D[(lane>>4) + 4 q][lane & 15] in register q
    const bool lower = pb.lower != 0;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int gj = J0 + wn * Cfg::kWN + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gi = I0 + wm * Cfg::kWM + i * 16 + (lane >> 4) + 4 * q;
                if (gi < pb.m && gj < pb.n && (!lower || gi >= gj)) {
                    double im = ci[i][j][q];
                    if (lower && gi == gj) im = 0.0;
                    pb.C[(int64_t)gj * pb.ldc + gi] = make_double2(cr[i][j][q], im);
                }
            }
        }
}

template <int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM * WGN * 64) void k_zgemm2_lh(const GemmProblem* __restrict__ probs,
                                                              const TileWork* __restrict__ work, int nwork)
{
    extern __shared__ __attribute__((aligned(1024))) char zlds[];
    for (int wid = blockIdx.x; wid < nwork; wid += gridDim.x) {
        zgemm2_tile<BM, BN, BK, WGM, WGN>(probs, work, wid, zlds);
        __syncthreads();   // the next tile's first stage overwrites LDS
    }
}

// Workgroups take tiles wid = blockIdx.x, blockIdx.x + gridDim.x, ...: with
// one workgroup per tile (grid = nwork) that is a single tile; with a grid of
// one workgroup per CU slot the launch is persistent and every workgroup
// walks the list, keeping the dispatch order (and so the XCD a tile runs on)
// of the one-per-tile launch.
template <int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM * WGN * 64) void k_zgemm_lh(const GemmProblem* __restrict__ probs,
                                                             const TileWork* __restrict__ work,
                                                             int nwork)
{
    using Cfg = ZgemmCfg<BM, BN, BK, WGM, WGN>;
    __shared__ double lds[2 * Cfg::kStageDoubles];
    for (int wid = blockIdx.x; wid < nwork; wid += gridDim.x) {
        zgemm_tile<BM, BN, BK, WGM, WGN>(probs, work, wid, lds);
        __syncthreads();   // the next tile's first stage overwrites LDS
    }
}

// Tile configurations (id -> template instance).

#define HSGEN_ZGEMM_CFGS(X)        \
    X(1, 128, 128, 8, 2, 4)         \
    X(2, 128, 128, 16, 2, 4)        \
    X(3, 64, 64, 8, 2, 2)           \
    X(4, 128, 64, 8, 2, 2)          \
    X(5, 128, 128, 8, 2, 2)         \
    X(6, 128, 128, 8, 4, 2)         \
    X(7, 128, 128, 16, 2, 2)        \
    X(8, 128, 128, 16, 4, 2)

// pipelined configurations (id -> k_zgemm2_lh instance)
#define HSGEN_ZGEMM2_CFGS(Y)       \
    Y(9, 128, 128, 16, 4, 2)        \
    Y(10, 128, 128, 16, 2, 4)       \
    Y(11, 128, 128, 16, 4, 4)       \
    Y(12, 64, 64, 16, 2, 2)         \
    Y(13, 64, 128, 8, 2, 4)         \
    Y(14, 128, 64, 8, 4, 2)

// the pipelined kernels address a tile's operand panel with 32-bit offsets
bool zgemm_offs32(int id)
{
#define Y(ID, BM, BN, BK, WGM, WGN) \
    if (id == ID) return true;
    HSGEN_ZGEMM2_CFGS(Y)
#undef Y
    return false;
}

ZgemmLaunchCfg zgemm_cfg(int id)
{
#define X(ID, BM, BN, BK, WGM, WGN) \
    if (id == ID) return ZgemmLaunchCfg{BM, BN, BK, WGM * WGN * 64};
    HSGEN_ZGEMM_CFGS(X)
    HSGEN_ZGEMM2_CFGS(X)
#undef X
    return ZgemmLaunchCfg{0, 0, 0, 0};
}

// Workgroups of configuration id that one CU holds at once (registers, LDS
// and waves permitting): a persistent launch sizes its grid by it.
int zgemm_blocks_per_cu(int id)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define X(ID, BM, BN, BK, WGM, WGN)                                                                   \
    if (id == ID)                                                                                     \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)k_zgemm_lh<BM, BN, BK, WGM, WGN>, \
                                                         WGM * WGN * 64, 0);
    HSGEN_ZGEMM_CFGS(X)
#undef X
#define Y(ID, BM, BN, BK, WGM, WGN)                                                                     \
    if (id == ID)                                                                                       \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)k_zgemm2_lh<BM, BN, BK, WGM, WGN>, \
                                                         WGM * WGN * 64, Zgemm2Cfg<BM, BN, BK, WGM, WGN>::kLdsBytes);
    HSGEN_ZGEMM2_CFGS(Y)
#undef Y
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

hipError_t launch_zgemm(int id, const GemmProblem* probs, const TileWork* work, int nwork,
                        hipStream_t st, int max_grid)
{
    if (nwork <= 0) return hipSuccess;
    const int grid = max_grid > 0 ? std::min(nwork, max_grid) : nwork;
#define X(ID, BM, BN, BK, WGM, WGN)                                                      \
    if (id == ID) {                                                                      \
        hipLaunchKernelGGL((k_zgemm_lh<BM, BN, BK, WGM, WGN>), dim3(grid),              \
                           dim3(WGM * WGN * 64), 0, st, probs, work, nwork);             \
        return hipGetLastError();                                                        \
    }
    HSGEN_ZGEMM_CFGS(X)
#undef X
#define Y(ID, BM, BN, BK, WGM, WGN)                                                                   \
    if (id == ID) {                                                                                   \
        constexpr int lds = Zgemm2Cfg<BM, BN, BK, WGM, WGN>::kLdsBytes;                               \
        /* per launch: the attribute is per device */                                                 \
        hipError_t e = hipFuncSetAttribute((const void*)k_zgemm2_lh<BM, BN, BK, WGM, WGN>,             \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);            \
        if (e != hipSuccess) return e;                                                                \
        hipLaunchKernelGGL((k_zgemm2_lh<BM, BN, BK, WGM, WGN>), dim3(grid), dim3(WGM * WGN * 64), lds, \
                           st, probs, work, nwork);                                                   \
        return hipGetLastError();                                                                     \
    }
    HSGEN_ZGEMM2_CFGS(Y)
#undef Y
    return hipErrorInvalidValue;
}

// Mirror the lower triangle into the upper one: C[j + i*ld] = conj(C[i + j*ld]).
__global__ void k_fill_upper(double2* C, int64_t ld, int n)
{
    const int j = blockIdx.y * 16 + threadIdx.y;
    const int i = blockIdx.x * 16 + threadIdx.x;
    if (i < n && j < n && i < j) {
        double2 v = C[(int64_t)i * ld + j];
        C[(int64_t)j * ld + i] = make_double2(v.x, -v.y);
    }
}

// Zero rows r0 .. r1-1 of ncols columns (leading dimension ld); a kernel
// rather than a memset so that it can sit in a captured graph.
__global__ void k_clear_rows(double2* X, int64_t ld, int64_t r0, int64_t r1)
{
    const int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < r1) X[(int64_t)blockIdx.y * ld + r] = make_double2(0.0, 0.0);
}

hipError_t launch_clear_rows(double2* X, int64_t ld, int64_t r0, int64_t r1, int ncols, hipStream_t st)
{
    if (r1 <= r0 || ncols <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_rows, dim3((unsigned)((r1 - r0 + 255) / 256), ncols), dim3(256), 0, st, X, ld,
                       r0, r1);
    return hipGetLastError();
}

hipError_t launch_fill_upper(double2* C, int64_t ld, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    dim3 grid((n + 15) / 16, (n + 15) / 16), block(16, 16);
    hipLaunchKernelGGL(k_fill_upper, grid, block, 0, st, C, ld, n);
    return hipGetLastError();
}

}  // namespace hsgen
