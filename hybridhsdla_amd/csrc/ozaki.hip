// Split-precision contraction on INT8 MFMA (Ozaki scheme II: integer
// residues modulo small pairwise-coprime moduli, exact INT32 products,
// Chinese-remainder recombination in FP64), or on BF16 MFMA with the same
// residues held as bf16 and exact FP32 accumulation.
//
// For C = L^H R (complex, depth kd):
//   1. per column j of each FP64 operand X: e_j with ||X_:j||_2 < 2^e_j,
//      X'_kj = rint(X_kj 2^(beta - e_j)), so |X'| < 2^beta and, rounding
//      included, ||X'_:j||_2 < 2^beta (1 + 2^-20) for beta >= 19 + log2(2 kd) / 2;
//   2. for each modulus p_m (<= 255): the symmetric residues X' mod p_m fit
//      INT8 (and bf16 exactly); with the (re, im) pairs interleaved,
//        Re C' = sum (Lr Rr + Li Ri),   Im C' = sum (-Li Rr + Lr Ri)
//      are each one dot product over the pairs, the second with (-Li, Lr)
//      formed in registers from L;
//   3. v_mfma_i32_16x16x64_i8 accumulates exactly in INT32, reduced mod p_m
//      every <= 2080 MFMAs (|products| <= 127^2) so it never overflows;
//      v_mfma_f32_16x16x32_bf16 accumulates exactly in FP32 while |acc| < 2^24,
//      reduced mod p_m every 32 MFMAs (32 x 32 x 127^2 + 254 < 2^24);
//   4. the epilogue stores r_m = C' mod p_m as one byte per real part, in the
//      order the MFMA fragments hold the tile (coalesced 128-byte stores);
//   5. the final pass sums r_m * (u_m / p_m) over the moduli (u_m =
//      (P/p_m)^-1 mod p_m, P the product of the moduli) with each weight split
//      into parts whose products and sums are exact, takes the centred
//      fractional part, multiplies by P and rescales by 2^(e_i + e_j - 2 beta).
// Both MFMA types see the same residues, so their H and S are bit for bit
// the same.  By Cauchy-Schwarz |C'_ij| <= ||L'_:i|| ||R'_:j|| < 2^(2 beta)
// (1 + 2^-19), so P > 2^(2 beta + 1.01) suffices whatever the depth; the host
// sets beta from the number of moduli (beta <= 52 keeps every X' an exact
// double).
#include <type_traits>

#include "hsgen_internal.h"

namespace hsgen {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(1))) const i4v* gi4ptr;

// ---------------------------------------------------------------------------
// 1. column exponents
// ---------------------------------------------------------------------------
// e_j = exponent of the column's 2-norm (||X_:j|| < 2^e_j), and the squared
// norm itself (for the error estimate)
__global__ __launch_bounds__(256) void k_colexp(const double2* __restrict__ X, int64_t ldx, int rows,
                                                int ncols, int32_t* __restrict__ e, double* __restrict__ ss)
{
    __shared__ double red2[256];
    const int col = blockIdx.x;
    if (col >= ncols) return;
    const double2* c = X + (int64_t)col * ldx;
    double q = 0.0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        double2 v = c[r];
        q = fma(v.x, v.x, fma(v.y, v.y, q));
    }
    red2[threadIdx.x] = q;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red2[threadIdx.x] += red2[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int ex = 0;
        // frexp gives 2^(ex-1) <= norm < 2^ex; the summation and the square root
        // are accurate to a few ulps, far inside the 2^-19 slack of the bound
        if (red2[0] > 0.0) (void)frexp(sqrt(red2[0]), &ex);
        e[col] = ex;
        ss[col] = red2[0];
    }
}

hipError_t launch_colexp(const double2* X, int64_t ldx, int rows, int ncols, int32_t* e, double* ss,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_colexp, dim3(ncols), dim3(256), 0, st, X, ldx, rows, ncols, e, ss);
    return hipGetLastError();
}

// Partial sums of |C_ij|^2 over the lower triangle (i >= j), one per block;
// grid (nblk), each block strides over columns.
__global__ __launch_bounds__(256) void k_tril_sumsq(const double2* __restrict__ C, int64_t ldc, int n,
                                                    double* __restrict__ part)
{
    __shared__ double red[256];
    double q = 0.0;
    for (int j = blockIdx.x; j < n; j += gridDim.x)
        for (int i = j + threadIdx.x; i < n; i += blockDim.x) {
            const double2 v = C[(int64_t)j * ldc + i];
            q = fma(v.x, v.x, fma(v.y, v.y, q));
        }
    red[threadIdx.x] = q;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

hipError_t launch_tril_sumsq(const double2* C, int64_t ldc, int n, double* part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_tril_sumsq, dim3(nblk), dim3(256), 0, st, C, ldc, n, part);
    return hipGetLastError();
}

// Statistics of a column sample Y (m x n, ld ldy) whose column j is column
// cols[j] of a square matrix X, one block per sample column:
//   part[3 j]     = sum |Y_ij|^2,
//   part[3 j + 1] = sum |X_i,cols[j] - Y_ij|^2   (0 when X is null),
//   part[3 j + 2] = |Y_cols[j],j|^2               (the diagonal element),
// over all rows i < m, or over i >= cols[j] when lower != 0.
__global__ __launch_bounds__(256) void k_block_cmp(const double2* __restrict__ X, int64_t ldx,
                                                   const double2* __restrict__ Y, int64_t ldy, int m,
                                                   const int32_t* __restrict__ cols, int lower,
                                                   double* __restrict__ part)
{
    __shared__ double ry[256], rd[256];
    const int j = blockIdx.x;
    const int cj = cols[j];
    const int i0 = lower ? cj : 0;
    double qy = 0.0, qd = 0.0;
    for (int i = i0 + threadIdx.x; i < m; i += blockDim.x) {
        const double2 y = Y[(int64_t)j * ldy + i];
        qy = fma(y.x, y.x, fma(y.y, y.y, qy));
        if (X) {
            const double2 x = X[(int64_t)cj * ldx + i];
            const double dx = x.x - y.x, dy = x.y - y.y;
            qd = fma(dx, dx, fma(dy, dy, qd));
        }
    }
    ry[threadIdx.x] = qy;
    rd[threadIdx.x] = qd;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ry[threadIdx.x] += ry[threadIdx.x + s];
            rd[threadIdx.x] += rd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * j] = ry[0];
        part[3 * j + 1] = rd[0];
        double dg = 0.0;
        if (cj < m) {
            const double2 y = Y[(int64_t)j * ldy + cj];
            dg = y.x * y.x + y.y * y.y;
        }
        part[3 * j + 2] = dg;
    }
}

// Y[:, j] = X[:, cols[j]] for j < n (rows values per column, same leading
// dimension ld for both)
__global__ __launch_bounds__(256) void k_gather_cols(const double2* __restrict__ X, int64_t ld,
                                                     const int32_t* __restrict__ cols, int rows,
                                                     double2* __restrict__ Y)
{
    const int j = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) Y[(int64_t)j * ld + r] = X[(int64_t)cols[j] * ld + r];
}

hipError_t launch_gather_cols(const double2* X, int64_t ld, const int32_t* cols, int rows, int n, double2* Y,
                              hipStream_t st)
{
    if (n <= 0 || rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_cols, dim3((rows + 255) / 256, n), dim3(256), 0, st, X, ld, cols, rows, Y);
    return hipGetLastError();
}

// acc[i] += parts[0][i], then parts[1][i], ... (in that order), n elements,
// the parts `n` apart
__global__ __launch_bounds__(256) void k_sum_parts(double2* __restrict__ acc, const double2* __restrict__ parts,
                                                   int64_t n, int nparts)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = acc[i];
        for (int c = 0; c < nparts; ++c) {
            const double2 w = parts[c * n + i];
            v.x += w.x;
            v.y += w.y;
        }
        acc[i] = v;
    }
}

hipError_t launch_sum_parts(double2* acc, const double2* parts, int64_t n, int nparts, hipStream_t st)
{
    if (n <= 0 || nparts <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)blocks), dim3(256), 0, st, acc, parts, n, nparts);
    return hipGetLastError();
}

hipError_t launch_block_cmp(const double2* X, int64_t ldx, const double2* Y, int64_t ldy, int m, int n,
                            const int32_t* cols, int lower, double* part, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_cmp, dim3(n), dim3(256), 0, st, X, ldx, Y, ldy, m, cols, lower, part);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 2. residues
// ---------------------------------------------------------------------------
// Symmetric residues.  An integer-valued double |x| < 2^53 is split once into
// x = xh 2^24 + xl (xl in [0, 2^24)); per modulus y = xh c + xl with
// c = 2^24 mod p in [-127, 127] is congruent to x and |y| < 2^36, so y / p is
// known to ~2^-23 and one rounding finds the symmetric residue y - p rint(y/p).
// Adding 2^-20 before the rounding sends the exact ties of the even modulus
// (y / p = k + 1/2) up, to -p/2 = lo; every other y / p lies at least 1/(2p)
// from a rounding boundary.  Five operations per value and modulus.
__device__ inline void split24(double x, double& xh, double& xl)
{
    xh = floor(x * 0x1p-24);
    xl = fma(-xh, 0x1p24, x);                 // exact
}

__device__ inline int sym_res(double xh, double xl, double c, double p, double invp)
{
    const double y = fma(xh, c, xl);          // exact: |y| < 2^36
    const double q = rint(fma(y, invp, 0x1p-20));
    return (int)fma(-q, p, y);                // exact
}

// bytes 0 of a, b, c, d -> one dword (a in byte 0)
__device__ inline uint32_t pack_bytes(int a, int b, int c, int d)
{
    const uint32_t ab = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0c0c0400u);
    const uint32_t cd = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x04000c0cu);
    return ab | cd;
}

// bf16 bits of an integer |r| <= 256 (exact: 8 significant bits)
__device__ inline uint32_t bf16_bits(int r)
{
    return __float_as_uint((float)r) >> 16;
}

// One lane per 4 consecutive rows of a column: 4 complex values in, per
// modulus one store of their 4 residue pairs (rows is a multiple of 4): 8
// bytes of INT8 pairs (ES = 1) or 16 bytes of bf16 pairs (ES = 2, no rotated
// copies).
template <int ES>
__global__ __launch_bounds__(256) void k_residues(OzResParams p)
{
    const int col = blockIdx.y;
    const int64_t r0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (r0 >= p.rows) return;
    const int sh = p.beta - p.e[col];
    const double2* xc = p.X + (int64_t)col * p.ldx;
    double hr[4], lr[4], hi_[4], li[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double2 v = r0 + k < p.rows_src ? xc[r0 + k] : make_double2(0.0, 0.0);
        split24(rint(ldexp(v.x, sh)), hr[k], lr[k]);
        split24(rint(ldexp(v.y, sh)), hi_[k], li[k]);
    }
    for (int m = 0; m < p.nmod; ++m) {
        const double pm = (double)p.p[m], ip = p.invp[m], c = p.c24[m];
        int rr[4], ri[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rr[k] = sym_res(hr[k], lr[k], c, pm, ip);
            ri[k] = sym_res(hi_[k], li[k], c, pm, ip);
        }
        if constexpr (ES == 2) {
            // bf16 (re, im) of row r0 + k at byte offset 4 k
            if (p.out[m])
                *(uint4*)(p.out[m] + (int64_t)col * p.ldo + 4 * r0) =
                    make_uint4(bf16_bits(rr[0]) | bf16_bits(ri[0]) << 16, bf16_bits(rr[1]) | bf16_bits(ri[1]) << 16,
                               bf16_bits(rr[2]) | bf16_bits(ri[2]) << 16, bf16_bits(rr[3]) | bf16_bits(ri[3]) << 16);
        } else {
            // bytes (re, im) of row r0 + k at byte offset 2 k
            if (p.out[m])
                *(uint2*)(p.out[m] + (int64_t)col * p.ldo + 2 * r0) =
                    make_uint2(pack_bytes(rr[0], ri[0], rr[1], ri[1]), pack_bytes(rr[2], ri[2], rr[3], ri[3]));
            if (p.out_rot[m]) {
                int a[4], b[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    a[k] = p.rot_kind == 1 ? -ri[k] : ri[k];
                    b[k] = p.rot_kind == 1 ? rr[k] : -rr[k];
                }
                *(uint2*)(p.out_rot[m] + (int64_t)col * p.ldo + 2 * r0) =
                    make_uint2(pack_bytes(a[0], b[0], a[1], b[1]), pack_bytes(a[2], b[2], a[3], b[3]));
            }
        }
    }
}

hipError_t launch_residues(const OzResParams& p, int ncols, hipStream_t st)
{
    if (p.esize != 1 && p.esize != 2) return hipErrorInvalidValue;
    if (p.rows % 4 != 0 || p.ldo % (8 * p.esize) != 0) return hipErrorInvalidValue;
    dim3 grid((p.rows / 4 + 255) / 256, ncols), block(256);
    if (p.esize == 2) {
        for (int m = 0; m < p.nmod; ++m)
            if (p.out_rot[m]) return hipErrorInvalidValue;   // bf16 kernels rotate in registers
        hipLaunchKernelGGL(k_residues<2>, grid, block, 0, st, p);
    } else {
        hipLaunchKernelGGL(k_residues<1>, grid, block, 0, st, p);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 3. INT8 MFMA GEMM with CRT epilogue
// ---------------------------------------------------------------------------
// Error-free sum of two doubles.  No multiply-add contraction may touch it
// (hipcc contracts by default, and a*b + c fused into one rounding breaks the
// error term it computes).
__device__ inline void two_sum(double a, double b, double& s, double& e)
{
#pragma clang fp contract(off)
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Position of a tile's element in its block of residue outputs: the order in
// which wave `wave` holds it (MFMA block (i, j), register q, lane lane), so
// that each store of a wave writes 128 contiguous bytes.
template <int RB, int CB>
__device__ __forceinline__ int igemm_frag_pos(int wave, int i, int j, int q, int lane)
{
    return (((wave * RB + i) * CB + j) * 4 + q) * 64 + lane;
}

// Epilogue of the INT8 kernels: C mod p of the tile as (re, im) bytes in
// [0, p), stored (first depth chunk) or added mod p to the previous chunks'.
// lane holds D[4*(lane>>4) + q][lane & 15] of MFMA block (i, j) in register q.
// Every element of the tile is written, padding included: the recombination
// reads only the lower-triangle elements inside the matrix.
template <int RB, int CB, int TILE>
__device__ __forceinline__ void store_residues(const IgemmModulus& md, int slot, int wave, int lane,
                                               const i4v (&are)[RB][CB], const i4v (&aim)[RB][CB])
{
    uchar2* out = (uchar2*)md.out + (int64_t)slot * TILE;
    const double pm = (double)md.p;
    const double invp = md.invp;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double xr = (double)are[i][j][q];
                double xi = (double)aim[i][j][q];
                int r = (int)fma(-floor(xr * invp), pm, xr);      // [0, p)
                int s = (int)fma(-floor(xi * invp), pm, xi);
                const int e = igemm_frag_pos<RB, CB>(wave, i, j, q, lane);
                if (!md.first_chunk) {
                    const uchar2 o = out[e];
                    r += o.x;
                    s += o.y;
                    r -= r >= md.p ? md.p : 0;
                    s -= s >= md.p ? md.p : 0;
                }
                out[e] = make_uchar2((unsigned char)r, (unsigned char)s);
            }
}

template <int BM, int BN, int BK, int WGM, int WGN>
struct IgemmCfg {
    static constexpr int kThreads = WGM * WGN * 64;
    static constexpr int kWM = BM / WGM;
    static constexpr int kWN = BN / WGN;
    static constexpr int kRB = kWM / 16;
    static constexpr int kCB = kWN / 16;
    static constexpr int kColBytes = 2 * BK;                  // bytes per operand column per stage
    static constexpr int kColStride = kColBytes + 16;         // + pad: conflict-free 16-lane reads
    static constexpr int kCols = 2 * BM + BN;                 // Lre, Lrot, R columns
    static constexpr int kStageBytes = kCols * kColStride;
    static constexpr int kChunks = kCols * (kColBytes / 16);  // 16-byte loads per stage
    static constexpr int kLoads = kChunks / kThreads;
    static_assert(BK % 32 == 0, "BK must be a multiple of one MFMA k-step (32 complex)");
    static_assert(kChunks % kThreads == 0, "stage must split evenly over threads");
};

template <int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(WGM * WGN * 64) void k_igemm_crt(const IgemmProblem* __restrict__ probs,
                                                              const TileWork* __restrict__ work, int nwork,
                                                              IgemmModulus md)
{
    using Cfg = IgemmCfg<BM, BN, BK, WGM, WGN>;
    constexpr int NT = Cfg::kThreads;
    constexpr int RB = Cfg::kRB, CB = Cfg::kCB;
    constexpr int CS = Cfg::kColStride;
    constexpr int NL = Cfg::kLoads;
    constexpr int CPC = Cfg::kColBytes / 16;     // 16-byte chunks per column
    extern __shared__ __attribute__((aligned(16))) char ilds[];

    const int wid = blockIdx.x;
    if (wid >= nwork) return;
    const TileWork tw = work[wid];
    const IgemmProblem pb = probs[tw.prob];
    const int I0 = tw.bi * BM;
    const int J0 = tw.bj * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WGM;
    const int wn = wave / WGM;

    gi4ptr src[NL];
    int ldsoff[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        const int e = tid + NT * r;
        const int c = e / CPC;
        const int q = e - c * CPC;
        const int8_t* p;
        if (c < BM) p = pb.Lre + (int64_t)(I0 + c) * pb.ldl + 16 * q;
        else if (c < 2 * BM) p = pb.Lrot + (int64_t)(I0 + c - BM) * pb.ldl + 16 * q;
        else p = pb.R + (int64_t)(J0 + c - 2 * BM) * pb.ldr + 16 * q;
        src[r] = (gi4ptr)(const void*)p;
        ldsoff[r] = c * CS + 16 * q;
    }

    i4v are[RB][CB], aim[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            are[i][j] = i4v{0, 0, 0, 0};
            aim[i][j] = i4v{0, 0, 0, 0};
        }

    const int nst = pb.kdepth / BK;
    const int stride_q = 2 * BK / 16;   // stage advance in 16-byte units
    i4v stage[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) stage[r] = src[r][0];
#pragma unroll
    for (int r = 0; r < NL; ++r) *(i4v*)&ilds[ldsoff[r]] = stage[r];
    __syncthreads();

    const int offA = (wm * Cfg::kWM + (lane & 15)) * CS + 16 * (lane >> 4);
    const int offAr = (BM + wm * Cfg::kWM + (lane & 15)) * CS + 16 * (lane >> 4);
    const int offB = (2 * BM + wn * Cfg::kWN + (lane & 15)) * CS + 16 * (lane >> 4);
    const double pm = (double)md.p;
    const double invp = md.invp;

    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        const int64_t qnext = (int64_t)min(s + 1, nst - 1) * stride_q;
#pragma unroll
        for (int r = 0; r < NL; ++r) stage[r] = src[r][qnext];
        __builtin_amdgcn_sched_barrier(0);
        const char* lb = ilds + buf * Cfg::kStageBytes;
#pragma unroll
        for (int ks = 0; ks < BK / 32; ++ks) {
            i4v a[RB], ar[RB], b[CB];
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                a[i] = *(const i4v*)&lb[offA + i * 16 * CS + 64 * ks];
                ar[i] = *(const i4v*)&lb[offAr + i * 16 * CS + 64 * ks];
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) b[j] = *(const i4v*)&lb[offB + j * 16 * CS + 64 * ks];
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    are[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], are[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    aim[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar[i], b[j], aim[i][j], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        char* nb = ilds + (buf ^ 1) * Cfg::kStageBytes;
#pragma unroll
        for (int r = 0; r < NL; ++r) *(i4v*)&nb[ldsoff[r]] = stage[r];
        // keep |acc| below 2^31: reduce mod p every chunk_stages stages
        if ((s + 1) % md.chunk_stages == 0 && s + 1 < nst) {
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        double x = (double)are[i][j][q];
                        are[i][j][q] = (int)fma(-floor(x * invp), pm, x);
                        x = (double)aim[i][j][q];
                        aim[i][j][q] = (int)fma(-floor(x * invp), pm, x);
                    }
        }
        __syncthreads();
    }

    store_residues<RB, CB, BM * BN>(md, tw.slot, wave, lane, are, aim);
}

// ---------------------------------------------------------------------------
// 3b. INT8 MFMA GEMM, pipelined: operands go straight from global memory into
// an NS-deep LDS ring (global_load_lds_dwordx4 writes lane l's 16 bytes at
// LDS base + 16 l), so NS-1 stages of loads are in flight while a stage is
// multiplied.  The (-Li, Lr) operand for the imaginary part is formed in
// registers from the (Lr, Li) fragment, so L is read once.
//
// LDS stage layout: column c (L columns first, then R) holds KS complex = 2 KS
// bytes = CPC chunks of 16 bytes; chunk q of column c sits at position
// q ^ f(c), f(c) = (c / (16 / CPC)) % CPC, which makes the 16-lane fragment
// reads (16 consecutive columns, one chunk) hit 16 distinct bank groups.
// ---------------------------------------------------------------------------
template <int BM, int BN, int KS, int WGM, int WGN, int NS, int ROTR, int ES = 1>
struct Igemm2Cfg {
    static constexpr int kThreads = WGM * WGN * 64;
    static constexpr int kWaves = WGM * WGN;
    static constexpr int kWM = BM / WGM;
    static constexpr int kWN = BN / WGN;
    static constexpr int kRB = kWM / 16;
    static constexpr int kCB = kWN / 16;
    static constexpr int kColBytes = 2 * KS * ES;                // KS complex residues per column
    static constexpr int kCPC = kColBytes / 16;                  // 16-byte chunks per column
    static constexpr int kKSteps = kColBytes / 64;               // MFMA k-steps (64 bytes of k) per stage
    static constexpr int kCols = BM + ((ROTR & 1) ? 2 : 1) * BN;  // L, R [, Rrot] columns
    static constexpr int kStageBytes = kCols * kColBytes;
    static constexpr int kWaveLoads = kCols * kCPC / 64;          // 1 KiB wave loads per stage
    static constexpr int kLoadsPerWave = kWaveLoads / kWaves;
    static constexpr int kLdsBytes = NS * kStageBytes;
    static_assert(KS == 32 || KS == 64, "stage depth: 32 or 64 complex");
    static_assert(ES == 1 || (ES == 2 && KS == 32 && ROTR == 0), "bf16: 32 complex per stage, rotation in registers");
    static_assert(kWaveLoads % kWaves == 0, "stage must split evenly over waves");
    static_assert(kLoadsPerWave * (NS - 2) <= 63, "vmcnt range");
};

// s_waitcnt immediate that waits for vmcnt <= N and nothing else
// (vmcnt bits [3:0] and [15:14], expcnt [6:4] = 7, lgkmcnt [11:8] = 15)
template <int N>
__device__ constexpr int vmcnt_only()
{
    return (N & 15) | (7 << 4) | (15 << 8) | (((N >> 4) & 3) << 14);
}

// ... vmcnt <= N and lgkmcnt = 0 (this wave's LDS reads have landed)
template <int N>
__device__ constexpr int vmcnt_lgkm0()
{
    return (N & 15) | (7 << 4) | (0 << 8) | (((N >> 4) & 3) << 14);
}

template <int CPC>
__device__ inline int chunk_swz(int c)
{
    return (c / (16 / CPC)) % CPC;
}

// (Lr, Li) byte pairs -> (-Li, Lr) byte pairs (mod 256 negation, no borrow
// between the two 16-bit halves)
__device__ inline int rot_pairs(int x)
{
    // bytes (Lr0, Li0, Lr1, Li1) -> (-Li0, Lr0, -Li1, Lr1): a shift puts Lr in
    // the odd bytes, then two sub-dword (SDWA) subtractions write 0 - Li into
    // the even bytes and keep the others (3 VALU operations instead of 5)
    int d = x << 8;
    const int z = 0;
    asm("v_sub_u16_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:BYTE_1"
        : "+v"(d)
        : "v"(z), "v"(x));
    asm("v_sub_u16_sdwa %0, %1, %2 dst_sel:BYTE_2 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:BYTE_3"
        : "+v"(d)
        : "v"(z), "v"(x));
    return d;
}

// (bf16 lanes: (Lr, Li) in the low and high halves of a dword) -> (-Li, Lr):
// swap the halves, then flip the sign of the new low half
__device__ inline int rot_bf16(int x)
{
    return (int)(__builtin_amdgcn_alignbit((uint32_t)x, (uint32_t)x, 16) ^ 0x8000u);
}

// ES = 1: INT8 residues on v_mfma_i32_16x16x64_i8; ES = 2: the same residues
// as bf16 on v_mfma_f32_16x16x32_bf16, FP32 accumulators (exact below 2^24,
// reduced mod p every md.chunk_stages stages).  A fragment is 16 bytes of
// one row either way (16 INT8 or 8 bf16 values of k), so the LDS layout,
// loads and fragment reads are shared; a stage holds KS complex.
template <int BM, int BN, int KS, int WGM, int WGN, int NS, int ROTR, int ES = 1>
__global__ __launch_bounds__(WGM * WGN * 64) void k_igemm2_crt(const IgemmProblem* __restrict__ probs,
                                                               const TileWork* __restrict__ work, int nwork,
                                                               IgemmModulus md)
{
    using Cfg = Igemm2Cfg<BM, BN, KS, WGM, WGN, NS, ROTR, ES>;
    using AccT = typename std::conditional<ES == 1, i4v, f4v>::type;
    constexpr int RB = Cfg::kRB, CB = Cfg::kCB;
    constexpr int CPC = Cfg::kCPC;
    constexpr int LPW = Cfg::kLoadsPerWave;
    constexpr int SB = Cfg::kStageBytes;
    extern __shared__ __attribute__((aligned(1024))) char ilds[];

    const int wid = blockIdx.x;
    if (wid >= nwork) return;
    const TileWork tw = work[wid];
    const IgemmProblem pb = probs[tw.prob];
    const int I0 = tw.bi * BM;
    const int J0 = tw.bj * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WGM;
    const int wn = wave / WGM;

    // this lane's source for each of the wave's loads of a stage (stage 0)
    // this lane's source for each of the wave's loads: a wave load fills 64
    // chunks = whole columns of one operand (the operands' column counts are
    // multiples of 64 / CPC), so its base is wave-uniform and each lane adds a
    // 32-bit offset (the host keeps (BM + BN) columns of an operand under 4 GiB)
    static_assert((BM * CPC) % 64 == 0 && ((BM + BN) * CPC) % 64 == 0, "a wave load must stay in one operand");
    const int8_t* base[LPW];
    uint32_t off[LPW];
#pragma unroll
    for (int r = 0; r < LPW; ++r) {
        const int g = wave * LPW + r;
        const int c0 = g * (64 / CPC);            // first column of this wave load (uniform)
        const int slot = g * 64 + lane;
        const int c = slot / CPC;
        const int q = (slot % CPC) ^ chunk_swz<CPC>(c);
        if (c0 < BM) {
            base[r] = pb.Lre + (int64_t)I0 * pb.ldl;
            off[r] = (uint32_t)(c * pb.ldl + 16 * q);
        } else if (c0 < BM + BN) {
            base[r] = pb.R + (int64_t)J0 * pb.ldr;
            off[r] = (uint32_t)((c - BM) * pb.ldr + 16 * q);
        } else {
            base[r] = pb.Rrot + (int64_t)J0 * pb.ldr;
            off[r] = (uint32_t)((c - BM - BN) * pb.ldr + 16 * q);
        }
    }
    typedef __attribute__((address_space(3))) char lchar;
    lchar* lbase = (lchar*)ilds;

    auto issue = [&](int stage_k, int buf) {
#pragma unroll
        for (int r = 0; r < LPW; ++r) {
            const __attribute__((address_space(1))) char* gb =
                (const __attribute__((address_space(1))) char*)(base[r] + (int64_t)stage_k * Cfg::kColBytes);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(gb + off[r]),
                                             (__attribute__((address_space(3))) void*)(lbase + buf * SB +
                                                                                       (wave * LPW + r) * 1024),
                                             16, 0, 0);
        }
    };

    AccT are[RB][CB], aim[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            are[i][j] = AccT{0, 0, 0, 0};
            aim[i][j] = AccT{0, 0, 0, 0};
        }

    const int nst = pb.kdepth / KS;
    // prologue: stages 0 .. NS-2 in flight
#pragma unroll
    for (int p = 0; p < NS - 1; ++p) issue(min(p, nst - 1), p);

    // fragment read offsets (bytes within a stage) for k-step 0; chunk of k-step ks is (lane>>4) + 4 ks
    int offA[RB], offB[CB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        const int c = wm * Cfg::kWM + i * 16 + (lane & 15);
        offA[i] = c * Cfg::kColBytes;
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int c = BM + wn * Cfg::kWN + j * 16 + (lane & 15);
        offB[j] = c * Cfg::kColBytes;
    }
    const int q0 = lane >> 4;
    const double pm = (double)md.p;
    const double invp = md.invp;

    // mod-p reduction of the accumulators after stage s when due
    auto reduce = [&](int s) {
        if (!(ROTR & 2) && (s + 1) % md.chunk_stages == 0 && s + 1 < nst) {
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if constexpr (ES == 1) {
                            double x = (double)are[i][j][q];
                            are[i][j][q] = (int)fma(-floor(x * invp), pm, x);
                            x = (double)aim[i][j][q];
                            aim[i][j][q] = (int)fma(-floor(x * invp), pm, x);
                        } else {
                            // exact integers below 2^24: into [0, p) in FP32 (the
                            // quotient may be one off; the fma itself is exact)
                            const float pf = (float)md.p, ipf = 1.0f / (float)md.p;
                            float x = are[i][j][q];
                            float r = fmaf(-floorf(x * ipf), pf, x);
                            r += r < 0.0f ? pf : 0.0f;
                            r -= r >= pf ? pf : 0.0f;
                            are[i][j][q] = r;
                            x = aim[i][j][q];
                            r = fmaf(-floorf(x * ipf), pf, x);
                            r += r < 0.0f ? pf : 0.0f;
                            r -= r >= pf ? pf : 0.0f;
                            aim[i][j][q] = r;
                        }
                    }
        }
    };
    auto read_frags = [&](int stage, int ks, i4v* A, i4v* B, i4v* BR) {
        const char* lb = ilds + (stage % NS) * SB;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int c = wm * Cfg::kWM + i * 16 + (lane & 15);
            A[i] = *(const i4v*)&lb[offA[i] + 16 * ((q0 + 4 * ks) ^ chunk_swz<CPC>(c))];
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int c = BM + wn * Cfg::kWN + j * 16 + (lane & 15);
            B[j] = *(const i4v*)&lb[offB[j] + 16 * ((q0 + 4 * ks) ^ chunk_swz<CPC>(c))];
            if (ROTR & 1) {
                const int cr = c + BN;
                BR[j] = *(const i4v*)&lb[offB[j] + BN * Cfg::kColBytes + 16 * ((q0 + 4 * ks) ^ chunk_swz<CPC>(cr))];
            }
        }
    };
    auto mfmas = [&](const i4v* a, const i4v* b, const i4v* br) {
        if constexpr (ES == 2) {
            i4v ar[RB];
#pragma unroll
            for (int i = 0; i < RB; ++i)
                ar[i] = i4v{rot_bf16(a[i][0]), rot_bf16(a[i][1]), rot_bf16(a[i][2]), rot_bf16(a[i][3])};
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    are[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a[i]),
                                                                        __builtin_bit_cast(bf8v, b[j]), are[i][j], 0,
                                                                        0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    aim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, ar[i]),
                                                                        __builtin_bit_cast(bf8v, b[j]), aim[i][j], 0,
                                                                        0, 0);
        } else {
            i4v ar[RB];
            if (!(ROTR & 1)) {
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    ar[i] = i4v{rot_pairs(a[i][0]), rot_pairs(a[i][1]), rot_pairs(a[i][2]), rot_pairs(a[i][3])};
            }
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    are[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], are[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    aim[i][j] = (ROTR & 1) ? __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], br[j], aim[i][j], 0, 0, 0)
                                           : __builtin_amdgcn_mfma_i32_16x16x64_i8(ar[i], b[j], aim[i][j], 0, 0, 0);
        }
    };

    if constexpr ((ROTR & 4) == 0) {
        for (int s = 0; s < nst; ++s) {
            // stage s landed (NS-2 younger stages may still be in flight), and
            // every wave is done with stage s-1, whose buffer is refilled next
            // (__syncthreads would wait for every LDS-DMA load in flight, i.e.
            // drain the prefetch ring; each wave instead waits for its own
            // loads of stage s and the bare barrier publishes them.  The
            // previous stage's LDS reads completed before their MFMAs issued.)
            __builtin_amdgcn_s_waitcnt(vmcnt_only<LPW * (NS - 2)>());
            asm volatile("s_barrier" ::: "memory");
            issue(min(s + NS - 1, nst - 1), (s + NS - 1) % NS);
#pragma unroll
            for (int ks = 0; ks < Cfg::kKSteps; ++ks) {
                i4v a[RB], b[CB], br[CB];
                read_frags(s, ks, a, b, br);
                mfmas(a, b, br);
            }
            reduce(s);
        }
    } else {
        // Fragments are read one k-step ahead: the reads of the next k-step
        // (across the stage boundary: after the barrier that publishes the
        // next stage) go out before this k-step's MFMAs, so the matrix pipe
        // keeps the MFMAs queued before the barrier busy while waves meet.
        constexpr int NKS = Cfg::kKSteps;
        __builtin_amdgcn_s_waitcnt(vmcnt_only<LPW * (NS - 2)>());
        asm volatile("s_barrier" ::: "memory");
        issue(min(NS - 1, nst - 1), (NS - 1) % NS);
        i4v fa[RB], fb[CB], fbr[CB];
        read_frags(0, 0, fa, fb, fbr);
        for (int s = 0; s < nst; ++s) {
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                i4v a[RB], b[CB], br[CB];
#pragma unroll
                for (int i = 0; i < RB; ++i) a[i] = fa[i];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    b[j] = fb[j];
                    br[j] = fbr[j];
                }
                if (ks + 1 < NKS) {
                    read_frags(s, ks + 1, fa, fb, fbr);
                } else if (s + 1 < nst) {
                    // stage s+1 landed (NS-2 younger stages in flight) and this
                    // wave's reads of stage s are complete; after the barrier no
                    // wave reads stage s, whose buffer takes stage s+NS
                    __builtin_amdgcn_s_waitcnt(vmcnt_lgkm0<LPW * (NS - 2)>());
                    asm volatile("s_barrier" ::: "memory");
                    issue(min(s + NS, nst - 1), (s + NS) % NS);
                    read_frags(s + 1, 0, fa, fb, fbr);
                }
                mfmas(a, b, br);
            }
            reduce(s);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    if constexpr (ES == 1) {
        store_residues<RB, CB, BM * BN>(md, tw.slot, wave, lane, are, aim);
    } else {
        i4v ire[RB][CB], iim[RB][CB];   // exact integers below 2^24
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ire[i][j][q] = (int)are[i][j][q];
                    iim[i][j][q] = (int)aim[i][j][q];
                }
        store_residues<RB, CB, BM * BN>(md, tw.slot, wave, lane, ire, iim);
    }
}

#define HSGEN_IGEMM_CFGS(X)        \
    X(1, 128, 128, 64, 2, 2)        \
    X(2, 128, 128, 32, 2, 2)        \
    X(3, 64, 64, 64, 2, 2)          \
    X(4, 128, 128, 64, 2, 4)        \
    X(5, 128, 128, 32, 2, 4)        \
    X(6, 128, 64, 64, 2, 2)

// pipelined kernels: (id, BM, BN, KS, WGM, WGN, NS, ROTR).  Every one of them
// compiles without register spills: the waits on vmcnt count the operand
// loads only, and scratch traffic in the loop would break that count.
#define HSGEN_IGEMM2_CFGS(Y)            \
    Y(11, 256, 128, 64, 4, 2, 3, 0)      \
    Y(12, 256, 128, 32, 4, 2, 4, 0)      \
    Y(13, 128, 128, 64, 2, 2, 4, 0)      \
    Y(14, 128, 128, 64, 2, 4, 4, 0)      \
    Y(15, 128, 256, 64, 2, 4, 3, 0)      \
    Y(16, 256, 128, 32, 4, 2, 5, 0)      \
    Y(21, 256, 128, 32, 4, 2, 4, 1)      \
    Y(22, 256, 128, 32, 4, 2, 3, 1)      \
    Y(23, 128, 128, 64, 2, 4, 3, 1)      \
    Y(24, 256, 128, 64, 4, 2, 2, 1)      \
    Y(25, 128, 256, 32, 2, 4, 3, 1)      \
    Y(31, 128, 256, 64, 2, 2, 3, 2)      \
    Y(32, 128, 256, 32, 2, 2, 4, 2)      \
    Y(33, 256, 128, 64, 4, 1, 3, 2)      \
    Y(34, 256, 128, 64, 4, 2, 3, 2)      \
    Y(35, 256, 128, 32, 4, 2, 4, 2)      \
    Y(41, 256, 128, 64, 4, 2, 3, 4)      \
    Y(42, 256, 128, 32, 4, 2, 4, 4)      \
    Y(43, 128, 128, 64, 2, 2, 4, 4)      \
    Y(44, 128, 128, 64, 2, 4, 4, 4)      \
    Y(45, 128, 256, 64, 2, 4, 3, 4)      \
    Y(46, 256, 128, 32, 4, 2, 5, 4)      \
    Y(47, 256, 128, 32, 4, 2, 6, 0)      \
    Y(48, 256, 128, 32, 4, 2, 6, 4)      \
    Y(49, 128, 256, 32, 2, 4, 6, 0)      \
    Y(51, 256, 192, 64, 4, 2, 2, 0)

// the same kernel on bf16 residues and BF16 MFMA: (id, BM, BN, KS, WGM, WGN, NS)
#define HSGEN_BF16_CFGS(Z)              \
    Z(61, 256, 192, 32, 4, 2, 2)         \
    Z(62, 256, 128, 32, 4, 2, 3)         \
    Z(63, 128, 256, 32, 2, 4, 3)

IgemmLaunchCfg igemm_cfg(int id)
{
#define X(ID, BM, BN, BK, WGM, WGN) \
    if (id == ID) return IgemmLaunchCfg{BM, BN, BK, WGM * WGN * 64, IgemmCfg<BM, BN, BK, WGM, WGN>::kStageBytes * 2, 1, 0, \
                                        WGM, WGN, 1};
    HSGEN_IGEMM_CFGS(X)
#undef X
#define Y(ID, BM, BN, KS, WGM, WGN, NS, ROTR) \
    if (id == ID) return IgemmLaunchCfg{BM, BN, KS, WGM * WGN * 64, Igemm2Cfg<BM, BN, KS, WGM, WGN, NS, ROTR>::kLdsBytes, \
                                        (ROTR & 1) ? 2 : 0, (ROTR & 2) ? 65536 : 0, WGM, WGN, 1};
    HSGEN_IGEMM2_CFGS(Y)
#undef Y
#define Z(ID, BM, BN, KS, WGM, WGN, NS) \
    if (id == ID) return IgemmLaunchCfg{BM, BN, KS, WGM * WGN * 64, Igemm2Cfg<BM, BN, KS, WGM, WGN, NS, 0, 2>::kLdsBytes, \
                                        0, 0, WGM, WGN, 2};
    HSGEN_BF16_CFGS(Z)
#undef Z
    return IgemmLaunchCfg{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
}

hipError_t launch_igemm(int id, const IgemmProblem* probs, const TileWork* work, int nwork,
                        const IgemmModulus& md, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
#define X(ID, BM, BN, BK, WGM, WGN)                                                            \
    if (id == ID) {                                                                            \
        constexpr int lds = IgemmCfg<BM, BN, BK, WGM, WGN>::kStageBytes * 2;                  \
        /* per launch: the attribute is per device */                                          \
        hipError_t e = hipFuncSetAttribute((const void*)k_igemm_crt<BM, BN, BK, WGM, WGN>,      \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);     \
        if (e != hipSuccess) return e;                                                         \
        hipLaunchKernelGGL((k_igemm_crt<BM, BN, BK, WGM, WGN>), dim3(nwork), dim3(WGM * WGN * 64), \
                           lds, st, probs, work, nwork, md);                                   \
        return hipGetLastError();                                                              \
    }
    HSGEN_IGEMM_CFGS(X)
#undef X
#define Y(ID, BM, BN, KS, WGM, WGN, NS, ROTR)                                                   \
    if (id == ID) {                                                                              \
        constexpr int lds = Igemm2Cfg<BM, BN, KS, WGM, WGN, NS, ROTR>::kLdsBytes;                \
        /* per launch: the attribute is per device */                                            \
        hipError_t e = hipFuncSetAttribute((const void*)k_igemm2_crt<BM, BN, KS, WGM, WGN, NS, ROTR>, \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);       \
        if (e != hipSuccess) return e;                                                           \
        hipLaunchKernelGGL((k_igemm2_crt<BM, BN, KS, WGM, WGN, NS, ROTR>), dim3(nwork),           \
                           dim3(WGM * WGN * 64), lds, st, probs, work, nwork, md);               \
        return hipGetLastError();                                                                \
    }
    HSGEN_IGEMM2_CFGS(Y)
#undef Y
#define Z(ID, BM, BN, KS, WGM, WGN, NS)                                                         \
    if (id == ID) {                                                                              \
        constexpr int lds = Igemm2Cfg<BM, BN, KS, WGM, WGN, NS, 0, 2>::kLdsBytes;                \
        hipError_t e = hipFuncSetAttribute((const void*)k_igemm2_crt<BM, BN, KS, WGM, WGN, NS, 0, 2>, \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);       \
        if (e != hipSuccess) return e;                                                           \
        hipLaunchKernelGGL((k_igemm2_crt<BM, BN, KS, WGM, WGN, NS, 0, 2>), dim3(nwork),           \
                           dim3(WGM * WGN * 64), lds, st, probs, work, nwork, md);               \
        return hipGetLastError();                                                                \
    }
    HSGEN_BF16_CFGS(Z)
#undef Z
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// 5. CRT finalisation: C = 2^(eL_i + eR_j - 2 beta) * P * centred_frac(acc)
// ---------------------------------------------------------------------------
// Recombination, one workgroup per tile of the INT8 work list: for each
// lower-triangle element inside the matrix, the residues r_m in [0, p_m) of
// all moduli give C' / P = centred_frac(sum_m r_m w_m), w_m = u_m / p_m,
// u_m = (P / p_m)^-1 mod p_m; then C = 2^(eL_i + eR_j - 2 beta) P frac.
//
// The host splits each weight as w_m = w1_m + w2_m + w3_m with w1_m a multiple
// of 2^-40 below 1, w2_m a multiple of 2^-80 below 2^-40 and w3_m < 2^-80.
// Then with r_m <= 254 and at most 20 moduli, sum r_m w1_m is a multiple of
// 2^-40 below 2^13 and sum r_m w2_m a multiple of 2^-80 below 2^-27.6: both
// fit 53 bits, so their fma chains are exact; only the third sum rounds, at
// ~2^-120 absolute.  Three fmas per residue.
//
// A lane takes 4 consecutive slot positions, which the fragment order puts in
// one row and 4 consecutive columns: per modulus one 8-byte load, all moduli
// loaded before the sums start, and the 16 rows x 16 columns that a wave
// covers are stored as 256 contiguous bytes per column.
__global__ __launch_bounds__(256) void k_crt_final(OzFinalParams p, const TileWork* __restrict__ work)
{
#pragma clang fp contract(off)
    const TileWork tw = work[blockIdx.x];
    const OzFinalProblem pb = p.prob[tw.prob];
    const int wm_n = p.bm / p.wgm, wn_n = p.bn / p.wgn;     // wave tile
    const int RB = wm_n / 16, CB = wn_n / 16;
    const int tile = p.bm * p.bn;                           // a multiple of 256
    const uint8_t* r0 = p.res + (int64_t)tw.slot * tile * 2;
    const int nmod = p.nmod;
    for (int e = 4 * threadIdx.x; e < tile; e += 4 * blockDim.x) {
        const int lane = e & 63;                            // a multiple of 4
        int t = e >> 6;
        const int q = t & 3;
        t >>= 2;
        const int j = t % CB;
        t /= CB;
        const int i = t % RB;
        const int wave = t / RB;
        const int gi = tw.bi * p.bm + (wave % p.wgm) * wm_n + i * 16 + 4 * (lane >> 4) + q;
        const int gj0 = tw.bj * p.bn + (wave / p.wgm) * wn_n + j * 16 + (lane & 15);
        // columns gj0 .. gj0 + 3 of row gi; none is kept when the first is not
        if (gi >= pb.m || gj0 >= pb.n || gi < gj0) continue;
        uint2 rv[kMaxModuli];
#pragma unroll
        for (int m = 0; m < kMaxModuli; ++m)
            rv[m] = *(const uint2*)(r0 + (int64_t)min(m, nmod - 1) * p.mod_stride + 2 * e);
        // c = 2 k + part: element gj0 + k, real (0) or imaginary (1) part
        double s1[8], s2[8], s3[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            s1[c] = 0.0;
            s2[c] = 0.0;
            s3[c] = 0.0;
        }
#pragma unroll
        for (int m = 0; m < kMaxModuli; ++m) {
            if (m < nmod) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const uint32_t w = c < 4 ? rv[m].x : rv[m].y;
                    const double r = (double)((w >> (8 * (c & 3))) & 0xffu);
                    s1[c] = fma(r, p.w1[m], s1[c]);         // exact
                    s2[c] = fma(r, p.w2[m], s2[c]);         // exact
                    s3[c] = fma(r, p.w3[m], s3[c]);
                }
            }
        }
        const int eLi = pb.eL[gi];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int gj = gj0 + k;
            if (gj >= pb.n || gi < gj) continue;
            const int sh = eLi + pb.eR[gj] - 2 * p.beta;
            double out[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int cc = 2 * k + c;
                const double f1 = s1[cc] - rint(s1[cc]);   // exact, in [-1/2, 1/2]
                double s, err0;
                two_sum(f1, s2[cc], s, err0);
                err0 += s3[cc];
                s -= rint(s);                              // exact: back into [-1/2, 1/2]
                double h, l;
                two_sum(s, err0, h, l);
                // (h + l) * (Phi + Plo)
                const double v = h * p.Phi;
                double err = fma(h, p.Phi, -v);
                err = fma(h, p.Plo, err);
                err = fma(l, p.Phi, err);
                out[c] = ldexp(v + err, sh);
            }
            if (gi == gj) out[1] = 0.0;
            pb.C[(int64_t)gj * pb.ldc + gi] = make_double2(out[0], out[1]);
        }
    }
}

hipError_t launch_crt_final(const OzFinalParams& p, const TileWork* work, int nwork, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    // 8-byte residue loads of 4 slot positions; every modulus slot read
    if (p.nmod < 1 || p.nmod > kMaxModuli || (p.bm * p.bn) % 256 != 0 || p.mod_stride % 8 != 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_crt_final, dim3(nwork), dim3(256), 0, st, p, work);
    return hipGetLastError();
}

}  // namespace hsgen
