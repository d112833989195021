// Internal declarations shared by the hsgen HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsgen {

constexpr int kMaxLmax = 24;            // largest supported l cutoff
constexpr int kAbgenMaxCls = 16;        // rotation / rmt classes of the class-shared generator
constexpr double kPi = 3.14159265358979323846264338327950288;

// Per-atom data of the A/B generator (AtomData fields, AtomData.h:16-52).
struct AtomDev {
    int32_t lmax;      // spherical cutoff of this atom
    int32_t nlm;       // (lmax+1)^2
    int32_t off;       // sum of nlm over preceding atoms (row offset / 2 in Z)
    int32_t cls;       // rotation class (bits 0-15) and rmt class (bits 16-31):
                       // atoms of one class have bitwise-equal rot / rmt
    double rmt;        // muffin-tin radius
    double pos[3];     // fractional position
    double rot[9];     // column-major R_atom * bmat
};

// Per-(atom, l) radial matching data (rad_match, types.h:35-43) plus the
// inverse Wronskian (AtomData.cpp:63-66).
struct RadDev {
    double u, u_dr, dE_u, dE_u_dr;
    double inv_wronk;
    double dE_u_norm;
    double sqrt_un;    // sqrt(dE_u_norm), taken on the host (h_and_s.cpp:710)
};

// One problem of the grouped complex GEMM  C = L^H R  (complex128).
//   L: kdepth x (>= m) columns, leading dimension ldl (complex units)
//   R: kdepth x (>= n) columns, leading dimension ldr
//   C: m x n, leading dimension ldc; element (i, j) at C[i + j*ldc]
// kdepth is a multiple of the kernel's BK and the operands are readable
// (zero padded) up to the tile-rounded column counts.  lower != 0: only
// tiles bi >= bj exist and only i >= j is stored, with Im(C_ii) := 0.
struct GemmProblem {
    const double2* L;
    const double2* R;
    double2* C;
    int64_t ldl, ldr, ldc;
    int32_t m, n, kdepth, lower;
};

struct TileWork {
    int32_t prob, bi, bj;
    int32_t slot;               // INT8 path: the tile's block of residue outputs; else 0
};

// Parameters of the A/B generator (abgen.hip).
struct AbgenParams {
    const AtomDev* atoms;
    const RadDev* rad;        // [natoms][nl]
    const double* ynorm;      // [L] ylm4 normalisation (incl. (-1)^m for m < 0)
    const int32_t* g;         // [ng][3]
    double kpt[3];
    double bmat[9];           // column-major
    double vol_prefac;        // 4 pi / sqrt(Omega)
    int32_t natoms, max_lmax, nl, L;
    int32_t ng, ncols;        // columns written (compact: padded; ref: ng)
    int32_t chunk;            // atoms per LDS chunk
    int32_t kdim;             // K = sum nlm (compact)
    // class-shared generator (nrot > 0): Y_lm once per rotation class and the
    // Bessel values once per rmt class, from the classes' first atoms
    int32_t nrot, nrmt;
    int32_t rot_rep[kAbgenMaxCls];
    int32_t rmt_rep[kAbgenMaxCls];
    // compact: Z = [A_a; B_a] and Zs = [A_a; N_a B_a] per atom, or, with
    // fold, only Z = [A_a; sqrt(N_a) B_a] (Zs unused)
    double2* Z;
    double2* Zs;
    int64_t ldz;
    int32_t fold;
    // ref
    double2* A;
    double2* B;
};

// Pack reference-layout A, B (L*a x G, ld = L*a) into the compact operands
// Z = [A_a; B_a] per atom and Zs = [A_a; N_a B_a], N = sqrt_u_norms^2, or,
// with fold, into Z = [A_a; sqrt_u_norms B_a] alone
// (h_and_s_stripped_v2.cpp:64-86 applies sqrt_u_norms to B before the herk).
struct PackParams {
    const double2* A;
    const double2* B;
    const AtomDev* atoms;
    const double* scale;    // [natoms][nl] B-row scale: sqrt_u_norms (fold) or their squares
    int32_t natoms, L, nl, ng, kdim;
    int64_t ldab;           // L * natoms
    double2* Z;
    double2* Zs;
    int64_t ldz;
    int32_t fold;
};

hipError_t launch_abgen(const AbgenParams& p, bool compact, hipStream_t st);
size_t abgen_lds_bytes(int chunk, int nl, int L);
size_t abgen_cls_lds_bytes(int chunk, int nl, int L, int nrot, int nrmt);
hipError_t launch_pack_ref(const PackParams& p, int ncols, hipStream_t st);

struct ZgemmLaunchCfg {
    int bm, bn, bk, threads;
};
ZgemmLaunchCfg zgemm_cfg(int id);
bool zgemm_offs32(int id);   // operand panels addressed with 32-bit offsets: BM or BN columns x ld < 4 GiB
hipError_t launch_zgemm(int id, const GemmProblem* probs, const TileWork* work, int nwork,
                        hipStream_t st, int max_grid = 0);
int zgemm_blocks_per_cu(int id);   // resident workgroups per CU (persistent grids)
hipError_t launch_clear_rows(double2* X, int64_t ld, int64_t r0, int64_t r1, int ncols, hipStream_t st);
hipError_t launch_fill_upper(double2* C, int64_t ld, int n, hipStream_t st);

// ---- split precision (ozaki.hip) ----
constexpr int kMaxModuli = 20;

struct OzResParams {
    const double2* X;           // FP64 operand, column-major
    int64_t ldx;
    const int32_t* e;           // per-column exponents
    int32_t rows;               // rows written (INT8 contraction depth)
    int32_t rows_src;           // rows present in X; rows_src..rows-1 are written as 0
    int32_t beta;
    int32_t nmod;
    int32_t p[kMaxModuli];
    double invp[kMaxModuli];
    double c24[kMaxModuli];      // 2^24 mod p, symmetric
    int8_t* out[kMaxModuli];     // (re, im) residue pairs, column stride ldo bytes, or null
    int8_t* out_rot[kMaxModuli]; // rotated pairs or null
    int32_t rot_kind;            // 1: (-im, re) (left operand), 2: (im, -re) (right operand)
    int64_t ldo;
    int32_t esize;               // bytes per real residue: 1 (INT8) or 2 (bf16)
};

struct IgemmProblem {
    const int8_t* Lre;          // (Lr, Li) residue pairs
    const int8_t* Lrot;         // (-Li, Lr) pairs (kernels that rotate L), else null
    const int8_t* R;            // (Rr, Ri) pairs
    const int8_t* Rrot;         // (Ri, -Rr) pairs (kernels that rotate R), else null
    int64_t ldl, ldr;           // bytes
    int32_t m, n, kdepth, lower;
};

// One modulus of an INT8 launch.  The kernel writes C mod p, as (re, im)
// byte pairs in [0, p), into out: tile slot s owns bytes [2 s BM BN,
// 2 (s+1) BM BN), in the order the MFMA fragments hold them (element
// (wave, i, j, q, lane) at (((wave RB + i) CB + j) 4 + q) 64 + lane, see
// igemm_frag_pos).  Launches after the first depth chunk add to what is there.
struct IgemmModulus {
    int32_t p;
    int32_t first_chunk;
    int32_t chunk_stages;
    int32_t pad;
    double invp;
    uint8_t* out;
};

// Recombination of the residue outputs of all moduli (one tile per block of
// the INT8 work list): problem prob of a tile writes C of oz_final.prob[prob].
struct OzFinalProblem {
    double2* C;
    int64_t ldc;
    const int32_t* eL;          // column exponents of the left / right operands
    const int32_t* eR;
    int32_t m, n;
};

struct IgemmLaunchCfg {
    int bm, bn, bk, threads, lds_bytes;
    int needs_rot;              // 1: reads precomputed (-Li, Lr); 2: reads precomputed (Ri, -Rr)
    int max_kchunk;             // > 0: no in-kernel mod reduction, launch at most this depth (complex)
    int wgm, wgn;               // waves along rows / columns of the tile
    int esize;                  // bytes per real residue operand: 1 (INT8 MFMA) or 2 (BF16 MFMA)
};

struct OzFinalParams {
    const uint8_t* res;         // residue outputs of modulus 0; modulus m at res + m mod_stride
    int64_t mod_stride;         // bytes
    const OzFinalProblem* prob; // [problem id of the INT8 work list], device memory
    int32_t nmod, beta;
    int32_t bm, bn, wgm, wgn;   // INT8 tile: the layout of a slot
    // u_m / p_m = w1 + w2 + w3 (u_m = (P / p_m)^-1 mod p_m): w1 a multiple of
    // 2^-40, w2 of 2^-80 below 2^-40, w3 below 2^-80
    double w1[kMaxModuli], w2[kMaxModuli], w3[kMaxModuli];
    double Phi, Plo;            // product of the moduli as double-double
};

hipError_t launch_colexp(const double2* X, int64_t ldx, int rows, int ncols, int32_t* e, double* ss,
                         hipStream_t st);
hipError_t launch_tril_sumsq(const double2* C, int64_t ldc, int n, double* part, int nblk, hipStream_t st);
hipError_t launch_block_cmp(const double2* X, int64_t ldx, const double2* Y, int64_t ldy, int m, int n,
                            const int32_t* cols, int lower, double* part, hipStream_t st);
hipError_t launch_sum_parts(double2* acc, const double2* parts, int64_t n, int nparts, hipStream_t st);
hipError_t launch_gather_cols(const double2* X, int64_t ld, const int32_t* cols, int rows, int n, double2* Y,
                              hipStream_t st);
hipError_t launch_residues(const OzResParams& p, int ncols, hipStream_t st);
IgemmLaunchCfg igemm_cfg(int id);
hipError_t launch_igemm(int id, const IgemmProblem* probs, const TileWork* work, int nwork,
                        const IgemmModulus& md, hipStream_t st);
hipError_t launch_crt_final(const OzFinalParams& p, const TileWork* work, int nwork, hipStream_t st);

}  // namespace hsgen
