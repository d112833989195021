/*
 * hsgen - MI355X-native FLAPW muffin-tin Hamiltonian (H) and overlap (S)
 * generation.  C ABI: plain pointers and sizes, no C++ or torch types.
 *
 * Complex numbers are complex128 stored as interleaved (re, im) doubles; a
 * "complex pointer" below is a double* to 2*n doubles.  Matrices are
 * column-major.  Output H and S hold the LOWER triangle (i >= j, diagonal
 * included) of the G x G Hermitian matrices; the strict upper triangle is left
 * untouched unless HSGEN_FILL_UPPER is set, exactly like the reference, which
 * calls zherk/zher2k/zgemmt with uplo = 'L' and never writes the upper part.
 *
 * Pointers to bulk arrays (A, B, H, S) may be host or device (HIP) memory;
 * the library inspects them and copies as needed.  Everything else
 * (system description, T-matrices, k-point lists) is host memory.
 *
 * Every entry point returns an hsgen_status; on failure no output is
 * guaranteed and hsgen_last_error() describes the cause.  Nothing calls
 * exit() or abort() (the reference does, see h_and_s.cpp:712-723).
 * A context binds one GPU and one HIP stream; calls on one context must not
 * overlap, separate contexts are independent.
 */
#ifndef HSGEN_H
#define HSGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HSGEN_OK = 0,
    HSGEN_ERR_ARG = -1,         /* invalid argument (shape, null, range) */
    HSGEN_ERR_HIP = -2,         /* HIP runtime error */
    HSGEN_ERR_NOMEM = -3,       /* device or host allocation failed */
    HSGEN_ERR_NODEV = -4,       /* no usable gfx950 device */
    HSGEN_ERR_UNSUPPORTED = -5, /* valid request this build cannot serve */
    HSGEN_ERR_ACCURACY = -6     /* split precision cannot reach the target */
} hsgen_status;

/* Arithmetic of the contraction (A/B generation, T application and the
 * final recombination are always FP64). */
typedef enum {
    HSGEN_PREC_FP64 = 0,      /* FP64 MFMA (v_mfma_f64_16x16x4) */
    HSGEN_PREC_OZAKI_INT8 = 1, /* split precision: INT8 slices on INT8 MFMA,
                                 exact INT32 accumulation, FP64 recombination;
                                 HSGEN_ERR_UNSUPPORTED if 2 (2K) bytes x 448 columns
                                 of an operand would exceed 4 GiB (K > 2.3e6) */
    HSGEN_PREC_OZAKI_BF16 = 2 /* the same slices (residues) held as bf16 on
                                 BF16 MFMA, exact FP32 accumulation: the same H
                                 and S bit for bit as HSGEN_PREC_OZAKI_INT8, at
                                 half its nominal MFMA rate and twice its slice
                                 memory (K > 1.1e6 unsupported) */
} hsgen_precision;

enum {
    HSGEN_FILL_UPPER = 1,     /* also write the upper triangle (Hermitian mirror) */
    HSGEN_SKIP_H = 2,         /* compute S only */
    HSGEN_SKIP_S = 4          /* compute H only */
};

typedef struct hsgen_ctx hsgen_ctx;   /* one GPU, one stream */
typedef struct hsgen_sys hsgen_sys;   /* device-resident system description */

/* Per-atom system description, the state of the reference's AtomData
 * (AtomData.h:16-52) after generate_matchdata() and augment_tmats():
 * everything the A/B generator and the contraction read. */
typedef struct {
    int32_t natoms;
    int32_t max_lmax;          /* max over lmax[]; sets the padded L = (max_lmax+1)^2 */
    const int32_t* lmax;       /* [natoms] spherical l cutoff (AtomData::get_lmax().first) */
    const double* rmt;         /* [natoms] muffin-tin radius (rad_mesh::rmt) */
    const double* pos;         /* [natoms][3] position in lattice (fractional) coordinates */
    const double* rotmat;      /* [natoms][9] column-major R_atom * bmat (AtomData::get_rotmat) */
    const double* match;       /* [natoms][max_lmax+1][5] = {u, u_dr, dE_u, dE_u_dr, dE_u_norm}
                                  (types.h:35-43); entries l > lmax[atom] are ignored */
    double omega;              /* unit-cell volume det(amat) (AtomData::vol_cell) */
    const double* bmat;        /* [9] column-major reciprocal basis incl. 2*pi */
    /* Final T-matrices of atom a (AtomData::t_matrices after augment_tmats):
     * n_a x n_a complex column-major, n_a = (lmax[a]+1)^2.  T_aa and T_bb are
     * Hermitian and only their lower triangles are read (the reference applies
     * them with zhemm uplo='L'); T_ab is read in full and T_ba = T_ab^H.
     * Atoms may share pointers (same type); the upload is deduplicated. */
    const double* const* t_aa;
    const double* const* t_ab;
    const double* const* t_bb;
} hsgen_system;

/* One k-point and its plane-wave set (k_mesh, types.h:14-21). */
typedef struct {
    double k[3];               /* reduced coordinates */
    int32_t ng;                /* number of G-vectors = matrix dimension */
    const int32_t* g;          /* [ng][3] integer reciprocal-lattice vectors, caller order kept */
} hsgen_kpoint;

typedef struct {
    int32_t precision;         /* hsgen_precision */
    int32_t slices;            /* split precision: residue slices per operand, i.e. the
                                  number of moduli (0 = chosen for target_rel_err from an error
                                  model and an FP64 sample of columns of H and S, then checked
                                  against the model with the exact norms, with a retry) */
    double target_rel_err;     /* split precision: target relative Frobenius error (0 = 1e-12) */
    int32_t flags;             /* HSGEN_FILL_UPPER | HSGEN_SKIP_H | HSGEN_SKIP_S */
    int32_t tile;              /* contraction tile config id (0 = the one hsgen_autotune chose
                                  for this size, else the built-in default for the size: smaller
                                  tiles for calls too small to fill the GPU with the large ones) */
} hsgen_opts;

typedef struct {
    double ms_total;           /* GPU time of the whole call (HIP events) */
    double ms_abgen;           /* A/B generation */
    double ms_tmul;            /* per-atom T application */
    double ms_slice;           /* split precision: slicing */
    double ms_contract;        /* H and S contraction */
    double flops_alg;          /* 16 G^2 K + 32 G sum L_a^2 (K = sum L_a) */
    double bytes_abgen;        /* algorithmic bytes written by A/B generation */
    int32_t ng;                /* G */
    int32_t kdim;              /* K = sum_a L_a */
    int32_t slices_used;       /* split precision: slices actually used */
    int32_t tile_used;         /* contraction tile config id actually used */
    double err_est_h;          /* split precision: estimated relative Frobenius error of H */
    double err_est_s;          /* ... and of S (0 in FP64) */
    double ms_igemm;           /* split precision: GPU time of the residue (INT8 or BF16) GEMM launches alone
                                  (last modulus set when an automatic retry ran) */
    double err_smp_h;          /* split precision: measured relative Frobenius error of H on
                                  64 evenly spaced columns (their lower parts) recomputed in FP64 */
    double err_smp_s;          /* ... and of S */
} hsgen_report;

/* Context: bind GPU `device` (HIP ordinal), create its stream.
 * Replaces dev_init()/dev_finalize() (flapw_opt/src/include/acc/wrappers.h:6-7). */
hsgen_status hsgen_init(hsgen_ctx** ctx, int32_t device);
hsgen_status hsgen_finalize(hsgen_ctx* ctx);
const char* hsgen_last_error(const hsgen_ctx* ctx);   /* never NULL */
const char* hsgen_status_string(hsgen_status s);      /* never NULL */
/* Frees cached device workspace (grown on demand by the calls below). */
hsgen_status hsgen_release_workspace(hsgen_ctx* ctx);

/* Upload a system once; reuse it for every k-point.  The host arrays in
 * `desc` are copied and may be freed after the call. */
hsgen_status hsgen_system_create(hsgen_ctx* ctx, const hsgen_system* desc,
                                 hsgen_sys** sys);
hsgen_status hsgen_system_destroy(hsgen_ctx* ctx, hsgen_sys* sys);

/* A/B matching coefficients for one k-point in the reference layout.
 * Replaces construct_AB (flapw_opt/src/include/core.h:15-16,
 * tensor_AB_match.cpp:63-129): A and B are L x natoms x ng complex with
 * lm fastest (ld = L*natoms, L = (max_lmax+1)^2); rows l > lmax[atom] are 0. */
hsgen_status hsgen_construct_ab(hsgen_ctx* ctx, const hsgen_sys* sys,
                                const hsgen_kpoint* kp, double* A, double* B);

/* Full H and S for one k-point: A/B generation, T application and the
 * contraction on the GPU.  Replaces create_HS2
 * (flapw_opt/src/h_and_s.cpp:591-598; STRIPPED branch :590-894).
 * H and S: ng x ng complex, column-major, leading dimension ld >= ng. */
hsgen_status hsgen_create_hs(hsgen_ctx* ctx, const hsgen_sys* sys,
                             const hsgen_kpoint* kp, const hsgen_opts* opts,
                             double* H, double* S, int64_t ld,
                             hsgen_report* report);

/* Many k-points of one system in one grouped, graph-captured launch
 * sequence (uneven ng per k).  Replaces the per-k loop of
 * generate_matrices (flapw_opt/src/h_and_s.cpp:914-931).
 * H[i], S[i]: kps[i].ng x kps[i].ng with leading dimension ld[i].
 * report may be NULL or point to nk reports (per-k figures are shares of
 * the single batched launch; ms_total is the whole batch in report[0]).
 * Split precision: the k-points of a memory batch share their INT8 launches
 * (one per modulus over the tiles of all k-points that use that many moduli);
 * each k-point gets the modulus count of its own error model, and its H, S
 * and report fields equal those of a hsgen_create_hs call, except ms_igemm:
 * the k-point's share (by ng^2) of the INT8 launches it shared. */
hsgen_status hsgen_create_hs_batched(hsgen_ctx* ctx, const hsgen_sys* sys,
                                     int32_t nk, const hsgen_kpoint* kps,
                                     const hsgen_opts* opts, double* const* H,
                                     double* const* S, const int64_t* ld,
                                     hsgen_report* report);

/* Autotuner: time every tile configuration of the contraction (FP64 GEMM,
 * or the INT8 GEMM of split precision at a fixed 10 moduli) on this
 * k-point and remember the fastest for problems of its size (half-octave
 * buckets of G and K).  Later calls with opts.tile = 0 use the remembered
 * configuration; explicit opts.tile values still win.  No reference
 * counterpart: it takes the place of the reference's per-machine GPU/CPU
 * split-ratio tuning (compute_ratio.cpp, wrappers.cpp:73-109).  *best_tile
 * and *best_ms (contraction GPU time of the winner) may be NULL. */
hsgen_status hsgen_autotune(hsgen_ctx* ctx, const hsgen_sys* sys, const hsgen_kpoint* kp,
                            int32_t precision, int32_t* best_tile, double* best_ms);

/* Contraction only, from precomputed A and B.  Replaces
 * create_HS2_stripped_v2 (flapw_opt/src/include/acc/h_and_s_stripped.h:24-34,
 * h_and_s_stripped_v2.cpp:100-108) with the same argument meaning:
 *   A, B        L*a x G complex, ld = L*a, atom at row offset atom*L
 *               (rows beyond t_size[atom] are not read)
 *   t_size      [a] T block size of each atom (myTMat::aa_size = ab_size = bb_size)
 *   t_aa/ab/bb  [a] pointers to t_size x t_size complex blocks (as myTMat)
 *   l_maxs, max_lmax, sqrt_u_norms   as the reference: sqrt_u_norms[atom*(max_lmax+1)+l]
 *               is sqrt(||u_dot_l||^2) of the atom
 *   H, S        G x G complex, leading dimension ld (lower triangle written)
 * Unlike the reference, A and B are const (not destroyed) and no W workspace
 * is needed. */
hsgen_status hsgen_contract_hs(hsgen_ctx* ctx, int32_t L, int32_t a, int32_t G,
                               const double* A, const double* B,
                               const int32_t* t_size,
                               const double* const* t_aa,
                               const double* const* t_ab,
                               const double* const* t_bb,
                               const int32_t* l_maxs, int32_t max_lmax,
                               const double* sqrt_u_norms,
                               const hsgen_opts* opts, double* H, double* S,
                               int64_t ld, hsgen_report* report);

/* Bytes of device scratch the allocation-free entry point below has per GPU. */
#define HSGEN_NOALLOC_SCRATCH_BYTES ((uint64_t)256 << 20)

/* hsgen_create_hs in FP64 with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * call's workspaces (a host G list's device copy, the grouped-GEMM work
 * lists, the operands: 2 x 16 B x (2K + 32) x roundup(ng, 128) for the
 * default tile, 3 x for a system with a zero ||u_dot||^2) live in one device
 * array of HSGEN_NOALLOC_SCRATCH_BYTES
 * per GPU, held by the library until the process ends; a problem that needs
 * more is HSGEN_ERR_ARG.  H and S must be device memory (HSGEN_ERR_ARG
 * otherwise); kp->g may be host or device memory.  opts->precision must be
 * HSGEN_PREC_FP64 (HSGEN_ERR_UNSUPPORTED otherwise); flags and tile as for
 * hsgen_create_hs, with the tile's T-block padding already in place (the
 * default tile's is made by hsgen_system_create).  Results equal
 * hsgen_create_hs's bit for bit.  Calls on one GPU share its array: calls
 * that overlap on one GPU, from any contexts, race. */
hsgen_status hsgen_create_hs_noalloc(hsgen_ctx* ctx, const hsgen_sys* sys,
                                     const hsgen_kpoint* kp, const hsgen_opts* opts,
                                     double* H, double* S, int64_t ld,
                                     hsgen_report* report);

/* The step after H/S: the generalized Hermitian eigenproblem H x = lambda S x
 * of one k-point.  No counterpart in the reference, which stops at H and S;
 * Fleur hands them to its own solvers (eigen.F:750,781).  H and S: n x n
 * lower triangles with leading dimension ld (host or device; not modified).
 * w[n] receives the eigenvalues in ascending order (host or device); V, if
 * not NULL, the S-orthonormal eigenvectors (n x n, ldv; host or device).
 * rocSOLVER's zhegvd on the context's stream; *ms (may be NULL) is its GPU
 * time.  HSGEN_ERR_ARG if S is not positive definite. */
hsgen_status hsgen_eigensolve(hsgen_ctx* ctx, int32_t n, const double* H, const double* S, int64_t ld,
                              double* w, double* V, int64_t ldv, double* ms);

/* Diagnostic, no counterpart in the reference: the sustained rate of the
 * INT8 MFMA instruction the split-precision GEMMs use, on every CU, with
 * operands that change from one instruction to the next (random bytes, as
 * residues are) and nothing else running.  The board's power limit lowers
 * the clock under this load, so *tops (INT8 TOP/s) is the ceiling those
 * GEMMs can reach on this board, below the nominal 5.0 POP/s.  iters <= 0:
 * 300000 iterations (about half a second).  *tops and *ms may be NULL. */
hsgen_status hsgen_int8_mfma_ceiling(hsgen_ctx* ctx, int32_t iters, double* tops, double* ms);

/* ---------------------------------------------------------------------------
 * System set-up on the host (no GPU needed): what turns a Fleur dump into the
 * hsgen_system above.
 * ------------------------------------------------------------------------ */

/* Scalar-relativistic radial function of angular momentum l at energy E and
 * its energy derivative for the spherical potential vr = r V(r) (Hartree) on
 * the logarithmic mesh r_i = rmt exp(dx (i + 1 - n)), i = 0..n-1 (n >= 7).
 * match[5] receives {u(R), u'(R), u_dot(R), u_dot'(R), <u_dot|u_dot>} at
 * R = rmt (the rad_match order, types.h:35-43).  f and g may be NULL, else
 * n x 2 column-major: the normalised large and small components of r u and of
 * r u_dot.  Replaces radial_solve (special_funcs.cpp:21-36) -> wrap_radfun
 * (libflerp.f90:50-93) -> Fleur's radfun/radsra/radsrd/intgr0
 * (radfun.f:3-59, radsra.f:10-120, radsrd.F:30-160, intgr.F:58-116). */
hsgen_status hsgen_radial_solve(int32_t l, double E, double dx, double rmt, int32_t n,
                                const double* vr, double* match, double* f, double* g);

/* Integer quadrature and multistep weights the radial solver derives from the
 * Lagrange interpolation polynomials: which = 0: 5 x 7 first-interval weights
 * over 60480 (row j integrates [t_j, t_j+1] from nodes 0..6); 1: 7 closed
 * Newton-Cotes weights over 140; 2: 6 Adams-Bashforth weights over 1440;
 * 3: 6 Adams-Moulton weights over 1440.  (The tables of intgr.F:39-50 and
 * radsra.f:92-104; exported for tests.) */
hsgen_status hsgen_quadrature_weights(int32_t which, double* w);

/* Final T-matrices of one atom from the dumped blocks of its type.
 * t4: 4 x n x n complex, n = (lmax+1)^2, column-major slices t_uu, t_dd,
 * t_ud, t_du of which only the lower triangles are read (the /tmat/<type>
 * dataset layout); E[lmax+1] the energy parameters, ddn[lmax+1] the
 * <u_dot|u_dot> of the atom.  Writes full n x n complex T_aa, T_ab, T_bb
 * (T_aa, T_bb Hermitian; T_ba = T_ab^H).  Replaces the TMat constructor,
 * TMat::zero_nonsph and AtomData::augment_tmats (types_helper.cpp:41-88,
 * AtomData.cpp:95-145).  HSGEN_ERR_ARG if the diagonals of t_ud and t_du
 * differ (they are one matrix). */
hsgen_status hsgen_tmat_assemble(int32_t lmax, int32_t lnonsph, const double* t4,
                                 const double* E, const double* ddn, double* t_aa,
                                 double* t_ab, double* t_bb);

#ifdef __cplusplus
}
#endif

#endif /* HSGEN_H */
