// Host-side state and helpers shared by the hsgen C-ABI translation units:
//   context.hip     contexts, device-resident systems, workspace, host copies
//   fp64_path.hip   tiling, the single-k and batched FP64 paths, construct_ab,
//                   create_hs, create_hs_batched, contract_hs
//   split_path.hip  split precision (Ozaki scheme II on INT8 MFMA), single-k
//                   and batched
//   tune_eig.hip    the tile autotuner and the generalized eigensolve
//   noalloc.hip     the allocation-free FP64 entry point
//
// Operand layout in HBM, per k-point (complex128, column-major, ld = ldz).
// Folded (every ||u_dot_l||^2 > 0, the usual case):
//   Y  = [A_1; N_1^1/2 B_1; A_2; ...]   rows 2*off_a .. 2*off_a + 2*nlm_a - 1
//   W  = [M'_1 Y_1; M'_2 Y_2; ...]      M'_a = D_a^-1/2 M_a D_a^-1/2
// so that   S = Y^H Y   and   H = Y^H W   (only i >= j stored); otherwise
//   Z  = [A_1; B_1; A_2; B_2; ...]
//   Zs = [A_1; N_1 B_1; ...]            N_a = diag(||u_dot_l||^2)
//   W  = [M_1 Z_1; M_2 Z_2; ...]        M_a = [[T_aa, T_ab], [T_ab^H, T_bb]]
// so that   S = Z^H Zs   and   H = Z^H W.  Y lives where Z does (ctx->Z).
// ldz = roundup(2K, BK) + BK, columns padded to the tile size; every padding
// element is zero so the kernels need no bounds checks along k.
#pragma once

#include <complex>
#include <cstdarg>
#include <list>
#include <map>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "../../include/hsgen.h"
#include "hsgen_internal.h"

namespace hsgen {

constexpr int kDefaultTile = 11;
// FP64 tile for problems too small to fill the GPU with 128 x 128 tiles:
// fewer than kSmallWaves tiles of 128 x 128 per CU (S and H lower tiles of
// all k-points of the call) take the 64 x 64 tile, whose workgroups are a
// quarter of the work and two of which a CU holds at once
constexpr int kSmallTile = 12;
constexpr int kSmallWaves = 6;
// INT8 configuration for calls with few S and H lower tiles of 256 x 192 (all
// k-points): 128 x 128 below kInt8FewTiles, 128 x 256 below kInt8SomeTiles;
// BF16: 256 x 128 below kInt8SomeTiles
constexpr int kInt8SmallCfg = 14, kInt8MidCfg = 15, kBf16SmallCfg = 62;
constexpr int kInt8FewTiles = 64, kInt8SomeTiles = 160;
constexpr int kFallbackTile = 8;   // register-staged, 64-bit addressing
constexpr int kIgemmCfg = 51;      // default INT8 kernel configuration (256 x 192 tiles)
constexpr int kBf16Cfg = 61;       // default BF16 kernel configuration (256 x 192 tiles)
constexpr int kSampleTile = 12;    // FP64 tile of the split-precision sample: 64 x 64, so 64 columns
constexpr size_t kMaxOzGraphs = 16; // cached split-precision launch-segment sets per context

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Where atom a's M_a block lives in the system's device allocation.
struct MLoc {
    int64_t off;       // offset of M_a (complex units)
    int32_t ld;        // leading dimension (2 nlm_a rounded up to the tile's BK)
    int32_t cols;      // padded column count (2 nlm_a rounded up to the tile's BM)
};

}  // namespace hsgen

struct hsgen_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    // workspace
    hsgen::DevBuf Z, Zs, W, Htmp, Stmp, A, B, g, probs, work, unorm2, atoms_tmp;
    // split precision
    hsgen::DevBuf oz_e, oz_res, oz_out, oz_probs, oz_work, oz_smp, oz_fprob;
    hipEvent_t oz_ev[2 * hsgen::kMaxModuli] = {};   // INT8 GEMM start/stop per residue group
    int n_cu = 0;                                    // compute units (persistent grids)
    // batched k-points: one captured graph per batch, replayed while the
    // call signature is unchanged
    struct BatchGraph {
        std::vector<int64_t> sig;
        hipGraphExec_t exec = nullptr;
        hsgen::DevBuf g, probs, work;
    };
    std::vector<BatchGraph> bgraphs;
    hsgen::DevBuf btmp;                              // device H, S of host-output k-points
    // batched split precision: captured launch segments of one evaluation
    // (per residue group its residue pass and its INT8 launches, then the
    // recombination) with their own problem and tile lists, replayed while
    // their key (modulus count, k-points, outputs, workspace) is unchanged
    struct OzGraph {
        std::vector<int64_t> key;
        hsgen::DevBuf probs, work, fprob;
        std::vector<hipGraphExec_t> seg;
    };
    std::list<OzGraph> ozgraphs;                     // oldest first, at most kMaxOzGraphs
    // autotuned tile configurations: (precision, size buckets) -> config id
    std::map<std::vector<int>, int> tuned;
    // generalized eigensolver
    rocblas_handle blas = nullptr;
    hsgen::DevBuf eig_a, eig_b, eig_d;
    // pinned staging for device -> host copies of H and S
    void* stage[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
};

struct hsgen_sys {
    uint64_t serial = 0;               // unique per created system
    int natoms = 0, max_lmax = 0, nl = 0, L = 0, kdim = 0;
    double omega = 0.0, vol_prefac = 0.0;
    double bmat[9];
    std::vector<hsgen::AtomDev> atoms; // host copy
    std::vector<int64_t> m_off;        // per atom: offset of M_a in d_M (complex units)
    std::vector<int32_t> m_ld;         // per atom: leading dim of M_a
    std::vector<int32_t> m_cols;       // per atom: padded column count
    int m_bk = 0, m_bm = 0;            // tile geometry M was padded for
    int64_t sum_nlm2 = 0;              // sum_a nlm_a^2
    // folded operands: Y = [A_a; sqrt(N_a) B_a] per atom (S = Y^H Y,
    // H = Y^H M' Y) when every N_l = ||u_dot_l||^2 is positive; otherwise
    // Z and Zs = [A_a; N_a B_a] (S = Z^H Zs, H = Z^H M Z)
    bool fold = false;
    std::vector<double> h_sqrtn;       // [natoms][nl] sqrt(N_l), 0 for l > lmax
    hsgen::AtomDev* d_atoms = nullptr;
    hsgen::RadDev* d_rad = nullptr;
    double* d_ynorm = nullptr;
    double2* d_M = nullptr;
    hsgen::MLoc* d_mloc = nullptr;     // [natoms] M_a layout, in the same allocation as d_M
    // host copies of the T blocks, so that M can be rebuilt for another tile
    // geometry
    std::vector<std::vector<std::complex<double>>> h_taa, h_tab, h_tbb;
    std::vector<int> h_owner;          // per atom: the atom whose copies hold its T
    // first atom of each rotation class and of each rmt class (AtomDev::cls)
    std::vector<int32_t> rot_rep, rmt_rep;
};

namespace hsgen {

// Record the error text on the context and return the status.
hsgen_status fail(hsgen_ctx* ctx, hsgen_status s, const char* fmt, ...);

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::hsgen::fail(ctx, HSGEN_ERR_HIP, "%s failed: %s (%s:%d)", #call,  \
                                 hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

inline int64_t roundup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// ---- context.hip ----
// Grow a workspace buffer to at least `bytes` (contents not kept).
hsgen_status ensure(hsgen_ctx* ctx, DevBuf& b, size_t bytes, bool zero_new);
void release(DevBuf& b);
bool is_device_ptr(const void* p);
// M_a blocks of every atom padded for tile geometry (bk, bm), deduplicated by
// T pointer triple, uploaded to s->d_M.
hsgen_status upload_M(hsgen_ctx* ctx, hsgen_sys* s, const int32_t* nlm, const double* const* taa,
                      const double* const* tab, const double* const* tbb, const double* rsc, int bk, int bm);
// Folded operands unless HSGEN_FOLD=0 (environment; tests and comparisons).
bool fold_enabled();
// N_l usable for folding (positive, 1/sqrt(N) far from overflow).
bool fold_ok(double n);
hsgen_status ensure_M(hsgen_ctx* ctx, hsgen_sys* s, int bk, int bm);
hsgen_status check_kpoint(hsgen_ctx* ctx, const hsgen_kpoint* kp);
hsgen_status upload_g(hsgen_ctx* ctx, const hsgen_kpoint* kp);
hsgen_status copy_lower_to_host(hsgen_ctx* ctx, const double2* dsrc, int n, double* H, int64_t ld, bool full);

// ---- fp64_path.hip ----
struct Geometry {
    int cfg;
    ZgemmLaunchCfg t;
    int64_t kpad;   // kdepth of the S/H contraction = roundup(2K, BK)
    int64_t ldz;    // kpad + BK
    int64_t gp;     // padded column count
    bool fold;      // operands Y (in Z) and W: S = Y^H Y; else Z, Zs, W
};

// FP64 operand arrays per column set: Y, W (folded) or Z, Zs, W.
inline int n_operands(const Geometry& g) { return g.fold ? 2 : 3; }
// The right operand of S: Zs, or Y itself (in ctx->Z) when folded.
inline double2* zs_ptr(hsgen_ctx* ctx, const Geometry& g) { return (double2*)(g.fold ? ctx->Z.p : ctx->Zs.p); }

// Split precision (INT8 or BF16 MFMA on the same residues).
inline bool is_split(int precision)
{
    return precision == HSGEN_PREC_OZAKI_INT8 || precision == HSGEN_PREC_OZAKI_BF16;
}
inline bool valid_precision(int precision) { return precision == HSGEN_PREC_FP64 || is_split(precision); }
// The split-precision kernel configuration of a call: opts.tile when it names
// one for the precision's MFMA type, else the precision's default.
int split_cfg(int precision, int tile);
// Column padding of split-precision operands: a multiple of both tile sides.
int64_t split_col_align(int precision, int tile);
Geometry geometry(int tile, int kdim, int ng, int64_t col_align = 1);
void add_hs_problems(std::vector<GemmProblem>& probs, std::vector<TileWork>& work, const Geometry& geo, int ng,
                     const double2* Z, const double2* Zs, const double2* W, double2* H, double2* S, int64_t ld,
                     int flags);
void xcd_permute(std::vector<TileWork>& work);
AbgenParams abgen_params(const hsgen_sys* s, const hsgen_kpoint* kp, const int32_t* dg);
int tapp_cfg(const Geometry& geo);
void add_tmul_problems(std::vector<GemmProblem>& probs, std::vector<TileWork>& work, const hsgen_sys* s,
                       const Geometry& geo, const ZgemmLaunchCfg& tc, int64_t ncols, const double2* Z, double2* W);
hsgen_status run_grouped(hsgen_ctx* ctx, int cfg, std::vector<GemmProblem>& probs, std::vector<TileWork>& work);
hsgen_status finish_hs(hsgen_ctx* ctx, hsgen_sys* s, const Geometry& geo, int ng, const hsgen_opts& o, double* H,
                       double* S, int64_t ld, hsgen_report* rep, double ms_front);

// ---- split_path.hip ----
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

// capture: the INT8 launch sequence of each modulus count runs as cached
// hipGraphs (the batched path), else eagerly.
hsgen_status run_ozaki(hsgen_ctx* ctx, const Geometry& geo, const std::vector<OzK>& ks, int flags, int icfg,
                       int nmod_req, double target, std::vector<OzResult>& res, bool capture);
void destroy_oz_graph(hsgen_ctx::OzGraph& g);
hsgen_status batched_split(hsgen_ctx* ctx, hsgen_sys* s, int nk, const hsgen_kpoint* kps, const hsgen_opts& o,
                           double* const* H, double* const* S, const int64_t* ld, hsgen_report* rep);

// ---- noalloc.hip ----
// Look up this context's device's scratch array once (the device current).
hsgen_status noalloc_setup(hsgen_ctx* ctx);

// ---- tune_eig.hip ----
// Size buckets of the autotuning table: half-octaves of G and K.
std::vector<int> tune_key(int precision, int ng, int kdim);
// Tile configuration for a call: the caller's, else the autotuned one for
// this size, else 0 (the built-in default).
// The tile of a call: opts.tile if set, else the autotuned one for the size,
// else the default by size over the call's k-points (kps[0 .. nk-1], or one
// k-point of ng when kps is null): see fp64_tile_by_size and int8_cfg_by_size.
int pick_tile(const hsgen_ctx* ctx, const hsgen_opts& o, int ng, int kdim, const hsgen_kpoint* kps = nullptr,
              int nk = 0);
// FP64 tile by size: kSmallTile when the S and H lower tiles of 128 x 128 of
// all the k-points number fewer than kSmallWaves per CU, else 0 (the default)
int fp64_tile_by_size(const hsgen_ctx* ctx, int ng, const hsgen_kpoint* kps = nullptr, int nk = 0);

}  // namespace hsgen
