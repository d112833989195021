// hsgen_driver: the reference's h_and_s executable (generate_matrices and
// main, flapw_opt/src/h_and_s.cpp:896-957) over the hsgen C ABI.  It reads a
// Fleur dump (FleurData, here hsgen_dump_*), builds H and S of its first
// k_ind_max k-points on the GPU (create_HS2, here hsgen_create_hs or one
// hsgen_create_hs_batched call) and writes them as HDF5Store does
// (/matrix_hamilton/<k+1>, /matrix_overlap/<k+1>).  Where the dump holds
// Fleur's own H and S it prints the Frobenius norms of the lower-triangle
// differences, the figures flapw_verify.py:52-66 prints, and their relative
// size; --check TOL makes a relative difference above TOL an error.
//
// usage: hsgen_driver DUMP.h5 [K_IND_MAX] [--precision fp64|int8|bf16]
//                     [--target REL] [--out PATH] [--check TOL] [--batched]
//
// Plain C++11 against include/hsgen.h and include/hsgen_io.h only: the code a
// C++ caller of the library writes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hsgen.h"
#include "hsgen_io.h"

namespace {

void usage()
{
    std::fprintf(stderr,
                 "usage: hsgen_driver DUMP.h5 [K_IND_MAX] [--precision fp64|int8|bf16] [--target REL]\n"
                 "                    [--out PATH] [--check TOL] [--batched]\n");
}

struct Args {
    std::string dump, out = "flapw_out.h5";
    int k_ind_max = 60;                 // the reference's default
    int precision = HSGEN_PREC_FP64;
    double target = 0.0;
    double check = -1.0;
    bool batched = false;
};

bool parse(int argc, char** argv, Args& a)
{
    int pos = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        const bool has_val = i + 1 < argc;
        if (s == "--precision" && has_val) {
            const std::string p = argv[++i];
            if (p == "fp64") a.precision = HSGEN_PREC_FP64;
            else if (p == "int8") a.precision = HSGEN_PREC_OZAKI_INT8;
            else if (p == "bf16") a.precision = HSGEN_PREC_OZAKI_BF16;
            else return false;
        } else if (s == "--target" && has_val) {
            a.target = std::atof(argv[++i]);
        } else if (s == "--out" && has_val) {
            a.out = argv[++i];
        } else if (s == "--check" && has_val) {
            a.check = std::atof(argv[++i]);
        } else if (s == "--batched") {
            a.batched = true;
        } else if (!s.empty() && s[0] == '-') {
            return false;
        } else if (pos == 0) {
            a.dump = s;
            ++pos;
        } else if (pos == 1) {
            a.k_ind_max = std::atoi(s.c_str());
            ++pos;
        } else {
            return false;
        }
    }
    return !a.dump.empty() && a.k_ind_max > 0;
}

// ||tril(X - R)||_F and ||tril R||_F of column-major n x n complex matrices
void tril_diff(const std::vector<double>& X, const std::vector<double>& R, int64_t n, double& diff, double& ref)
{
    double d = 0.0, r = 0.0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = j; i < n; ++i) {
            const int64_t e = 2 * (i + j * n);
            const double dr = X[e] - R[e], di = X[e + 1] - R[e + 1];
            d += dr * dr + di * di;
            r += R[e] * R[e] + R[e + 1] * R[e + 1];
        }
    diff = std::sqrt(d);
    ref = std::sqrt(r);
}

}  // namespace

int main(int argc, char** argv)
{
    Args args;
    if (!parse(argc, argv, args)) {
        usage();
        return 2;
    }
    hsgen_dump* dump = nullptr;
    if (hsgen_dump_open(args.dump.c_str(), &dump) != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
        return 1;
    }
    hsgen_dump_info info;
    hsgen_system desc;
    if (hsgen_dump_get_info(dump, &info) != HSGEN_OK || hsgen_dump_get_system(dump, &desc) != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
        hsgen_dump_close(dump);
        return 1;
    }
    const int nk = std::min(args.k_ind_max, (int)info.num_kpts);
    std::printf("dump %s: %d atoms, %d types, max lmax %d, %d k-points (%d with Fleur H/S); running %d\n",
                args.dump.c_str(), info.num_atoms, info.num_types, info.max_lmax, info.num_kpts, info.num_hs, nk);

    // k-points (get_kpoint / get_g_mesh)
    std::vector<std::vector<int32_t> > gs(nk);
    std::vector<hsgen_kpoint> kps(nk);
    for (int k = 0; k < nk; ++k) {
        hsgen_kpoint& kp = kps[k];
        int32_t ng = 0;
        if (hsgen_dump_get_kpoint(dump, k, kp.k, &ng, nullptr) != HSGEN_OK) {
            std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
            hsgen_dump_close(dump);
            return 1;
        }
        gs[k].resize(3 * (size_t)std::max(ng, 1));
        if (hsgen_dump_get_kpoint(dump, k, nullptr, nullptr, gs[k].data()) != HSGEN_OK) {
            std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
            hsgen_dump_close(dump);
            return 1;
        }
        kp.ng = ng;
        kp.g = gs[k].data();
    }

    int rc = 0;
    hsgen_ctx* ctx = nullptr;
    hsgen_sys* sys = nullptr;
    hsgen_store* store = nullptr;
    std::vector<std::vector<double> > H(nk), S(nk);
    std::vector<hsgen_report> rep(nk);
    std::vector<double> secs(nk, 0.0);
    hsgen_opts opts;
    std::memset(&opts, 0, sizeof opts);
    opts.precision = args.precision;
    opts.target_rel_err = args.target;
    for (int k = 0; k < nk; ++k) {
        H[k].assign(2 * (size_t)kps[k].ng * kps[k].ng, 0.0);
        S[k].assign(2 * (size_t)kps[k].ng * kps[k].ng, 0.0);
    }
    // the device context replaces dev_init (h_and_s.cpp:793-796)
    const hsgen_status ist = hsgen_init(&ctx, 0);
    if (ist != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: no context on device 0: %s\n", hsgen_status_string(ist));
        rc = 1;
    } else if (hsgen_system_create(ctx, &desc, &sys) != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_last_error(ctx));
        rc = 1;
    } else if (args.batched) {
        std::vector<double*> hp(nk), sp(nk);
        std::vector<int64_t> ld(nk);
        for (int k = 0; k < nk; ++k) {
            hp[k] = H[k].data();
            sp[k] = S[k].data();
            ld[k] = kps[k].ng;
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (hsgen_create_hs_batched(ctx, sys, nk, kps.data(), &opts, hp.data(), sp.data(), ld.data(),
                                    rep.data()) != HSGEN_OK) {
            std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_last_error(ctx));
            rc = 1;
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("batched call over %d k-points: %.4f s\n", nk, dt);
        for (int k = 0; k < nk; ++k) secs[k] = dt / nk;
    } else {
        for (int k = 0; k < nk && rc == 0; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            if (hsgen_create_hs(ctx, sys, &kps[k], &opts, H[k].data(), S[k].data(), kps[k].ng, &rep[k]) !=
                HSGEN_OK) {
                std::fprintf(stderr, "hsgen_driver: k-point %d: %s\n", k, hsgen_last_error(ctx));
                rc = 1;
            }
            secs[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    // H/S output (HDF5Store::add_hsmat) and the comparison with Fleur's
    if (rc == 0 && hsgen_store_open(args.out.c_str(), &store) != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
        rc = 1;
    }
    for (int k = 0; k < nk && rc == 0; ++k) {
        const int64_t n = kps[k].ng;
        if (hsgen_store_add_hs(store, k, (int32_t)n, H[k].data(), S[k].data(), n) != HSGEN_OK) {
            std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
            rc = 1;
            break;
        }
        std::printf("k %d: ng %lld, %.4f s%s", k, (long long)n, secs[k], args.batched ? " (share of the call)" : "");
        if (args.precision != HSGEN_PREC_FP64)
            std::printf(", %d slices, model rel err H %.3e S %.3e", rep[k].slices_used, rep[k].err_est_h,
                        rep[k].err_est_s);
        if (k < info.num_hs) {
            int64_t nf = 0;
            hsgen_dump_get_hs_packed(dump, k, &nf, nullptr, nullptr);
            if (nf != n) {
                std::printf("\n");
                std::fprintf(stderr, "hsgen_driver: k-point %d: Fleur's H/S are %lld x %lld, not %lld\n", k,
                             (long long)nf, (long long)nf, (long long)n);
                rc = 1;
                break;
            }
            std::vector<double> hpk((size_t)n * (n + 1)), spk((size_t)n * (n + 1));
            std::vector<double> Hf(2 * (size_t)n * n, 0.0), Sf(2 * (size_t)n * n, 0.0);
            if (hsgen_dump_get_hs_packed(dump, k, &nf, hpk.data(), spk.data()) != HSGEN_OK ||
                hsgen_unpack_lower(n, hpk.data(), Hf.data(), n) != HSGEN_OK ||
                hsgen_unpack_lower(n, spk.data(), Sf.data(), n) != HSGEN_OK) {
                std::printf("\n");
                std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
                rc = 1;
                break;
            }
            double dh, rh, ds, rs;
            tril_diff(H[k], Hf, n, dh, rh);
            tril_diff(S[k], Sf, n, ds, rs);
            const double eh = rh > 0.0 ? dh / rh : dh, es = rs > 0.0 ? ds / rs : ds;
            std::printf("; vs Fleur ||tril dH|| %.3e ||tril dS|| %.3e, rel H %.3e S %.3e", dh, ds, eh, es);
            if (args.check >= 0.0 && !(eh <= args.check && es <= args.check)) {
                std::printf("\n");
                std::fprintf(stderr, "hsgen_driver: k-point %d differs from Fleur's H/S by more than %g\n", k,
                             args.check);
                rc = 1;
                break;
            }
        }
        std::printf("\n");
    }
    if (store && hsgen_store_close(store) != HSGEN_OK) {
        std::fprintf(stderr, "hsgen_driver: %s\n", hsgen_io_last_error());
        rc = 1;
    }
    if (rc == 0) std::printf("H, S of %d k-points written to %s\n", nk, args.out.c_str());
    if (sys) hsgen_system_destroy(ctx, sys);
    if (ctx) hsgen_finalize(ctx);
    hsgen_dump_close(dump);
    return rc;
}
