// Host-side system setup: final per-atom T-matrices from the four blocks a
// Fleur dump holds per atom type.
//
// The dump stores, for each type, the lower triangles of t_uu, t_dd, t_ud and
// t_du (n x n complex, n = (lmax+1)^2, column-major; the upper triangles
// are not meaningful).  The contraction needs
//   T_aa = t_uu + t_uu^H - diag,  T_bb likewise from t_dd   (Hermitian),
//   T_ab = lower(t_ud) + strictly-upper(t_du^H),  diag from t_ud,
// (types_helper.cpp:41-65), with every row and column of index
// >= (lnonsph+1)^2 cleared (types_helper.cpp:73-88), plus the spherical
// diagonal terms Fleur keeps out of the dumped matrices: E_l on T_aa,
// E_l <u_dot|u_dot>_l on T_bb and 1/2 on T_ab (AtomData.cpp:95-145).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "hsgen.h"

namespace {

struct Cx {
    double re, im;
};

}  // namespace

extern "C" {

hsgen_status hsgen_tmat_assemble(int32_t lmax, int32_t lnonsph, const double* t4, const double* E,
                                 const double* ddn, double* t_aa, double* t_ab, double* t_bb)
{
    if (lmax < 0 || lnonsph < 0 || lnonsph > lmax || !t4 || !E || !ddn || !t_aa || !t_ab || !t_bb)
        return HSGEN_ERR_ARG;
    const int64_t n = (int64_t)(lmax + 1) * (lmax + 1);
    const Cx* uu = (const Cx*)t4;
    const Cx* dd = uu + n * n;
    const Cx* ud = dd + n * n;
    const Cx* du = ud + n * n;
    Cx* aa = (Cx*)t_aa;
    Cx* ab = (Cx*)t_ab;
    Cx* bb = (Cx*)t_bb;
    // the two mixed blocks must agree on the diagonal (they are one matrix)
    for (int64_t i = 0; i < n; ++i) {
        const Cx x = ud[i * n + i], y = du[i * n + i];
        const double sc = 1.0 + std::fabs(x.re) + std::fabs(x.im);
        if (std::fabs(x.re - y.re) > 1e-12 * sc || std::fabs(x.im - y.im) > 1e-12 * sc) return HSGEN_ERR_ARG;
    }
    const int64_t ns = (int64_t)(lnonsph + 1) * (lnonsph + 1);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            const int64_t o = j * n + i;
            if (i >= ns || j >= ns) {
                aa[o] = bb[o] = ab[o] = Cx{0.0, 0.0};
                continue;
            }
            const int64_t t = i * n + j;      // transposed position
            if (i > j) {
                aa[o] = uu[o];
                bb[o] = dd[o];
                ab[o] = ud[o];
            } else if (i < j) {
                aa[o] = Cx{uu[t].re, -uu[t].im};
                bb[o] = Cx{dd[t].re, -dd[t].im};
                ab[o] = Cx{du[t].re, -du[t].im};
            } else {
                aa[o] = uu[o];
                bb[o] = dd[o];
                ab[o] = ud[o];
            }
        }
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) {
            const int64_t lm = (int64_t)l * (l + 1) + m;
            const int64_t o = lm * n + lm;
            aa[o].re += E[l];
            bb[o].re += E[l] * ddn[l];
            ab[o].re += 0.5;
        }
    return HSGEN_OK;
}

}  // extern "C"
