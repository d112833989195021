/*
 * CPU oracle for FLAPW muffin-tin H/S generation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference
 * algorithm of the HybridHSDLA hot path so that the HIP implementation can be
 * checked against it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only to compare results with or as the
 * timed CPU baseline - never as the product path.
 *
 * Restated functions and the reference lines they follow:
 *   orc_sphbes        flapw_opt/external/compat/fleur/sphbes.f:7-76
 *   orc_dsphbs        flapw_opt/external/compat/fleur/dsphbs.f:11-47
 *   orc_ylm4          flapw_opt/external/compat/fleur/ylm4.f:4-139
 *   orc_kmesh_rk      flapw_opt/src/types_helper.cpp:20-37
 *   orc_construct_ab  flapw_opt/src/tensor_AB_match.cpp:63-129
 *                     (+ AtomData.cpp:63-66 inverse Wronskian,
 *                        FleurData.cpp:46 cell volume)
 *   orc_create_hs2_v2 flapw_opt/src/h_and_s_stripped_v2.cpp:40-382
 *                     (+ wrappers.cpp:171-257 CPU branches: zherk,
 *                        zher2k, zgemmt) - the reference's BLAS sequence,
 *                     run on the host OpenBLAS loaded with dlopen
 *   orc_naive_hs      flapw_opt/src/h_and_s.cpp:84-162 (element-wise sums)
 *
 * How each function is pinned (tests/test_oracle.py):
 *   sphbes/dsphbs  scipy.special.spherical_jn tables incl. the x<1e-3 and
 *                  x<1e-4 series branches, closed forms j0, j1, Wronskian
 *   ylm4           scipy.special.sph_harm_y tables, sum rule, m<->-m symmetry,
 *                  addition theorem, z-axis special case
 *   construct_ab   Legendre (addition-theorem) form of the spherical overlap
 *                  (hssphn.F:677-681 form), brute force on tiny inputs
 *   create_hs2_v2  orc_naive_hs on tiny inputs, T=I / T=0 special cases,
 *                  Hermiticity, atom permutation invariance
 *
 * Complex numbers are C99 double complex, memory-compatible with the
 * interleaved (re, im) complex128 layout used everywhere else.
 */
#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------ */
/* Spherical Bessel functions, sphbes.f:7-76                           */
/* ------------------------------------------------------------------ */
int orc_sphbes(int lmax, double x, double* fj)
{
    const double small = 1.0e-3;
    if (lmax < 1) return -1;          /* sphbes.f:29 STOP 'sphbes1' */
    if (x < 0.0) return -2;           /* sphbes.f:30 STOP 'sphbes2' */
    double xx = x * x;
    double xinv = 0.0;
    if (x >= small) {
        xinv = 1.0 / x;
        fj[0] = sin(x) * xinv;
        fj[1] = (fj[0] - cos(x)) * xinv;
    } else {
        /* series branch, sphbes.f:37-44 */
        fj[0] = 1.0 - xx / 6.0 * (1.0 - xx / 20.0 * (1.0 - xx / 42.0));
        fj[1] = (x / 3.0) * (1.0 - xx / 10.0 * (1.0 - xx / 28.0));
        double fac = xx / 15.0;
        for (int l = 2; l <= lmax; ++l) {
            fj[l] = fac * (1.0 - xx / (4 * l + 6));
            fac = x * fac / (2 * l + 3);
        }
        return 0;
    }
    if ((double)lmax < x) {
        /* upward recurrence, sphbes.f:46-50 */
        for (int l = 2; l <= lmax; ++l)
            fj[l] = (2 * l - 1) * xinv * fj[l - 1] - fj[l - 2];
    } else if (lmax >= 2) {
        /* Miller downward recurrence, sphbes.f:52-72 */
        int n = (int)(lmax + 10 + x);
        double* aux = (double*)malloc((size_t)(n + 1) * sizeof(double));
        aux[n] = 0.0;
        aux[n - 1] = 1.0;
        for (int i = n - 1; i >= 1; --i)
            aux[i - 1] = (2 * i + 1) * xinv * aux[i] - aux[i + 1];
        int mn = 0;
        if (fabs(fj[0]) < fabs(fj[1])) mn = 1;
        double quot = fj[mn] / aux[mn];
        for (int l = 2; l <= lmax; ++l) fj[l] = aux[l] * quot;
        free(aux);
    }
    return 0;
}

/* Derivative dj_l/dx, dsphbs.f:11-47 */
void orc_dsphbs(int lmax, double x, const double* fj, double* dfj)
{
    const double xlim = 1.0e-4;
    dfj[0] = -fj[1];
    if (x < xlim) {
        double x2 = 0.5 * x * x;
        double fac = 1.0 / 3.0;
        for (int l = 1; l <= lmax; ++l) {
            dfj[l] = fac * (l - x2 * (l + 2) / (2 * l + 3));
            fac = x * fac / (2 * l + 3);
        }
    } else {
        for (int l = 1; l <= lmax; ++l)
            dfj[l] = fj[l - 1] - (l + 1) * fj[l] / x;
    }
}

/* ------------------------------------------------------------------ */
/* Spherical harmonics, ylm4.f:4-139                                   */
/* ylm[l*(l+1)+m], l = 0..lmax, m = -l..l                              */
/* ------------------------------------------------------------------ */
void orc_ylm4(int lmax, const double* v, cplx* ylm)
{
    const double small = 1.0e-12;
    int n = lmax + 1;
    double* p = (double*)calloc((size_t)n * n, sizeof(double)); /* p[l*n+m] */
    double* c = (double*)calloc((size_t)n, sizeof(double));
    double* s = (double*)calloc((size_t)n, sizeof(double));
    double* ynorm = (double*)calloc((size_t)n * n, sizeof(double));

    /* normalisation, ylm4.f:236-262 */
    double fpi = 4.0 * ORC_PI;
    for (int l = 0; l <= lmax; ++l) {
        int lm0 = l * (l + 1);
        double cd = 1.0;
        double a = sqrt((2 * l + 1) / fpi);
        ynorm[lm0] = a;
        for (int m = 1; m <= l; ++m) {
            cd = cd / ((double)(l + 1 - m) * (double)(l + m));
            ynorm[lm0 + m] = a * sqrt(cd);
            ynorm[lm0 - m] = ((m & 1) ? -1.0 : 1.0) * ynorm[lm0 + m];
        }
    }

    double x = v[0], y = v[1], z = v[2];
    double xy = x * x + y * y;
    double r = sqrt(xy + z * z);
    double rxy = sqrt(xy);
    double cth, sth, cph, sph;
    if (r > small) { cth = z / r; sth = rxy / r; }
    else { sth = 0.0; cth = 1.0; }
    if (rxy > small) { cph = x / rxy; sph = y / rxy; }
    else { cph = 1.0; sph = 0.0; }

    /* associated Legendre functions for m >= 0, ylm4.f:193-206 */
    double fac = 1.0;
    for (int m = 0; m <= lmax - 1; ++m) {
        fac = -(m + m - 1) * fac;
        p[m * n + m] = fac;
        p[(m + 1) * n + m] = (m + m + 1) * cth * fac;
        for (int l = m + 2; l <= lmax; ++l)
            p[l * n + m] = ((l + l - 1) * cth * p[(l - 1) * n + m]
                            - (l + m - 1) * p[(l - 2) * n + m]) / (l - m);
        fac = fac * sth;
    }
    p[lmax * n + lmax] = -(lmax + lmax - 1) * fac;

    /* cos(m phi), sin(m phi), ylm4.f:208-217 */
    s[0] = 0.0;
    c[0] = 1.0;
    if (lmax >= 1) { s[1] = sph; c[1] = cph; }
    double cph2 = cph + cph;
    for (int m = 2; m <= lmax; ++m) {
        s[m] = cph2 * s[m - 1] - s[m - 2];
        c[m] = cph2 * c[m - 1] - c[m - 2];
    }

    /* ylm4.f:219-230 */
    for (int l = 0; l <= lmax; ++l)
        ylm[l * (l + 1)] = ynorm[l * (l + 1)] * p[l * n + 0];
    for (int m = 1; m <= lmax; ++m) {
        for (int l = m; l <= lmax; ++l) {
            int lm0 = l * (l + 1);
            cplx ylms = p[l * n + m] * (c[m] + I * s[m]);
            ylm[lm0 + m] = ynorm[lm0 + m] * ylms;
            ylm[lm0 - m] = conj(ylms) * ynorm[lm0 - m];
        }
    }
    free(p); free(c); free(s); free(ynorm);
}

/* ------------------------------------------------------------------ */
/* |k+g| for each g, types_helper.cpp:20-37                            */
/* bmat column-major 3x3; g is ng x 3 (triples contiguous)             */
/* ------------------------------------------------------------------ */
void orc_kmesh_rk(const double* bmat, const double* kpt, int ng,
                  const int32_t* g, double* r_k)
{
    for (int i = 0; i < ng; ++i) {
        double K[3], c[3];
        for (int d = 0; d < 3; ++d) K[d] = kpt[d] + g[3 * i + d];
        for (int r = 0; r < 3; ++r)
            c[r] = bmat[r + 0] * K[0] + bmat[r + 3] * K[1] + bmat[r + 6] * K[2];
        r_k[i] = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
}

/* ------------------------------------------------------------------ */
/* A/B matching coefficients, tensor_AB_match.cpp:63-129               */
/* A, B: L x natoms x ng, lm fastest, L = (max_lmax+1)^2, zero-filled. */
/* match: per (atom, l) 5 doubles {u, u_dr, dE_u, dE_u_dr, dE_u_norm}  */
/*        at match[(atom*(max_lmax+1) + l)*5].                         */
/* rotmat: per atom column-major 3x3 (R_atom * bmat).                  */
/* ------------------------------------------------------------------ */
int orc_construct_ab(int natoms, int max_lmax, const int32_t* lmax,
                     const double* rmt, const double* pos,
                     const double* rotmat, const double* match, double omega,
                     const double* bmat, const double* kpt, int ng,
                     const int32_t* g, cplx* A, cplx* B)
{
    const int L = (max_lmax + 1) * (max_lmax + 1);
    const int nl = max_lmax + 1;
    memset(A, 0, sizeof(cplx) * (size_t)L * natoms * ng);
    memset(B, 0, sizeof(cplx) * (size_t)L * natoms * ng);
    double* r_k = (double*)malloc(sizeof(double) * (size_t)(ng > 0 ? ng : 1));
    orc_kmesh_rk(bmat, kpt, ng, g, r_k);
    double* fj = (double*)malloc(sizeof(double) * nl);
    double* dfj = (double*)malloc(sizeof(double) * nl);
    cplx* ylm = (cplx*)malloc(sizeof(cplx) * L);
    const double vol_prefac = 4.0 * ORC_PI / sqrt(omega);
    int rc = 0;
    for (int i = 0; i < ng; ++i) {
        double K[3];
        for (int d = 0; d < 3; ++d) K[d] = kpt[d] + g[3 * i + d];
        double K_norm = r_k[i];
        for (int j = 0; j < natoms; ++j) {
            const double* R = rotmat + 9 * j;
            double Kb[3];
            /* R^T K with R column-major: (R^T K)_c = sum_r R[r + 3c] K[r] */
            for (int cc = 0; cc < 3; ++cc)
                Kb[cc] = R[0 + 3 * cc] * K[0] + R[1 + 3 * cc] * K[1] + R[2 + 3 * cc] * K[2];
            double nrm = sqrt(Kb[0] * Kb[0] + Kb[1] * Kb[1] + Kb[2] * Kb[2]);
            if (nrm != 0.0) { Kb[0] /= nrm; Kb[1] /= nrm; Kb[2] /= nrm; }
            const double* p = pos + 3 * j;
            double kp = K[0] * p[0] + K[1] * p[1] + K[2] * p[2];
            cplx phase = cexp(I * (2.0 * ORC_PI * kp));
            rc = orc_sphbes(max_lmax, rmt[j] * K_norm, fj);
            if (rc) goto done;
            orc_dsphbs(max_lmax, rmt[j] * K_norm, fj, dfj);
            orc_ylm4(max_lmax, Kb, ylm);
            for (int l = 0; l <= lmax[j]; ++l) {
                const double* mt = match + ((size_t)j * nl + l) * 5;
                double u = mt[0], u_dr = mt[1], dE_u = mt[2], dE_u_dr = mt[3];
                double inv_wronk = 1.0 / (dE_u * u_dr - u * dE_u_dr);
                double A_u = dE_u * K_norm * dfj[l] - dE_u_dr * fj[l];
                double B_u = -u * K_norm * dfj[l] + u_dr * fj[l];
                cplx common = vol_prefac * phase * inv_wronk;
                for (int m = -l; m <= l; ++m) {
                    int lm = l * (l + 1) + m;
                    size_t idx = (size_t)lm + (size_t)L * ((size_t)j + (size_t)natoms * i);
                    A[idx] = conj(ylm[lm]) * common * A_u;
                    B[idx] = conj(ylm[lm]) * common * B_u;
                }
            }
        }
    }
done:
    free(r_k); free(fj); free(dfj); free(ylm);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Host BLAS, loaded at run time (OpenBLAS, optionally symbol-prefixed) */
/* ------------------------------------------------------------------ */
typedef void (*zherk_t)(const char*, const char*, const int*, const int*,
                        const double*, const cplx*, const int*,
                        const double*, cplx*, const int*);
typedef void (*zher2k_t)(const char*, const char*, const int*, const int*,
                         const cplx*, const cplx*, const int*, const cplx*,
                         const int*, const double*, cplx*, const int*);
typedef void (*zgemmt_t)(const char*, const char*, const char*, const int*,
                         const int*, const cplx*, const cplx*, const int*,
                         const cplx*, const int*, const cplx*, cplx*,
                         const int*);
typedef void (*zhemm_t)(const char*, const char*, const int*, const int*,
                        const cplx*, const cplx*, const int*, const cplx*,
                        const int*, const cplx*, cplx*, const int*);
typedef void (*zgemm_t)(const char*, const char*, const int*, const int*,
                        const int*, const cplx*, const cplx*, const int*,
                        const cplx*, const int*, const cplx*, cplx*,
                        const int*);
typedef void (*zlacpy_t)(const char*, const int*, const int*, const cplx*,
                         const int*, cplx*, const int*);
typedef void (*setnt_t)(int);

static struct {
    void* h;
    zherk_t zherk; zher2k_t zher2k; zgemmt_t zgemmt; zhemm_t zhemm;
    zgemm_t zgemm; zlacpy_t zlacpy; setnt_t set_num_threads;
} g_blas;

/* Returns 0 on success. prefix is "" for plain OpenBLAS or "scipy_" for the
 * scipy wheel's copy. */
int orc_blas_load(const char* path, const char* prefix)
{
    char name[128];
    if (g_blas.h) return 0;
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
#define LOAD(field, sym)                                              \
    snprintf(name, sizeof name, "%s%s", prefix, sym);                 \
    *(void**)(&g_blas.field) = dlsym(h, name);                        \
    if (!g_blas.field) { dlclose(h); return -2; }
    LOAD(zherk, "zherk_");
    LOAD(zher2k, "zher2k_");
    LOAD(zgemmt, "zgemmt_");
    LOAD(zhemm, "zhemm_");
    LOAD(zgemm, "zgemm_");
    LOAD(zlacpy, "zlacpy_");
#undef LOAD
    snprintf(name, sizeof name, "%sopenblas_set_num_threads", prefix);
    *(void**)(&g_blas.set_num_threads) = dlsym(h, name);
    g_blas.h = h;
    return 0;
}

int orc_blas_set_threads(int n)
{
    if (!g_blas.h || !g_blas.set_num_threads) return -1;
    g_blas.set_num_threads(n);
    return 0;
}

/* ------------------------------------------------------------------ */
/* The reference contraction, h_and_s_stripped_v2.cpp:100-382.         */
/* Same BLAS call sequence, same argument semantics.  A, B, W are       */
/* destroyed as in the reference (compaction and reuse as X).          */
/* T of atom a: n_a x n_a column-major blocks at T_all + t_off[3a+k],   */
/* k = 0 (aa), 1 (ab), 2 (bb), n_a = t_size[a].                        */
/* ------------------------------------------------------------------ */

/* apply_u_norm_s, h_and_s_stripped_v2.cpp:64-86 */
static void orc_apply_u_norm_s(int L, int a, int G, const cplx* B, cplx* C,
                               const int32_t* lmaxs, int max_lmax,
                               const double* u_norms)
{
    double* lin = (double*)calloc((size_t)L * a, sizeof(double));
    for (int at = 0; at < a; ++at)
        for (int l = 0; l <= lmaxs[at]; ++l) {
            double s = u_norms[at * (max_lmax + 1) + l];
            for (int lm = l * l; lm <= l * l + 2 * l; ++lm) lin[lm + at * L] = s;
        }
    size_t lda = (size_t)L * a;
    for (int col = 0; col < G; ++col)
        for (size_t row = 0; row < lda; ++row)
            C[col * lda + row] = B[col * lda + row] * lin[row];
    free(lin);
}

int orc_create_hs2_v2(int L, int a, int G, cplx* A, cplx* B, cplx* H,
                      cplx* S, const int32_t* lmaxs, int max_lmax,
                      const double* sqrt_u_norms, const int32_t* t_size,
                      const cplx* T_all, const int64_t* t_off, cplx* W)
{
    if (!g_blas.h) return -1;
    const cplx c0 = 0.0, c1 = 1.0, cp5 = 0.5;
    const double dONE = 1.0, dZERO = 0.0;
    int ldA = L * a, ldB = L * a, ldH = G, ldS = G, ldTmp = L;
    int La = L * a;
    cplx* Tmp = (cplx*)malloc(sizeof(cplx) * (size_t)L * (G > 0 ? G : 1));

    /* S = A^H A, :158 */
    g_blas.zherk("L", "C", &G, &La, &dONE, A, &ldA, &dZERO, S, &ldS);
    /* W = sqrt(N) B, :167 */
    orc_apply_u_norm_s(L, a, G, B, W, lmaxs, max_lmax, sqrt_u_norms);
    /* S += W^H W, :177 */
    g_blas.zherk("L", "C", &G, &La, &dONE, W, &ldB, &dONE, S, &ldS);

    /* per-atom X build, :193-231 */
    cplx* X = W;
    int ldX = ldA;
    int X_height = 0;
    for (int atom = 0; atom < a; ++atom) {
        memset(Tmp, 0, sizeof(cplx) * (size_t)L * G);
        int n = t_size[atom];
        const cplx* Tab = T_all + t_off[3 * atom + 1];
        const cplx* Tbb = T_all + t_off[3 * atom + 2];
        cplx* B_atom = B + (size_t)atom * L;
        cplx* A_atom = A + (size_t)atom * L;
        g_blas.zhemm("L", "L", &n, &G, &cp5, Tbb, &n, B_atom, &ldB, &c0, Tmp, &ldTmp);
        g_blas.zgemm("C", "N", &n, &G, &n, &c1, Tab, &n, A_atom, &ldA, &c1, Tmp, &ldTmp);
        int Tmp_size = n;
        g_blas.zlacpy("A", &Tmp_size, &G, Tmp, &ldTmp, X + X_height, &ldX);
        if (X_height != atom * L) {
            g_blas.zlacpy("A", &Tmp_size, &G, B_atom, &ldB, Tmp, &ldTmp);
            g_blas.zlacpy("A", &Tmp_size, &G, Tmp, &ldTmp, B + X_height, &ldB);
        }
        X_height += Tmp_size;
    }
    /* H = B^H X + X^H B, :300 */
    double zero = 0.0;
    g_blas.zher2k("L", "C", &G, &X_height, &c1, B, &ldB, X, &ldX, &zero, H, &ldH);

    /* per-atom T_aa A into B's storage, :321-349 */
    X = B;
    int X_nhpd_height = 0;
    for (int atom = 0; atom < a; ++atom) {
        int n = t_size[atom];
        const cplx* Taa = T_all + t_off[3 * atom + 0];
        cplx* A_atom = A + (size_t)atom * L;
        int idx = X_nhpd_height;
        cplx* A_idx = A + idx;
        if (A_idx != A_atom) {
            g_blas.zlacpy("A", &n, &G, A_atom, &ldA, Tmp, &ldTmp);
            g_blas.zlacpy("A", &n, &G, Tmp, &ldTmp, A_idx, &ldA);
        }
        g_blas.zhemm("L", "L", &n, &G, &c1, Taa, &n, A_idx, &ldA, &c0, X + idx, &ldA);
        X_nhpd_height += n;
    }
    /* H += A^H X (lower only), :362 via zgemmt (wrappers.cpp:204) */
    cplx cbeta = 1.0;
    g_blas.zgemmt("L", "C", "N", &G, &X_nhpd_height, &c1, A, &ldA, X, &ldX,
                  &cbeta, H, &ldH);
    free(Tmp);
    (void)ldS;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Element-wise sums, h_and_s.cpp:84-162 (naive overlap / Hamiltonian). */
/* Full (not only lower) H and S, column-major G x G.  T blocks as in  */
/* orc_create_hs2_v2 but used as full matrices; T_ba = T_ab^H.         */
/* dE_u_norm[atom*(max_lmax+1)+l] is ||u_dot||^2 (not its root).        */
/* ------------------------------------------------------------------ */
void orc_naive_hs(int L, int a, int G, const cplx* A, const cplx* B,
                  int max_lmax, const double* dE_u_norm,
                  const int32_t* t_size, const cplx* T_all,
                  const int64_t* t_off, cplx* H, cplx* S)
{
    for (int g = 0; g < G; ++g) {
        for (int gp = 0; gp < G; ++gp) {
            cplx s = 0.0;
            for (int at = 0; at < a; ++at)
                for (int l = 0; l <= max_lmax; ++l) {
                    double nrm = dE_u_norm[at * (max_lmax + 1) + l];
                    for (int m = -l; m <= l; ++m) {
                        int lm = l * (l + 1) + m;
                        size_t i1 = lm + (size_t)L * (at + (size_t)a * gp);
                        size_t i2 = lm + (size_t)L * (at + (size_t)a * g);
                        s += conj(A[i1]) * A[i2] + conj(B[i1]) * B[i2] * nrm;
                    }
                }
            S[gp + (size_t)G * g] = s;
        }
    }
    for (int g = 0; g < G; ++g) {
        for (int gp = 0; gp < G; ++gp) {
            cplx h = 0.0;
            for (int at = 0; at < a; ++at) {
                int n = t_size[at];
                const cplx* Taa = T_all + t_off[3 * at + 0];
                const cplx* Tab = T_all + t_off[3 * at + 1];
                const cplx* Tbb = T_all + t_off[3 * at + 2];
                for (int lm = 0; lm < n; ++lm)
                    for (int lmp = 0; lmp < n; ++lmp) {
                        size_t ig = (size_t)L * (at + (size_t)a * g);
                        size_t igp = (size_t)L * (at + (size_t)a * gp);
                        cplx tba = conj(Tab[lm + (size_t)n * lmp]); /* T_ba(lmp,lm) */
                        h += conj(A[lmp + ig]) * Taa[lmp + (size_t)n * lm] * A[lm + igp];
                        h += conj(B[lmp + ig]) * Tbb[lmp + (size_t)n * lm] * B[lm + igp];
                        h += conj(A[lmp + ig]) * Tab[lmp + (size_t)n * lm] * B[lm + igp];
                        h += conj(B[lmp + ig]) * tba * A[lm + igp];
                    }
            }
            H[g + (size_t)G * gp] = h;
        }
    }
}
