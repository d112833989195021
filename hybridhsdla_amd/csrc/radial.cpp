// Host-side system setup: the scalar-relativistic radial solver that turns a
// spherical muffin-tin potential and an energy parameter into the matching
// data {u, u', u_dot, u_dot', <u_dot|u_dot>} consumed by the A/B generator.
//
// It is a sequential ODE integration over a few hundred mesh points per
// (atom type, l): microseconds of host work once per system, with no
// parallelism a GPU could use, so it runs on the host like the rest of the
// set-up.
//
// Units and equations follow Fleur (Hartree atomic units, c = 137.0359895,
// fleur constants.f:15-25).  With x = ln r and the mesh r_i = r_0 e^(i h),
// the large (p = r u) and small (q) components satisfy
//     dp/dx = M q + p,          M = 2 r + (r E - vr) / c^2 = 2 r (1 + (E - V) / 2c^2),
//     dq/dx = (l(l+1)/M + vr - r E) p - q,          vr = r V(r),
// integrated outward (radsra.f:25-120): a 4th-order Runge-Kutta start for the
// first five steps, then a 6-step Adams-Bashforth predictor with an
// Adams-Moulton corrector iterated to a relative change below 1e-6.  The
// energy derivative solves the same system with the normalised function as
// source (radsrd.F:30-160, with its RELATIVISTIC_CORRECTIONS factor
// 1 + l(l+1) / (c^2 M^2) on the large-component source), is orthogonalised
// to u and normalised by <u_dot|u_dot>.  Integrals over the mesh follow
// intgr0 (intgr.F:58-116): a power-law start from 0 to r_0, 7-point Lagrange
// weights for the first (n - 1) mod 6 intervals, 7-point closed Newton-Cotes
// panels for the rest.  All quadrature and multistep weights are derived
// here from the Lagrange interpolation polynomials (exact integers over
// their common denominators) rather than tabulated.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hsgen.h"

namespace {

constexpr double kCLight = 137.0359895;   // speed of light in Hartree atomic units

// Integer weights W_i and denominator D with
//   integral_a^b of the interpolation polynomial through f(t_0..t_{m-1}) dt
//     = sum_i W_i f(t_i) / D,
// for integer nodes t_i and integer limits a, b (exact: the polynomial
// coefficients are computed in long double over small integers and each
// weight is rounded to the nearest multiple of 1/D).
void lagrange_weights(const int* t, int m, int a, int b, long long den, double* w)
{
    for (int i = 0; i < m; ++i) {
        // numerator polynomial prod_{k != i} (x - t_k), coefficients c[0..m-1]
        long double c[16] = {1.0L};
        int deg = 0;
        long double dnm = 1.0L;
        for (int k = 0; k < m; ++k) {
            if (k == i) continue;
            for (int d = deg + 1; d > 0; --d) c[d] = c[d - 1] - (long double)t[k] * c[d];
            c[0] = -(long double)t[k] * c[0];
            ++deg;
            dnm *= (long double)(t[i] - t[k]);
        }
        long double s = 0.0L;
        for (int d = 0; d <= deg; ++d) {
            long double pb = 1.0L, pa = 1.0L;
            for (int e = 0; e <= d; ++e) {
                pb *= b;
                pa *= a;
            }
            s += c[d] * (pb - pa) / (long double)(d + 1);
        }
        w[i] = (double)llroundl(s / dnm * (long double)den);
    }
}

struct Weights {
    double lag[5][7];     // first-interval Lagrange weights over 60480 (interval j: [j, j+1])
    double nc7[7];        // closed 7-point Newton-Cotes over 140
    double ab[6];         // Adams-Bashforth, nodes 0, -1, ..., -5 over [0, 1], over 1440
    double am[6];         // Adams-Moulton, nodes 1, 0, ..., -4 over [0, 1], over 1440
    Weights()
    {
        const int seven[7] = {0, 1, 2, 3, 4, 5, 6};
        for (int j = 0; j < 5; ++j) lagrange_weights(seven, 7, j, j + 1, 60480, lag[j]);
        lagrange_weights(seven, 7, 0, 6, 140, nc7);
        const int abn[6] = {0, -1, -2, -3, -4, -5};
        lagrange_weights(abn, 6, 0, 1, 1440, ab);
        const int amn[6] = {1, 0, -1, -2, -3, -4};
        lagrange_weights(amn, 6, 0, 1, 1440, am);
    }
};

const Weights& weights()
{
    static const Weights w;
    return w;
}

// Definite integral of y over the whole logarithmic mesh r_i = r0 e^(i h),
// i = 0..n-1 (intgr.F intgr0).  y is integrated in x = ln r, i.e. the
// integrand is r y(r).
double mesh_integral(const double* y, double r0, double h, int n)
{
    const Weights& W = weights();
    double z = 0.0;
    // 0 .. r0 from the leading power law y ~ r^(alpha - 1)
    if (y[0] * y[1] > 0.0) {
        const double alpha = 1.0 + std::log(y[1] / y[0]) / h;
        if (alpha > 0.0) z = r0 * y[0] / alpha;
    }
    const int nsteps = (n - 1) / 6;
    int i0 = n - 6 * nsteps - 1;          // first point of the Newton-Cotes panels (0-based)
    const double dr = std::exp(h);
    double r[7];
    r[0] = r0;
    for (int i = 1; i < 7; ++i) r[i] = dr * r[i - 1];
    if (i0 > 0) {
        double yr[7];
        for (int i = 0; i < 7; ++i) yr[i] = r[i] * y[i];
        for (int j = 0; j < i0; ++j) {
            double d = 0.0;
            for (int i = 0; i < 7; ++i) d += W.lag[j][i] * yr[i];
            z += h * d / 60480.0;
        }
    }
    r[0] = r[i0];
    for (int s = 0; s < nsteps; ++s) {
        for (int i = 1; i < 7; ++i) r[i] = dr * r[i - 1];
        double d = 0.0;
        for (int i = 0; i < 7; ++i) d += (h * W.nc7[i] * r[i] / 140.0) * y[i0 + i];
        z += d;
        i0 += 6;
        r[0] = r[6];
    }
    return z;
}

struct Equation {
    int l;
    double e, h, r0;
    const double* vr;
    int n;
    double cin, cin2, fl1;
};

// Source of the inhomogeneous equation for the energy derivative: the
// normalised radial function (p, q), or none.
struct Source {
    const double* p = nullptr;
    const double* q = nullptr;
};

// Large-component source factor with the relativistic correction.
inline double corr(const Equation& eq, double rm)
{
    return 1.0 + eq.cin2 * eq.fl1 / (rm * rm);
}

// Outward integration.  On entry p[0], q[0], dp[0], dq[0] hold the start
// values; on return p, q, dp, dq (derivatives in x) are filled for all n
// points, and the mesh radius of the last point (by repeated multiplication,
// as the integrator advances) is returned.
double integrate(const Equation& eq, const Source& src, double* p, double* q, double* dp, double* dq, int* nodes)
{
    const Weights& W = weights();
    const double h = eq.h;
    const double dr = std::exp(h);
    const double drh = std::sqrt(dr);
    const double* vr = eq.vr;
    const bool inh = src.p != nullptr;
    double r = eq.r0;
    for (int i = 0; i < 5; ++i) {   // Runge-Kutta start
        const double rh2 = drh * r;
        const double rh = dr * r;
        const double sk1 = h * dp[i];
        const double sl1 = h * dq[i];
        double rve = 0.5 * (vr[i] + vr[i + 1]) - rh2 * eq.e;
        double rm = 2.0 * rh2 - eq.cin2 * rve;
        const double t1 = inh ? 0.5 * rh2 * eq.cin * (src.q[i] + src.q[i + 1]) : 0.0;
        const double t2 = inh ? 0.5 * rh2 * (src.p[i] + src.p[i + 1]) * corr(eq, rm) : 0.0;
        double yn = p[i] + 0.5 * sk1;
        double zn = q[i] + 0.5 * sl1;
        const double sk2 = h * (rm * zn + yn + t1);
        const double sl2 = h * ((eq.fl1 / rm + rve) * yn - zn - t2);
        yn = p[i] + 0.5 * sk2;
        zn = q[i] + 0.5 * sl2;
        const double sk3 = h * (rm * zn + yn + t1);
        const double sl3 = h * ((eq.fl1 / rm + rve) * yn - zn - t2);
        rve = vr[i + 1] - rh * eq.e;
        rm = 2.0 * rh - eq.cin2 * rve;
        const double u1 = inh ? rh * eq.cin * src.q[i + 1] : 0.0;
        const double u2 = inh ? rh * src.p[i + 1] * corr(eq, rm) : 0.0;
        yn = p[i] + sk3;
        zn = q[i] + sl3;
        const double sk4 = h * (rm * zn + yn + u1);
        const double sl4 = h * ((eq.fl1 / rm + rve) * yn - zn - u2);
        p[i + 1] = p[i] + (sk1 + 2.0 * sk2 + 2.0 * sk3 + sk4) / 6.0;
        q[i + 1] = q[i] + (sl1 + 2.0 * sl2 + 2.0 * sl3 + sl4) / 6.0;
        dp[i + 1] = rm * q[i + 1] + p[i + 1] + u1;
        dq[i + 1] = (eq.fl1 / rm + rve) * p[i + 1] - q[i + 1] - u2;
        r = rh;
    }
    int nd = 0;
    const double eps = 1e-6;
    for (int i = 5; i < eq.n - 1; ++i) {   // predictor-corrector
        r = r * dr;
        double sp = 0.0, sq = 0.0;
        for (int k = 0; k < 6; ++k) {
            sp += W.ab[k] * dp[i - k];
            sq += W.ab[k] * dq[i - k];
        }
        double p0 = p[i] + h * sp / 1440.0;
        double q0 = q[i] + h * sq / 1440.0;
        const double rve = vr[i + 1] - r * eq.e;
        const double rm = 2.0 * r - eq.cin2 * rve;
        const double t1 = inh ? eq.cin * r * src.q[i + 1] : 0.0;
        const double t2 = inh ? r * src.p[i + 1] * corr(eq, rm) : 0.0;
        double p1p = rm * q0 + p0 + t1;
        double q1p = (eq.fl1 / rm + rve) * p0 - q0 - t2;
        double p1 = p0, q1 = q0;
        for (int it = 0; it < 5; ++it) {
            double cp = W.am[0] * p1p, cq = W.am[0] * q1p;
            for (int k = 1; k < 6; ++k) {
                cp += W.am[k] * dp[i + 1 - k];
                cq += W.am[k] * dq[i + 1 - k];
            }
            p1 = p[i] + h * cp / 1440.0;
            q1 = q[i] + h * cq / 1440.0;
            p1p = rm * q1 + p1 + t1;
            q1p = (eq.fl1 / rm + rve) * p1 - q1 - t2;
            const double erp = std::fabs(p1 - p0) / (std::fabs(p1) + std::fabs(h * p1p));
            const double erq = std::fabs(q1 - q0) / (std::fabs(q1) + std::fabs(h * p1p));
            if (erp < eps && erq < eps) break;
            p0 = p1;
            q0 = q1;
        }
        p[i + 1] = p1;
        q[i + 1] = q1;
        dp[i + 1] = p1p;
        dq[i + 1] = q1p;
        if ((p[i + 1] < 0.0) != (p[i] < 0.0)) ++nd;
    }
    if (nodes) *nodes = nd;
    return r;
}

struct RadialResult {
    double us, dus, uds, duds, ddn;
};

// u and u_dot at the last mesh point; f, g (n x 2 each, or null) receive the
// normalised (large, small) components of r u and r u_dot.
RadialResult solve(int l, double e, const double* vr, double r0, double h, int n, double* f, double* g)
{
    Equation eq;
    eq.l = l;
    eq.e = e;
    eq.h = h;
    eq.r0 = r0;
    eq.vr = vr;
    eq.n = n;
    eq.cin = 1.0 / kCLight;
    eq.cin2 = eq.cin * eq.cin;
    eq.fl1 = (double)(l * (l + 1));
    std::vector<double> buf(8 * (size_t)n);
    double* p = buf.data();
    double* q = p + n;
    double* pe = q + n;
    double* qe = pe + n;
    double* dp = qe + n;
    double* dq = dp + n;
    double* w = dq + n;          // integrand scratch
    // --- u: regular start p ~ r^(l+1), q / p from the small-r limit
    {
        const double s = std::sqrt(eq.fl1 + 1.0 - eq.cin2 * vr[0] * vr[0]) - 1.0;
        const double rm = 2.0 * r0 + eq.cin2 * (r0 * e - vr[0]);
        p[0] = std::pow(r0, l + 1);
        q[0] = s * p[0] / rm;
        dp[0] = rm * q[0] + p[0];
        dq[0] = (eq.fl1 / rm + vr[0] - r0 * e) * p[0] - q[0];
    }
    const double r = integrate(eq, Source{}, p, q, dp, dq, nullptr);
    for (int i = 0; i < n; ++i) q[i] = eq.cin * q[i];
    for (int i = 0; i < n; ++i) w[i] = p[i] * p[i] + q[i] * q[i];
    const double t = 1.0 / std::sqrt(mesh_integral(w, r0, h, n));
    RadialResult res;
    res.us = t * p[n - 1] / r;
    res.dus = t * (dp[n - 1] - p[n - 1]) / (r * r);
    for (int i = 0; i < n; ++i) {
        p[i] = t * p[i];
        q[i] = t * q[i];
    }
    // --- u_dot: inhomogeneous equation with (p, q) as source, zero start
    {
        const double rm = 2.0 * r0 + eq.cin2 * (r0 * e - vr[0]);
        pe[0] = 0.0;
        qe[0] = 0.0;
        dp[0] = r0 * eq.cin * q[0];
        dq[0] = -r0 * p[0] * corr(eq, rm);
    }
    Source src;
    src.p = p;
    src.q = q;
    (void)integrate(eq, src, pe, qe, dp, dq, nullptr);
    for (int i = 0; i < n; ++i) qe[i] = eq.cin * qe[i];
    for (int i = 0; i < n; ++i) w[i] = p[i] * pe[i] + q[i] * qe[i];
    const double ov = mesh_integral(w, r0, h, n);     // <u|u_dot> before orthogonalisation
    double dud = (dp[n - 1] - pe[n - 1]) / (r * r);
    for (int i = 0; i < n; ++i) {
        pe[i] = pe[i] - ov * p[i];
        qe[i] = qe[i] - ov * q[i];
    }
    res.uds = pe[n - 1] / r;
    res.duds = dud - ov * res.dus;
    for (int i = 0; i < n; ++i) w[i] = pe[i] * pe[i] + qe[i] * qe[i];
    res.ddn = mesh_integral(w, r0, h, n);
    if (f) {
        std::memcpy(f, p, sizeof(double) * n);
        std::memcpy(f + n, q, sizeof(double) * n);
    }
    if (g) {
        std::memcpy(g, pe, sizeof(double) * n);
        std::memcpy(g + n, qe, sizeof(double) * n);
    }
    return res;
}

}  // namespace

extern "C" {

hsgen_status hsgen_radial_solve(int32_t l, double E, double dx, double rmt, int32_t n, const double* vr,
                                double* match, double* f, double* g)
{
    if (!vr || !match || n < 7 || l < 0 || !(dx > 0.0) || !(rmt > 0.0)) return HSGEN_ERR_ARG;
    // first mesh point as the reference's Fleur wrapper sets it (libflerp.f90:69)
    const double r0 = rmt * std::exp(dx * (double)(1 - n));
    const RadialResult r = solve(l, E, vr, r0, dx, n, f, g);
    match[0] = r.us;
    match[1] = r.dus;
    match[2] = r.uds;
    match[3] = r.duds;
    match[4] = r.ddn;
    for (int i = 0; i < 5; ++i)
        if (!std::isfinite(match[i])) return HSGEN_ERR_ARG;
    return HSGEN_OK;
}

hsgen_status hsgen_quadrature_weights(int32_t which, double* w)
{
    if (!w) return HSGEN_ERR_ARG;
    const Weights& W = weights();
    switch (which) {
    case 0: std::memcpy(w, W.lag, sizeof W.lag); return HSGEN_OK;
    case 1: std::memcpy(w, W.nc7, sizeof W.nc7); return HSGEN_OK;
    case 2: std::memcpy(w, W.ab, sizeof W.ab); return HSGEN_OK;
    case 3: std::memcpy(w, W.am, sizeof W.am); return HSGEN_OK;
    }
    return HSGEN_ERR_ARG;
}

}  // extern "C"
