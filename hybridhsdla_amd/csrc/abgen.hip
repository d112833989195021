// A/B matching-coefficient generator on the GPU (FP64).
//
// Computes, per plane wave K = k + g and atom a (tensor_AB_match.cpp:63-129):
//   A_lm = (4 pi / sqrt(Omega)) e^{i 2 pi K.tau_a} (1/W_l) conj(Y_lm(R_a^T K)) a_l
//   B_lm = (same)                                                     b_l
//   a_l  =  dE_u |K| j_l'(R_MT|K|) - dE_u_dr j_l(R_MT|K|)
//   b_l  = -u    |K| j_l'(R_MT|K|) + u_dr    j_l(R_MT|K|)
// with j_l by the Fleur sphbes/dsphbs scheme (series for x < 1e-3, upward
// recurrence for lmax < x, Miller downward recurrence otherwise) and Y_lm by
// the ylm4 upward recurrence in l (Condon-Shortley phase), all in FP64.
//
// One workgroup owns one G column and walks the atoms in chunks:
//   phase 1a  one thread per atom: direction, Bessel values, phase, radial
//             factors (common * a_l, common * b_l) -> LDS
//   phase 1b  one thread per (atom, m >= 0): Legendre recurrence in l,
//             conj(Y_{l,+-m}) -> LDS
//   phase 2   all threads: one output element each, contiguous rows, so the
//             column is written with full 64-lane coalescing.
// Output layouts:
//   compact (create_hs): Z[col*ldz + 2*off_a + r], r < nlm_a -> A_a(lm=r),
//             r >= nlm_a -> B_a(lm=r-nlm_a); Zs the same with B rows scaled
//             by N_l = ||u_dot_l||^2 (S = Z^H Zs).  Folded (fold != 0): Z alone
//             with B rows scaled by sqrt(N_l) (S = Z^H Z).  Rows >= 2K and
//             columns >= ng are written as zeros (tile padding).
//   ref:      A[col*L*natoms + a*L + lm], B likewise, rows lm >= nlm_a zero
//             (the reference's arma::cx_cube(L, natoms, G) layout).
#include "hsgen_internal.h"

namespace hsgen {


// Fleur sphbes (sphbes.f:7-76) + dsphbs (dsphbs.f:11-47) for one x; writes
// fj[l], dfj[l] for l = 0..lmax into strided LDS slots.  lmax >= 1, x >= 0.
__device__ inline void bessel_lds(int lmax, double x, double* fj, double* dfj)
{
    const double small = 1.0e-3;
    double xx = x * x;
    if (x >= small) {
        double xinv = 1.0 / x;
        double s, c;
        sincos(x, &s, &c);
        double f0 = s * xinv;
        double f1 = (f0 - c) * xinv;
        fj[0] = f0;
        fj[1] = f1;
        if ((double)lmax < x) {
            double fm2 = f0, fm1 = f1;
            for (int l = 2; l <= lmax; ++l) {
                double f = (2 * l - 1) * xinv * fm1 - fm2;
                fj[l] = f;
                fm2 = fm1;
                fm1 = f;
            }
        } else if (lmax >= 2) {
            // Miller: aux(n) = 0, aux(n-1) = 1, aux(i-1) = (2i+1)/x aux(i) - aux(i+1).
            // n - 1 >= lmax + 9, so every aux(l), l <= lmax, comes out of the loop;
            // they are parked in fj[] and rescaled afterwards.
            int n = (int)(lmax + 10 + x);
            double anext = 0.0, acur = 1.0;
            for (int i = n - 1; i >= 1; --i) {
                double aprev = (2 * i + 1) * xinv * acur - anext;   // aux(i-1)
                if (i - 1 <= lmax) fj[i - 1] = aprev;
                anext = acur;
                acur = aprev;
            }
            // normalise with j0 or j1, whichever is larger in magnitude
            int mn = (fabs(f0) < fabs(f1)) ? 1 : 0;
            double quot = mn ? f1 / fj[1] : f0 / fj[0];
            fj[0] = f0;
            fj[1] = f1;
            for (int l = 2; l <= lmax; ++l) fj[l] = fj[l] * quot;
        }
        // x >= 1e-3 > 1e-4: dsphbs recurrence branch
        dfj[0] = -fj[1];
        for (int l = 1; l <= lmax; ++l) dfj[l] = fj[l - 1] - (l + 1) * fj[l] / x;
    } else {
        fj[0] = 1.0 - xx / 6.0 * (1.0 - xx / 20.0 * (1.0 - xx / 42.0));
        fj[1] = (x / 3.0) * (1.0 - xx / 10.0 * (1.0 - xx / 28.0));
        double fac = xx / 15.0;
        for (int l = 2; l <= lmax; ++l) {
            fj[l] = fac * (1.0 - xx / (4 * l + 6));
            fac = x * fac / (2 * l + 3);
        }
        dfj[0] = -fj[1];
        if (x < 1.0e-4) {
            double x2 = 0.5 * x * x;
            double f = 1.0 / 3.0;
            for (int l = 1; l <= lmax; ++l) {
                dfj[l] = f * (l - x2 * (l + 2) / (2 * l + 3));
                f = x * f / (2 * l + 3);
            }
        } else {
            for (int l = 1; l <= lmax; ++l) dfj[l] = fj[l - 1] - (l + 1) * fj[l] / x;
        }
    }
}

template <bool kCompact>
__global__ __launch_bounds__(256) void k_abgen(AbgenParams p)
{
    extern __shared__ double smem[];
    const int col = blockIdx.x;
    const int tid = threadIdx.x;
    const int nl = p.nl;
    const int L = p.L;
    const int chunk = p.chunk;
    // LDS carve-up (doubles)
    double* s_rf = smem;                               // [chunk][nl][4]  fa.re fa.im fb.re fb.im
    double* s_dir = s_rf + chunk * nl * 4;             // [chunk][4]      cth sth cph sph
    double2* s_yl = (double2*)(s_dir + chunk * 4);     // [chunk][L]      conj(Y_lm)
    double* s_bes = (double*)(s_yl + chunk * L);       // [chunk][2*nl]   fj, dfj scratch
    int* s_off = (int*)(s_bes + chunk * 2 * nl);       // [chunk+1]       row prefix (compact)

    if (kCompact && col >= p.ng) {
        // padding column: zeros
        double2 z = make_double2(0.0, 0.0);
        double2* zc = p.Z + (int64_t)col * p.ldz;
        for (int64_t r = tid; r < p.ldz; r += blockDim.x) zc[r] = z;
        if (!p.fold) {
            double2* zsc = p.Zs + (int64_t)col * p.ldz;
            for (int64_t r = tid; r < p.ldz; r += blockDim.x) zsc[r] = z;
        }
        return;
    }

    // K = k + g (reduced), |K| = || bmat K ||  (types_helper.cpp:33)
    double K0 = p.kpt[0] + p.g[3 * col + 0];
    double K1 = p.kpt[1] + p.g[3 * col + 1];
    double K2 = p.kpt[2] + p.g[3 * col + 2];
    double c0 = p.bmat[0] * K0 + p.bmat[3] * K1 + p.bmat[6] * K2;
    double c1 = p.bmat[1] * K0 + p.bmat[4] * K1 + p.bmat[7] * K2;
    double c2 = p.bmat[2] * K0 + p.bmat[5] * K1 + p.bmat[8] * K2;
    const double Knorm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);

    for (int a0 = 0; a0 < p.natoms; a0 += chunk) {
        const int na = min(chunk, p.natoms - a0);
        // ---- phase 1a: per-atom scalars ----
        if (tid < na) {
            const AtomDev& at = p.atoms[a0 + tid];
            const double* R = at.rot;
            double kb0 = R[0] * K0 + R[1] * K1 + R[2] * K2;   // R^T K
            double kb1 = R[3] * K0 + R[4] * K1 + R[5] * K2;
            double kb2 = R[6] * K0 + R[7] * K1 + R[8] * K2;
            double nrm = sqrt(kb0 * kb0 + kb1 * kb1 + kb2 * kb2);
            if (nrm != 0.0) { kb0 /= nrm; kb1 /= nrm; kb2 /= nrm; }
            // ylm4 angles (ylm4.f:170-191)
            const double small = 1.0e-12;
            double xy = kb0 * kb0 + kb1 * kb1;
            double r = sqrt(xy + kb2 * kb2);
            double rxy = sqrt(xy);
            double cth, sth, cph, sph;
            if (r > small) { cth = kb2 / r; sth = rxy / r; } else { sth = 0.0; cth = 1.0; }
            if (rxy > small) { cph = kb0 / rxy; sph = kb1 / rxy; } else { cph = 1.0; sph = 0.0; }
            double* dir = s_dir + tid * 4;
            dir[0] = cth; dir[1] = sth; dir[2] = cph; dir[3] = sph;
            // phase e^{i 2 pi K.tau}
            double kp = K0 * at.pos[0] + K1 * at.pos[1] + K2 * at.pos[2];
            double ps, pc;
            sincos(2.0 * kPi * kp, &ps, &pc);
            // Bessel with the global max_lmax, as construct_AB does
            double* fj = s_bes + tid * 2 * nl;
            double* dfj = fj + nl;
            bessel_lds(p.max_lmax, at.rmt * Knorm, fj, dfj);
            const RadDev* rd = p.rad + (int64_t)(a0 + tid) * nl;
            double* rf = s_rf + tid * nl * 4;
            for (int l = 0; l <= at.lmax; ++l) {
                RadDev m = rd[l];
                double Au = m.dE_u * Knorm * dfj[l] - m.dE_u_dr * fj[l];
                double Bu = -m.u * Knorm * dfj[l] + m.u_dr * fj[l];
                double cr = p.vol_prefac * pc * m.inv_wronk;
                double ci = p.vol_prefac * ps * m.inv_wronk;
                // folded operands: the B rows' factor carries sqrt(||u_dot_l||^2)
                const double bs = (kCompact && p.fold) ? m.sqrt_un : 1.0;
                rf[4 * l + 0] = cr * Au;
                rf[4 * l + 1] = ci * Au;
                rf[4 * l + 2] = cr * Bu * bs;
                rf[4 * l + 3] = ci * Bu * bs;
            }
        }
        if (kCompact && tid <= na) {
            // rows of the chunk's atoms from their global offsets, one lane each
            const int o = a0 + tid < p.natoms ? p.atoms[a0 + tid].off : p.kdim;
            s_off[tid] = 2 * (o - p.atoms[a0].off);
        }
        __syncthreads();
        // ---- phase 1b: conj(Y_lm) by recurrence in l for each (atom, m >= 0) ----
        for (int q = tid; q < na * nl; q += blockDim.x) {
            const int t = q / nl;
            const int m = q - t * nl;
            const int lmax = p.atoms[a0 + t].lmax;
            if (m > lmax) continue;
            const double* dir = s_dir + t * 4;
            const double cth = dir[0], sth = dir[1], cph = dir[2], sph = dir[3];
            // cos(m phi), sin(m phi) by the Chebyshev recurrence (ylm4.f:208-217)
            double cm = 1.0, sm = 0.0;
            if (m >= 1) {
                double cmm1 = 1.0, smm1 = 0.0;   // m-2 terms
                cm = cph; sm = sph;
                double cph2 = cph + cph;
                for (int mm = 2; mm <= m; ++mm) {
                    double cn = cph2 * cm - cmm1;
                    double sn = cph2 * sm - smm1;
                    cmm1 = cm; smm1 = sm;
                    cm = cn; sm = sn;
                }
            }
            // p(m,m) (ylm4.f:194-206)
            double fac = 1.0;
            for (int mm = 0; mm < m; ++mm) { fac = -(mm + mm - 1) * fac; fac = fac * sth; }
            fac = -(m + m - 1) * fac;
            double2* yl = s_yl + t * L;
            double pl2 = 0.0, pl1 = fac;      // p(l-2,m), p(l-1,m)
            for (int l = m; l <= lmax; ++l) {
                double pv;
                if (l == m) pv = fac;
                else if (l == m + 1) pv = (m + m + 1) * cth * fac;
                else pv = ((l + l - 1) * cth * pl1 - (l + m - 1) * pl2) / (l - m);
                if (l > m) { pl2 = pl1; pl1 = pv; }
                const int lm0 = l * (l + 1);
                if (m == 0) {
                    // Y_l0 real; conj is itself
                    yl[lm0] = make_double2(p.ynorm[lm0] * pv, 0.0);
                } else {
                    double yr = pv * cm, yi = pv * sm;        // ylms
                    double np = p.ynorm[lm0 + m], nm = p.ynorm[lm0 - m];
                    // Y_{l,m} = np*ylms -> conj = np*(yr, -yi)
                    yl[lm0 + m] = make_double2(np * yr, -np * yi);
                    // Y_{l,-m} = conj(ylms)*nm -> conj = nm*(yr, yi)
                    yl[lm0 - m] = make_double2(yr * nm, yi * nm);
                }
            }
        }
        __syncthreads();
        // ---- phase 2: coalesced output ----
        if (kCompact) {
            const int nrows = s_off[na];
            const int64_t base = (int64_t)col * p.ldz + 2 * (int64_t)p.atoms[a0].off;
            for (int q = tid; q < nrows; q += blockDim.x) {
                int t = 0;
                while (t + 1 < na && s_off[t + 1] <= q) ++t;
                const int nlm = p.atoms[a0 + t].nlm;
                int r = q - s_off[t];
                const bool isB = r >= nlm;
                const int lm = isB ? r - nlm : r;
                const int l = (int)sqrt((double)lm + 0.5);
                const double2 y = s_yl[t * L + lm];
                const double* rf = s_rf + (t * nl + l) * 4 + (isB ? 2 : 0);
                double2 v = make_double2(y.x * rf[0] - y.y * rf[1], y.x * rf[1] + y.y * rf[0]);
                p.Z[base + q] = v;
                if (!p.fold) {
                    double sc = isB ? p.rad[(int64_t)(a0 + t) * nl + l].dE_u_norm : 1.0;
                    p.Zs[base + q] = make_double2(v.x * sc, v.y * sc);
                }
            }
        } else {
            const int64_t ld = (int64_t)L * p.natoms;
            for (int q = tid; q < na * L; q += blockDim.x) {
                const int t = q / L;
                const int lm = q - t * L;
                const int nlm = p.atoms[a0 + t].nlm;
                const int64_t idx = (int64_t)col * ld + (int64_t)(a0 + t) * L + lm;
                if (lm < nlm) {
                    const int l = (int)sqrt((double)lm + 0.5);
                    const double2 y = s_yl[t * L + lm];
                    const double* rf = s_rf + (t * nl + l) * 4;
                    p.A[idx] = make_double2(y.x * rf[0] - y.y * rf[1], y.x * rf[1] + y.y * rf[0]);
                    p.B[idx] = make_double2(y.x * rf[2] - y.y * rf[3], y.x * rf[3] + y.y * rf[2]);
                } else {
                    p.A[idx] = make_double2(0.0, 0.0);
                    p.B[idx] = make_double2(0.0, 0.0);
                }
            }
        }
        __syncthreads();
    }
    if (kCompact) {
        const int64_t r0 = 2 * (int64_t)p.kdim;
        double2 z = make_double2(0.0, 0.0);
        for (int64_t r = r0 + tid; r < p.ldz; r += blockDim.x) {
            p.Z[(int64_t)col * p.ldz + r] = z;
            if (!p.fold) p.Zs[(int64_t)col * p.ldz + r] = z;
        }
    }
}

// conj(Y_lm) for l <= lmax at one m >= 0 of the direction (cth, sth, cph,
// sph), by the ylm4 recurrences (ylm4.f:194-217); yl indexed by lm = l(l+1)+m.
__device__ inline void ylm_column(int m, int lmax, double cth, double sth, double cph, double sph,
                                  const double* __restrict__ ynorm, double2* yl)
{
    // cos(m phi), sin(m phi) by the Chebyshev recurrence (ylm4.f:208-217)
    double cm = 1.0, sm = 0.0;
    if (m >= 1) {
        double cmm1 = 1.0, smm1 = 0.0;   // m-2 terms
        cm = cph;
        sm = sph;
        const double cph2 = cph + cph;
        for (int mm = 2; mm <= m; ++mm) {
            const double cn = cph2 * cm - cmm1;
            const double sn = cph2 * sm - smm1;
            cmm1 = cm;
            smm1 = sm;
            cm = cn;
            sm = sn;
        }
    }
    // p(m,m) (ylm4.f:194-206)
    double fac = 1.0;
    for (int mm = 0; mm < m; ++mm) {
        fac = -(mm + mm - 1) * fac;
        fac = fac * sth;
    }
    fac = -(m + m - 1) * fac;
    double pl2 = 0.0, pl1 = fac;      // p(l-2,m), p(l-1,m)
    for (int l = m; l <= lmax; ++l) {
        double pv;
        if (l == m) pv = fac;
        else if (l == m + 1) pv = (m + m + 1) * cth * fac;
        else pv = ((l + l - 1) * cth * pl1 - (l + m - 1) * pl2) / (l - m);
        if (l > m) {
            pl2 = pl1;
            pl1 = pv;
        }
        const int lm0 = l * (l + 1);
        if (m == 0) {
            yl[lm0] = make_double2(ynorm[lm0] * pv, 0.0);
        } else {
            const double yr = pv * cm, yi = pv * sm;
            const double np = ynorm[lm0 + m], nm = ynorm[lm0 - m];
            yl[lm0 + m] = make_double2(np * yr, -np * yi);   // conj(Y_{l,m})
            yl[lm0 - m] = make_double2(yr * nm, yi * nm);    // conj(Y_{l,-m})
        }
    }
}

// l of lm = l(l+1)+m, |m| <= l
__device__ inline int l_of_lm(int lm)
{
    int l = (int)sqrtf((float)lm + 0.5f);
    if ((l + 1) * (l + 1) <= lm) ++l;
    if (l * l > lm) --l;
    return l;
}

// Class-shared generator: Y_lm depends on the atom only through its
// rotation, the Bessel values only through its rmt (and the global
// max_lmax).  Per G column they are computed once per rotation class and
// once per rmt class; each chunk of atoms then needs only its phases and
// radial factors before the coalesced stores.  Same arithmetic per element
// as k_abgen.
template <bool kCompact>
__global__ __launch_bounds__(256) void k_abgen_cls(AbgenParams p)
{
    extern __shared__ double smem[];
    const int col = blockIdx.x;
    const int tid = threadIdx.x;
    const int nl = p.nl;
    const int L = p.L;
    const int chunk = p.chunk;
    double2* s_y = (double2*)smem;                       // [nrot][L]    conj(Y_lm)
    double* s_bes = (double*)(s_y + p.nrot * L);         // [nrmt][2 nl] fj, dfj
    double* s_dir = s_bes + p.nrmt * 2 * nl;             // [nrot][4]    cth sth cph sph
    double* s_rf = s_dir + p.nrot * 4;                   // [chunk][nl][4]
    double* s_un = s_rf + chunk * nl * 4;                // [chunk][nl]  B-row scale (Zs, or Z folded)
    double* s_ph = s_un + chunk * nl;                    // [chunk][2]   cos, sin of the phase 2 pi K.tau
    int* s_off = (int*)(s_ph + chunk * 2);               // [chunk+1]    row prefix (compact)
    int* s_rc = s_off + chunk + 1;                       // [chunk]      rotation class
    unsigned char* s_l = (unsigned char*)(s_rc + chunk); // [L]          l of lm

    if (kCompact && col >= p.ng) {
        const double2 z = make_double2(0.0, 0.0);
        double2* zc = p.Z + (int64_t)col * p.ldz;
        for (int64_t r = tid; r < p.ldz; r += blockDim.x) zc[r] = z;
        if (!p.fold) {
            double2* zsc = p.Zs + (int64_t)col * p.ldz;
            for (int64_t r = tid; r < p.ldz; r += blockDim.x) zsc[r] = z;
        }
        return;
    }
    const double K0 = p.kpt[0] + p.g[3 * col + 0];
    const double K1 = p.kpt[1] + p.g[3 * col + 1];
    const double K2 = p.kpt[2] + p.g[3 * col + 2];
    const double c0 = p.bmat[0] * K0 + p.bmat[3] * K1 + p.bmat[6] * K2;
    const double c1 = p.bmat[1] * K0 + p.bmat[4] * K1 + p.bmat[7] * K2;
    const double c2 = p.bmat[2] * K0 + p.bmat[5] * K1 + p.bmat[8] * K2;
    const double Knorm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);

    for (int lm = tid; lm < L; lm += blockDim.x) s_l[lm] = (unsigned char)l_of_lm(lm);
    // per class: Bessel values (threads 0..nrmt-1), directions (64..64+nrot-1)
    if (tid < p.nrmt) {
        double* fj = s_bes + tid * 2 * nl;
        bessel_lds(p.max_lmax, p.atoms[p.rmt_rep[tid]].rmt * Knorm, fj, fj + nl);
    } else if (tid >= 64 && tid - 64 < p.nrot) {
        const int c = tid - 64;
        const double* R = p.atoms[p.rot_rep[c]].rot;
        double kb0 = R[0] * K0 + R[1] * K1 + R[2] * K2;   // R^T K
        double kb1 = R[3] * K0 + R[4] * K1 + R[5] * K2;
        double kb2 = R[6] * K0 + R[7] * K1 + R[8] * K2;
        const double nrm = sqrt(kb0 * kb0 + kb1 * kb1 + kb2 * kb2);
        if (nrm != 0.0) {
            kb0 /= nrm;
            kb1 /= nrm;
            kb2 /= nrm;
        }
        // ylm4 angles (ylm4.f:170-191)
        const double small = 1.0e-12;
        const double xy = kb0 * kb0 + kb1 * kb1;
        const double r = sqrt(xy + kb2 * kb2);
        const double rxy = sqrt(xy);
        double cth, sth, cph, sph;
        if (r > small) {
            cth = kb2 / r;
            sth = rxy / r;
        } else {
            sth = 0.0;
            cth = 1.0;
        }
        if (rxy > small) {
            cph = kb0 / rxy;
            sph = kb1 / rxy;
        } else {
            cph = 1.0;
            sph = 0.0;
        }
        double* dir = s_dir + c * 4;
        dir[0] = cth;
        dir[1] = sth;
        dir[2] = cph;
        dir[3] = sph;
    }
    __syncthreads();
    // conj(Y_lm) per (rotation class, m >= 0), up to max_lmax
    for (int q = tid; q < p.nrot * nl; q += blockDim.x) {
        const int c = q / nl;
        const int m = q - c * nl;
        const double* dir = s_dir + c * 4;
        ylm_column(m, p.max_lmax, dir[0], dir[1], dir[2], dir[3], p.ynorm, s_y + c * L);
    }

    for (int a0 = 0; a0 < p.natoms; a0 += chunk) {
        const int na = min(chunk, p.natoms - a0);
        // per atom, one lane each: the phase e^{i 2 pi K.tau}, the rotation
        // class and the atom's first row in the chunk (compact layout)
        if (tid < na) {
            const AtomDev& at = p.atoms[a0 + tid];
            const double kp = K0 * at.pos[0] + K1 * at.pos[1] + K2 * at.pos[2];
            sincos(2.0 * kPi * kp, &s_ph[2 * tid + 1], &s_ph[2 * tid]);
            s_rc[tid] = at.cls & 0xffff;
        }
        if (kCompact && tid <= na) {
            const int o = a0 + tid < p.natoms ? p.atoms[a0 + tid].off : p.kdim;
            s_off[tid] = 2 * (o - p.atoms[a0].off);
        }
        __syncthreads();
        // radial factors (common * a_l, common * b_l) per (atom, l)
        for (int q = tid; q < na * nl; q += blockDim.x) {
            const int t = q / nl;
            const int l = q - t * nl;
            const AtomDev& at = p.atoms[a0 + t];
            if (l > at.lmax) continue;
            const double pc = s_ph[2 * t], ps = s_ph[2 * t + 1];
            const double* fj = s_bes + (at.cls >> 16) * 2 * nl;
            const double* dfj = fj + nl;
            const RadDev m = p.rad[(int64_t)(a0 + t) * nl + l];
            const double Au = m.dE_u * Knorm * dfj[l] - m.dE_u_dr * fj[l];
            const double Bu = -m.u * Knorm * dfj[l] + m.u_dr * fj[l];
            const double cr = p.vol_prefac * pc * m.inv_wronk;
            const double ci = p.vol_prefac * ps * m.inv_wronk;
            // folded operands: the B rows' factor carries sqrt(||u_dot_l||^2)
            const double bs = (kCompact && p.fold) ? m.sqrt_un : 1.0;
            double* rf = s_rf + (t * nl + l) * 4;
            rf[0] = cr * Au;
            rf[1] = ci * Au;
            rf[2] = cr * Bu * bs;
            rf[3] = ci * Bu * bs;
            s_un[t * nl + l] = m.dE_u_norm;
        }
        __syncthreads();
        if (kCompact) {
            const int nrows = s_off[na];
            const int64_t base = (int64_t)col * p.ldz + 2 * (int64_t)p.atoms[a0].off;
            // the atom t of row q and its first row lo, last row + 1 hi,
            // rotation class rc, held in registers: q only grows, so t does,
            // and they are re-read only when q passes hi
            int t = 0, lo = 0, hi = s_off[1], rc = s_rc[0];
            for (int q = tid; q < nrows; q += blockDim.x) {
                if (q >= hi) {
                    do {
                        ++t;
                        lo = hi;
                        hi = s_off[t + 1];
                    } while (q >= hi);
                    rc = s_rc[t];
                }
                const int nlm2 = (hi - lo) >> 1;
                const int r = q - lo;
                const bool isB = r >= nlm2;
                const int lm = isB ? r - nlm2 : r;
                const int l = s_l[lm];
                const double2 y = s_y[rc * L + lm];
                const double* rf = s_rf + (t * nl + l) * 4 + (isB ? 2 : 0);
                const double2 v = make_double2(y.x * rf[0] - y.y * rf[1], y.x * rf[1] + y.y * rf[0]);
                p.Z[base + q] = v;
                if (!p.fold) {
                    const double sc = isB ? s_un[t * nl + l] : 1.0;
                    p.Zs[base + q] = make_double2(v.x * sc, v.y * sc);
                }
            }
        } else {
            const int64_t ld = (int64_t)L * p.natoms;
            for (int q = tid; q < na * L; q += blockDim.x) {
                const int t = q / L;
                const int lm = q - t * L;
                const int64_t idx = (int64_t)col * ld + (int64_t)(a0 + t) * L + lm;
                const int l = s_l[lm];
                if (l <= p.atoms[a0 + t].lmax) {
                    const double2 y = s_y[s_rc[t] * L + lm];
                    const double* rf = s_rf + (t * nl + l) * 4;
                    p.A[idx] = make_double2(y.x * rf[0] - y.y * rf[1], y.x * rf[1] + y.y * rf[0]);
                    p.B[idx] = make_double2(y.x * rf[2] - y.y * rf[3], y.x * rf[3] + y.y * rf[2]);
                } else {
                    p.A[idx] = make_double2(0.0, 0.0);
                    p.B[idx] = make_double2(0.0, 0.0);
                }
            }
        }
        __syncthreads();
    }
    if (kCompact) {
        const int64_t r0 = 2 * (int64_t)p.kdim;
        const double2 z = make_double2(0.0, 0.0);
        for (int64_t r = r0 + tid; r < p.ldz; r += blockDim.x) {
            p.Z[(int64_t)col * p.ldz + r] = z;
            if (!p.fold) p.Zs[(int64_t)col * p.ldz + r] = z;
        }
    }
}

size_t abgen_cls_lds_bytes(int chunk, int nl, int L, int nrot, int nrmt)
{
    return 16 * (size_t)nrot * L +
           sizeof(double) * ((size_t)nrmt * 2 * nl + (size_t)nrot * 4 + (size_t)chunk * nl * 5 + (size_t)chunk * 2) +
           sizeof(int) * (size_t)(2 * chunk + 1) + (size_t)L;
}

size_t abgen_lds_bytes(int chunk, int nl, int L)
{
    return sizeof(double) * ((size_t)chunk * nl * 4 + (size_t)chunk * 4 + (size_t)chunk * 2 * L +
                              (size_t)chunk * 2 * nl) +
           sizeof(int) * (size_t)(chunk + 1);
}

hipError_t launch_abgen(const AbgenParams& p, bool compact, hipStream_t st)
{
    dim3 grid(p.ncols), block(256);
    if (p.nrot > 0) {
        const size_t lds = abgen_cls_lds_bytes(p.chunk, p.nl, p.L, p.nrot, p.nrmt);
        if (compact)
            hipLaunchKernelGGL(k_abgen_cls<true>, grid, block, lds, st, p);
        else
            hipLaunchKernelGGL(k_abgen_cls<false>, grid, block, lds, st, p);
        return hipGetLastError();
    }
    size_t lds = abgen_lds_bytes(p.chunk, p.nl, p.L);
    if (compact)
        hipLaunchKernelGGL(k_abgen<true>, grid, block, lds, st, p);
    else
        hipLaunchKernelGGL(k_abgen<false>, grid, block, lds, st, p);
    return hipGetLastError();
}


__global__ __launch_bounds__(256) void k_pack_ref(PackParams p)
{
    const int col = blockIdx.x;
    const double2 zero = make_double2(0.0, 0.0);
    double2* zc = p.Z + (int64_t)col * p.ldz;
    double2* zsc = p.fold ? nullptr : p.Zs + (int64_t)col * p.ldz;
    if (col >= p.ng) {
        for (int64_t r = threadIdx.x; r < p.ldz; r += blockDim.x) {
            zc[r] = zero;
            if (zsc) zsc[r] = zero;
        }
        return;
    }
    // walk atoms: each atom contributes 2*nlm rows
    for (int a = 0; a < p.natoms; ++a) {
        const AtomDev at = p.atoms[a];
        const int nlm = at.nlm;
        for (int r = threadIdx.x; r < 2 * nlm; r += blockDim.x) {
            const bool isB = r >= nlm;
            const int lm = isB ? r - nlm : r;
            const int64_t src = (int64_t)col * p.ldab + (int64_t)a * p.L + lm;
            double2 v = isB ? p.B[src] : p.A[src];
            double sc = 1.0;
            if (isB) {
                const int l = (int)sqrt((double)lm + 0.5);
                sc = p.scale[a * p.nl + l];
            }
            if (zsc) {
                zc[2 * (int64_t)at.off + r] = v;
                zsc[2 * (int64_t)at.off + r] = make_double2(v.x * sc, v.y * sc);
            } else {
                zc[2 * (int64_t)at.off + r] = make_double2(v.x * sc, v.y * sc);
            }
        }
    }
    for (int64_t r = 2 * (int64_t)p.kdim + threadIdx.x; r < p.ldz; r += blockDim.x) {
        zc[r] = zero;
        if (zsc) zsc[r] = zero;
    }
}

hipError_t launch_pack_ref(const PackParams& p, int ncols, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_ref, dim3(ncols), dim3(256), 0, st, p);
    return hipGetLastError();
}

}  // namespace hsgen
