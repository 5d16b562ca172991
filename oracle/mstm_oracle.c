/*
 * mstm_oracle.c -- CPU FP64 restatement of the MSTM solution hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.  Each routine cites the reference routine it restates
 * (paths relative to the MSTM tree, december2023/).
 *
 * Independent pins (tests/test_oracle_*.py):
 *   rotcoef        -> Jacobi closed form + orthogonality (scipy), ref dumps
 *   cg             -> exact rational Clebsch-Gordan (python fractions)
 *   riccati        -> scipy.special.spherical_jn / spherical_yn
 *   axial/pair     -> dense gentranmatrix (independent formula) and ref dumps
 *   mie            -> textbook TE/TM Lorenz-Mie (scipy) via the L/R transform
 *   cbicg          -> dense LU solve of (I - T H) a = T f; ref dumps
 *   fft path       -> brute-force node sums (this file) vs reference dumps
 */
#include "mstm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CI (I)
/* the reference's pi is a widened real(4) literal: mstm-modules-33.f90:18 */
static const double PI_R4 = (double)3.14159265358979323846f;

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int iabs_(int a) { return a < 0 ? -a : a; }
static inline double sgn1(int k) { return (k & 1) ? -1.0 : 1.0; } /* (-1)^k */
static inline double fnr(int n) { return sqrt((double)n); }

/* ------------------------------------------------------------------ */
/* sqrt of binomial coefficient, numconstants::init bcof                */
/* (mstm-modules-33.f90:29-48): bcof(n,l) = ((n+l)!/(n! l!))^(1/2)     */
/* ------------------------------------------------------------------ */
double orc_bcof(int n, int l)
{
    long double c = 1.0L;
    int a = imin(n, l), b = imax(n, l);
    for (int k = 1; k <= a; k++) c = c * (long double)(b + k) / (long double)k;
    return (double)sqrtl(c);
}

/* Clebsch-Gordan <j1 m1 j2 m2 | J m1+m2>, Racah's closed sum in long
 * double.  Stands in for vcfunc/vcfuncuprec (mstm-modules-33.f90:950-1056),
 * which return the same Condon-Shortley coefficients by recurrence.     */
static long double lfact_tab[512];
static int lfact_init = 0;
static long double lfac(int n)
{
    if (!lfact_init) {
        lfact_tab[0] = 1.0L;
        for (int k = 1; k < 512; k++) lfact_tab[k] = lfact_tab[k - 1] * (long double)k;
        lfact_init = 1;
    }
    return lfact_tab[n];
}
double orc_cg(int j1, int m1, int j2, int m2, int J)
{
    int M = m1 + m2;
    if (iabs_(m1) > j1 || iabs_(m2) > j2 || iabs_(M) > J) return 0.0;
    if (J < iabs_(j1 - j2) || J > j1 + j2) return 0.0;
    long double pre = (2.0L * J + 1.0L) * lfac(J + j1 - j2) * lfac(J - j1 + j2) * lfac(j1 + j2 - J) /
                      lfac(j1 + j2 + J + 1);
    pre *= lfac(J + M) * lfac(J - M) * lfac(j1 - m1) * lfac(j1 + m1) * lfac(j2 - m2) * lfac(j2 + m2);
    int kmin = imax(0, imax(j2 - J - m1, j1 - J + m2));
    int kmax = imin(j1 + j2 - J, imin(j1 - m1, j2 + m2));
    long double s = 0.0L;
    for (int k = kmin; k <= kmax; k++) {
        long double den = lfac(k) * lfac(j1 + j2 - J - k) * lfac(j1 - m1 - k) * lfac(j2 + m2 - k) *
                          lfac(J - j2 + m1 + k) * lfac(J - j1 - m2 + k);
        s += ((k & 1) ? -1.0L : 1.0L) / den;
    }
    return (double)(sqrtl(pre) * s);
}

/* ------------------------------------------------------------------ */
/* Wigner d: rotcoef (mstm-modules-33.f90:1098-1175).                   */
/* dc[(m+kmax) + (2kmax+1)*(n(n+1)+k)] = d^n_{mk}(beta), cbe = cos beta */
/* ------------------------------------------------------------------ */
void orc_rotcoef(double cbe, int kmax, int nmax, double *dc)
{
    int ld = 2 * kmax + 1;
#define DC(m, kn) dc[(m) + kmax + ld * (kn)]
    memset(dc, 0, sizeof(double) * ld * (nmax + 1) * (nmax + 1));
    double sbe = (fabs(cbe) >= 1.0) ? 0.0 : sqrt(fabs((1.0 + cbe) * (1.0 - cbe)));
    double cbe2 = 0.5 * (1.0 + cbe), sbe2 = 0.5 * (1.0 - cbe);
    double *dk0 = (double *)calloc(2 * nmax + 3, sizeof(double)) + nmax + 1;
    double *dk01 = (double *)calloc(2 * nmax + 3, sizeof(double)) + nmax + 1;
    int sgn = 1;
    double sben = 1.0;
    dk0[0] = 1.0;
    dk01[0] = 0.0;
    DC(0, 0) = 1.0;
    for (int n = 1; n <= nmax; n++) {
        int knmax = imin(n, kmax), nn1 = n * (n + 1);
        sgn = -sgn;
        sben = sben * sbe / 2.0;
        if (sben < 1e-30) sben = 0.0;
        dk0[n] = sgn * sben * orc_bcof(n, n);
        dk0[-n] = sgn * dk0[n];
        dk01[n] = 0.0;
        dk01[-n] = 0.0;
        DC(0, nn1 + n) = dk0[n];
        DC(0, nn1 - n) = dk0[-n];
        for (int k = -n + 1; k <= n - 1; k++) {
            double dkt = dk01[k];
            dk01[k] = dk0[k];
            dk0[k] = (cbe * (double)(2 * n - 1) * dk01[k] - fnr(n - k - 1) * fnr(n + k - 1) * dkt) /
                     (fnr(n + k) * fnr(n - k));
            DC(0, nn1 + k) = dk0[k];
        }
        int im = 1;
        for (int m = 1; m <= knmax; m++) {
            im = -im;
            double fmn = 1.0 / fnr(n - m + 1) / fnr(n + m);
            double dkm0 = 0.0, dkm1;
            for (int k = -n; k <= n; k++) {
                int kn = nn1 + k;
                dkm1 = dkm0;
                dkm0 = DC(m - 1, kn);
                double dkn1 = (k == n) ? 0.0 : DC(m - 1, kn + 1);
                DC(m, kn) = (fnr(n + k) * fnr(n - k + 1) * cbe2 * dkm1 -
                             fnr(n - k) * fnr(n + k + 1) * sbe2 * dkn1 - (double)k * sbe * DC(m - 1, kn)) *
                            fmn;
                DC(-m, nn1 - k) = DC(m, kn) * sgn1(k) * im;
            }
        }
    }
    free(dk0 - nmax - 1);
    free(dk01 - nmax - 1);
#undef DC
}

/* normalizedlegendre (mstm-modules-33.f90:1063-1085):
 * dc[(m+mmax) + (2mmax+1)*n], n = 0..nmax                                */
void orc_normalizedlegendre(double cbe, int mmax, int nmax, double *dc)
{
    int ld = 2 * mmax + 1;
#define DL(m, n) dc[(m) + mmax + ld * (n)]
    double sbe = sqrt((1.0 + cbe) * (1.0 - cbe));
    memset(dc, 0, sizeof(double) * ld * (nmax + 1));
    for (int m = 0; m <= mmax; m++) {
        DL(m, m) = sgn1(m) * pow(0.5 * sbe, m) * orc_bcof(m, m);
        if (m == nmax) break;
        DL(m, m + 1) = fnr(2 * m + 1) * cbe * DL(m, m);
        for (int n = m + 1; n <= nmax - 1; n++)
            DL(m, n + 1) = (-fnr(n - m) * fnr(n + m) * DL(m, n - 1) + (double)(2 * n + 1) * cbe * DL(m, n)) /
                           (fnr(n + 1 - m) * fnr(n + 1 + m));
    }
    for (int m = 1; m <= mmax; m++)
        for (int n = m; n <= nmax; n++) DL(-m, n) = sgn1(m) * DL(m, n);
#undef DL
}

/* ------------------------------------------------------------------ */
/* Riccati-Bessel / Hankel functions of real argument.                  */
/* psi_n = x j_n(x) by Miller's downward recurrence (x < n) or upward    */
/* recurrence (x >= n); chi_n = x y_n(x) upward; xi_n = psi_n + i chi_n. */
/* Stand-in for ricbessel/richankel/cricbessel/crichankel               */
/* (mstm-modules-33.f90:162-301), which use the same definitions.       */
/* ------------------------------------------------------------------ */
void orc_riccati_psi(int n, double x, double *psi)
{
    psi[0] = sin(x);
    if (n == 0) return;
    if (x >= (double)n) {
        psi[1] = sin(x) / x - cos(x);
        for (int k = 1; k < n; k++) psi[k + 1] = (double)(2 * k + 1) / x * psi[k] - psi[k - 1];
        return;
    }
    int ns = n + 30 + (int)(2.0 * x);
    double fp1 = 0.0, f = 1e-300, fm1;
    double *tmp = malloc(sizeof(double) * (n + 1));
    double scale = 1.0;
    for (int k = ns; k >= 1; k--) {
        fm1 = (double)(2 * k + 1) / x * f - fp1; /* f_{k-1} */
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e250) { /* rescale everything stored so far */
            f *= 1e-250;
            fp1 *= 1e-250;
            for (int j = k; j <= n; j++) tmp[j] *= 1e-250;
        }
        if (k - 1 <= n) tmp[k - 1] = f;
    }
    /* normalise with the larger of psi_0, psi_1 */
    double p0 = sin(x), p1 = sin(x) / x - cos(x);
    if (fabs(p0) > fabs(p1)) scale = p0 / tmp[0];
    else scale = p1 / tmp[1];
    for (int k = 0; k <= n; k++) psi[k] = tmp[k] * scale;
    (void)fm1;
    free(tmp);
}

void orc_riccati_xi(int n, double x, zc *xi)
{
    double *psi = malloc(sizeof(double) * (n + 2));
    orc_riccati_psi(n, x, psi);
    double c0 = -cos(x), c1 = -cos(x) / x - sin(x);
    xi[0] = psi[0] + CI * c0;
    if (n >= 1) xi[1] = psi[1] + CI * c1;
    for (int k = 1; k < n; k++) {
        double c2 = (double)(2 * k + 1) / x * c1 - c0;
        xi[k + 1] = psi[k + 1] + CI * c2;
        c0 = c1;
        c1 = c2;
    }
    free(psi);
}

/* logarithmic derivative D_k(z) = psi_k'(z)/psi_k(z), k=0..n, complex z,
 * downward recurrence.  mieoa uses psi_{n-1}/psi_n - n/z, which is D_n.  */
void orc_logderiv(int n, double zr, double zi, zc *d)
{
    zc z = zr + CI * zi;
    int ns = imax(n, (int)cabs(z)) + 30;
    zc dk = 0.0;
    for (int k = ns; k >= 1; k--) {
        zc kz = (double)k / z;
        zc dkm1 = kz - 1.0 / (dk + kz);
        if (k <= n) d[k] = dk;
        dk = dkm1;
    }
    d[0] = dk;
}

/* ------------------------------------------------------------------ */
/* axial translation: axialtrancoefinit + axialtrancoefrecurrence       */
/* (mstm-modules-33.f90:1566-1722), packed by m-blocks 0,-1,1,-2,2,...  */
/* (moffset :1966-1978, atcdim :1956-1962).                             */
/* ------------------------------------------------------------------ */
int orc_atcdim(int ntot, int ltot)
{
    int s = 0;
    for (int ma = 0; ma <= imin(ntot, ltot); ma++) {
        int m1 = imax(1, ma);
        int nb = 2 * (ntot - m1 + 1) * (ltot - m1 + 1);
        s += (ma == 0) ? nb : 2 * nb;
    }
    return s;
}
int orc_moffset(int m, int ntot, int ltot)
{
    int off = 0;
    for (int ma = 0; ma <= imin(ntot, ltot); ma++) {
        int m1 = imax(1, ma);
        int nb = 2 * (ntot - m1 + 1) * (ltot - m1 + 1);
        if (ma == 0) {
            if (m == 0) return off;
            off += nb;
        } else {
            if (m == -ma) return off;
            off += nb;
            if (m == ma) return off;
            off += nb;
        }
    }
    return -1;
}

/* constants of the diamond recurrence (axialtrancoefinit :1680-1722) */
static double fnm1c(int m, int l)
{
    if (l < imax(1, iabs_(m))) return 0.0;
    return fnr(l - 1) * fnr(l + 1) * fnr(l - m) * fnr(l + m) / fnr(2 * l - 1) / fnr(2 * l + 1) / (double)l;
}
static double fnc(int m, int l)
{
    if (l < imax(1, iabs_(m))) return 0.0;
    return (double)m / (double)l / (double)(l + 1);
}
static double fnp1c(int m, int l)
{
    if (l < imax(1, iabs_(m))) return 0.0;
    return fnr(l) * fnr(l + 2) * fnr(l + 1 - m) * fnr(l + 1 + m) / fnr(2 * l + 1) / fnr(2 * l + 3) /
           (double)(l + 1);
}
/* vcc_const(n, l(l+1)+m, w), cached for n,l < VCMAX */
#define VCMAX 24
static double vcc_tab[VCMAX][VCMAX][2 * VCMAX + 1][2 * VCMAX + 1];
static unsigned char vcc_set[VCMAX][VCMAX][2 * VCMAX + 1][2 * VCMAX + 1];
static double vccc_eval(int n, int l, int m, int w);
static double vccc(int n, int l, int m, int w)
{
    if (n < VCMAX && l < VCMAX) {
        if (!vcc_set[n][l][m + VCMAX][w]) {
            vcc_tab[n][l][m + VCMAX][w] = vccc_eval(n, l, m, w);
            vcc_set[n][l][m + VCMAX][w] = 1;
        }
        return vcc_tab[n][l][m + VCMAX][w];
    }
    return vccc_eval(n, l, m, w);
}
static double vccc_eval(int n, int l, int m, int w)
{
    double c1 = fnr(2 * n + 1) * fnr(2 * l + 1);
    double c2 = -c1 * sgn1(iabs_(m));
    double vc1 = orc_cg(n, -m, l, m, w);
    double vc2 = orc_cg(n, -1, l, 1, w);
    int k = (n - l + w) & 3;
    double s = (k == 0 || k == 1) ? 1.0 : -1.0; /* Re(i^k)+Im(i^k) */
    return c2 * vc1 * vc2 * s;
}

void orc_axialtrancoef(int itype, double r, const zc *ri, int nmax, int lmax, zc *ac)
{
    int nlmin = imin(nmax, lmax);
    int dim = orc_atcdim(nmax, lmax);
    zc *act = calloc((size_t)(nmax + 1) * (lmax + 1) * 2, sizeof(zc));
#define ACT(n, l, p) act[(n) + (nmax + 1) * ((l) + (lmax + 1) * (p))]
    if (r <= 1e-12) {
        memset(ac, 0, sizeof(zc) * dim);
        if (itype != 1) { free(act); return; }
        int iadd = 0;
        for (int ma = 0; ma <= nlmin; ma++) {
            int m1 = imax(1, ma);
            for (int m = -ma; m <= ma; m += 2 * m1) {
                for (int p = 0; p < 2; p++)
                    for (int l = m1; l <= lmax; l++)
                        for (int n = m1; n <= nmax; n++) ac[iadd++] = (n == l) ? 1.0 : 0.0;
            }
        }
        free(act);
        return;
    }
    int nw = nmax + lmax;
    zc *xi = malloc(sizeof(zc) * (nw + 1) * 2);
    for (int p = 0; p < 2; p++) {
        zc z = r * ri[p];
        double zr = creal(z);
        if (fabs(cimag(z)) > 0.0) {
            fprintf(stderr, "oracle: complex medium index unsupported\n");
            abort();
        }
        if (itype == 1) {
            double *psi = malloc(sizeof(double) * (nw + 1));
            orc_riccati_psi(nw, zr, psi);
            for (int w = 0; w <= nw; w++) xi[w + (nw + 1) * p] = psi[w] / zr;
            free(psi);
        } else {
            orc_riccati_xi(nw, zr, xi + (nw + 1) * p);
            for (int w = 0; w <= nw; w++) xi[w + (nw + 1) * p] /= zr;
        }
    }
    int iadd0 = 0;
    for (int ma = 0; ma <= nlmin; ma++) {
        int m1 = imax(1, ma);
        for (int m = -ma; m <= ma; m += 2 * m1) {
            /* seed row n = m1 (all l) and column l = lmax (n > m1) */
            for (int pass = 0; pass < 2; pass++) {
                int cnt = pass == 0 ? (lmax - m1 + 1) : (nmax - m1);
                for (int q = 0; q < cnt; q++) {
                    int n = pass == 0 ? m1 : m1 + 1 + q;
                    int l = pass == 0 ? m1 + q : lmax;
                    int wmin = iabs_(n - l), wmax = n + l;
                    for (int p = 0; p < 2; p++) {
                        zc se = 0.0, so = 0.0;
                        for (int w = wmin; w <= wmax; w += 2) se += vccc(n, l, m, w) * xi[w + (nw + 1) * p];
                        for (int w = wmin + 1; w <= wmax - 1; w += 2)
                            so += vccc(n, l, m, w) * xi[w + (nw + 1) * p];
                        ACT(n, l, p) = (p == 0) ? se + CI * so : se - CI * so;
                    }
                }
            }
            if (m1 < nlmin) {
                for (int n = m1; n <= nmax - 1; n++) {
                    for (int p = 0; p < 2; p++) {
                        double sp = (p == 0) ? 1.0 : -1.0;
                        for (int l = m1; l <= lmax - 1; l++) {
                            zc v = -ACT(n, l + 1, p) * fnp1c(m, l) + sp * (fnc(m, l) - fnc(m, n)) * CI * ACT(n, l, p);
                            if (l >= m1 + 1) v += ACT(n, l - 1, p) * fnm1c(m, l);
                            if (n > m1) v += ACT(n - 1, l, p) * fnm1c(m, n);
                            ACT(n + 1, l, p) = v / fnp1c(m, n);
                        }
                    }
                }
            }
            for (int p = 0; p < 2; p++)
                for (int l = m1; l <= lmax; l++)
                    for (int n = m1; n <= nmax; n++) ac[iadd0++] = ACT(n, l, p);
        }
    }
#undef ACT
    free(xi);
    free(act);
}

/* ------------------------------------------------------------------ */
/* dense translation matrix: gentrancoefconstants + gentranmatrix       */
/* (mstm-modules-33.f90:1724-1888), index model 2.                      */
/* ac[(kla) + nblkt*((mna) + nblks*p)], kla target (l,k), mna source.   */
/* ------------------------------------------------------------------ */
/* 1-based address of (m,n) in a block of order L (amnaddress model 2) */
static int amnaddr(int m, int n, int L) { return (m >= 0) ? (n - 1) * (L + 2) + m + 1 : (-m - 1) * (L + 2) + n + 2; }

void orc_gentranmatrix(int nodr_s, int nodr_t, const double *xp, const zc *ri, int itype, zc *ac)
{
    int nblks = nodr_s * (nodr_s + 2), nblkt = nodr_t * (nodr_t + 2);
    int wmx = nodr_s + nodr_t;
    double r = xp[0] * xp[0] + xp[1] * xp[1] + xp[2] * xp[2];
    memset(ac, 0, sizeof(zc) * nblks * nblkt * 2);
    if (r == 0.0) {
        if (itype == 1) {
            int nb = imin(nblks, nblkt);
            /* the reference sets ac(n,n) for the first min(nblk) addresses */
            for (int q = 0; q < nb; q++) {
                ac[q + nblkt * q] = 1.0;
                ac[q + nblkt * (q + nblks)] = 1.0;
            }
        }
        return;
    }
    r = sqrt(r);
    double ct = xp[2] / r;
    zc ephi = (xp[0] == 0.0 && xp[1] == 0.0) ? 1.0 : (xp[0] + CI * xp[1]) / sqrt(xp[0] * xp[0] + xp[1] * xp[1]);
    zc *ephim = malloc(sizeof(zc) * (2 * wmx + 1));
    ephim[wmx] = 1.0;
    for (int m = 1; m <= wmx; m++) {
        ephim[wmx + m] = ephi * ephim[wmx + m - 1];
        ephim[wmx - m] = conj(ephim[wmx + m]);
    }
    double *ymn = malloc(sizeof(double) * (2 * wmx + 1) * (wmx + 1));
    orc_normalizedlegendre(ct, wmx, wmx, ymn);
    zc *hn = malloc(sizeof(zc) * (wmx + 1) * 2);
    for (int p = 0; p < 2; p++) {
        zc rri = r * ri[p];
        if (itype == 3) {
            zc e = cexp(CI * rri);
            hn[0 + (wmx + 1) * p] = -CI * e / rri;
            if (wmx >= 1) hn[1 + (wmx + 1) * p] = -e * (CI + rri) / rri / rri;
            for (int n = 2; n <= wmx; n++)
                hn[n + (wmx + 1) * p] = (double)(2 * n - 1) / rri * hn[n - 1 + (wmx + 1) * p] - hn[n - 2 + (wmx + 1) * p];
        } else {
            double *psi = malloc(sizeof(double) * (wmx + 1));
            orc_riccati_psi(wmx, creal(rri), psi);
            for (int n = 0; n <= wmx; n++) hn[n + (wmx + 1) * p] = psi[n] / rri;
            free(psi);
        }
    }
    for (int n = 1; n <= nodr_s; n++) {
        for (int m = -n; m <= n; m++) {
            int mna = amnaddr(m, n, nodr_s) - 1;
            for (int l = 1; l <= nodr_t; l++) {
                int wmax = n + l;
                for (int k = -l; k <= l; k++) {
                    int kla = amnaddr(k, l, nodr_t) - 1;
                    int v = m - k;
                    int wmin = imax(iabs_(v), iabs_(n - l));
                    /* tran_coef(kl,mn,w): first slot (l,k) plays the role of
                     * (n,m) in gentrancoefconstants :1740-1756 */
                    zc c = -cpow(CI, l - n) * fnr(2 * l + 1) * fnr(2 * n + 1);
                    double m1m = sgn1(iabs_(k));
                    zc a1 = 0, a2 = 0, b1 = 0, b2 = 0;
                    for (int w = wmax; w >= wmin; w--) {
                        double vc1 = orc_cg(l, -k, n, m, w);
                        double vc2 = orc_cg(l, -1, n, 1, w);
                        zc a = cpow(CI, w) * c * m1m * vc1 * vc2;
                        double tc = ((wmax - w) % 2 == 0) ? creal(a) : cimag(a);
                        double y = ymn[v + wmx + (2 * wmx + 1) * w];
                        if ((wmax - w) % 2 == 0) {
                            a1 += hn[w] * y * tc;
                            a2 += hn[w + wmx + 1] * y * tc;
                        } else {
                            b1 += hn[w] * y * tc;
                            b2 += hn[w + wmx + 1] * y * tc;
                        }
                    }
                    ac[kla + nblkt * mna] = ephim[v + wmx] * (a1 + CI * b1);
                    ac[kla + nblkt * (mna + nblks)] = ephim[v + wmx] * (a2 - CI * b2);
                }
            }
        }
    }
    free(ephim);
    free(ymn);
    free(hn);
}

/* ------------------------------------------------------------------ */
/* shiftcoefficient (mstm-modules-33.f90:6789-6820) with msign=mflip=-1: */
/* a'(m,n) = (-1)^m a(-m,n).  In place allowed.                         */
/* ------------------------------------------------------------------ */
static void mflip(int L, zc *a)
{
    int bl = L * (L + 2);
    for (int p = 0; p < 2; p++)
        for (int n = 1; n <= L; n++)
            for (int m = 1; m <= n; m++) {
                int ip = amnaddr(m, n, L) - 1 + p * bl, in = amnaddr(-m, n, L) - 1 + p * bl;
                zc t = a[ip];
                double s = sgn1(m);
                a[ip] = s * a[in];
                a[in] = s * t;
            }
}

/* ------------------------------------------------------------------ */
/* coefficient_translation (mstm-modules-33.f90:6610-6787).             */
/* Tables are regenerated per call (store_translation_matrix=.false.    */
/* default, :5124).  The tables are built in the orientation of the     */
/* sop=.false. call: nodrs = source of that call, nodrt = its target.   */
/* ------------------------------------------------------------------ */
void orc_pair_translate(int itype, const double *rtran, const zc *ri, int noa, int nog, int sop, int top,
                        int force_rot, const zc *ain, zc *gout)
{
    int nodrs = sop ? nog : noa, nodrt = sop ? noa : nog;
    int nmin = imin(noa, nog), nmax = imax(noa, nog);
    int blka = noa * (noa + 2), blkg = nog * (nog + 2);
    double r2 = rtran[0] * rtran[0] + rtran[1] * rtran[1] + rtran[2] * rtran[2];
    int shift = (sop != top);
    int rot = force_rot || (nmin >= 3); /* translation_switch_order = 3, :5126 */
    zc *a = malloc(sizeof(zc) * 2 * blka);
    zc *g = calloc(2 * blkg, sizeof(zc));
    memcpy(a, ain, sizeof(zc) * 2 * blka);
    if (r2 < 1e-12) {
        /* zero translation: mtransfer for regular waves only (:6691-6696) */
        if (itype == 1) {
            for (int p = 0; p < 2; p++)
                for (int n = 1; n <= nmin; n++)
                    for (int m = -n; m <= n; m++)
                        gout[amnaddr(m, n, nog) - 1 + p * blkg] += a[amnaddr(m, n, noa) - 1 + p * blka];
        }
        free(a);
        free(g);
        return;
    }
    if (shift) mflip(noa, a);
    if (rot) {
        double r = sqrt(r2), ct = rtran[2] / r;
        zc ephi = (rtran[0] == 0.0 && rtran[1] == 0.0)
                      ? 1.0
                      : (rtran[0] + CI * rtran[1]) / sqrt(rtran[0] * rtran[0] + rtran[1] * rtran[1]);
        int kd = 2 * nmin + 1;
        double *rot_mat = malloc(sizeof(double) * kd * (nmax + 1) * (nmax + 1));
        orc_rotcoef(ct, nmin, nmax, rot_mat);
#define RM(k, kn) rot_mat[(k) + nmin + kd * (kn)]
        zc *phi = malloc(sizeof(zc) * (2 * nmax + 1));
        phi[nmax] = 1.0;
        for (int m = 1; m <= nmax; m++) {
            phi[nmax + m] = ephi * phi[nmax + m - 1];
            phi[nmax - m] = conj(phi[nmax + m]);
        }
        int tdim = orc_atcdim(nodrt, nodrs);
        zc *zmat = malloc(sizeof(zc) * tdim);
        orc_axialtrancoef(itype, r, ri, nodrt, nodrs, zmat);
        int im = sop ? -1 : 1;
        /* a_tt(m,n,p): m in -noa..noa */
        int lda = 2 * noa + 1, ldg = 2 * nog + 1;
        zc *att = calloc((size_t)lda * (noa + 1) * 2, sizeof(zc));
        zc *gtt = calloc((size_t)ldg * (nog + 1) * 2, sizeof(zc));
#define ATT(m, n, p) att[(m) + noa + lda * ((n) + (noa + 1) * (p))]
#define GTT(m, n, p) gtt[(m) + nog + ldg * ((n) + (nog + 1) * (p))]
        for (int p = 0; p < 2; p++)
            for (int n = 1; n <= noa; n++)
                for (int m = -n; m <= n; m++) ATT(m, n, p) = a[amnaddr(m, n, noa) - 1 + p * blka] * phi[nmax + im * m];
        /* forward rotation */
        zc *tmp = malloc(sizeof(zc) * (2 * nmax + 1));
        for (int p = 0; p < 2; p++)
            for (int n = 1; n <= noa; n++) {
                int n1 = imin(n, nog);
                for (int k = -n1; k <= n1; k++) {
                    zc s = 0;
                    for (int m = -n; m <= n; m++) s += RM(k, n * (n + 1) + m) * ATT(m, n, p);
                    tmp[k + n1] = s;
                }
                for (int k = -n1; k <= n1; k++) ATT(k, n, p) = tmp[k + n1];
            }
        /* axial translation per m */
        for (int m = -nmin; m <= nmin; m++) {
            int m1 = imax(1, iabs_(m));
            int off = orc_moffset(m, nodrt, nodrs);
            for (int p = 0; p < 2; p++) {
                for (int l = m1; l <= nog; l++) {
                    zc s = 0;
                    for (int n = m1; n <= noa; n++) {
                        zc z;
                        if (!sop) /* atc(l,n): rows target (nog), cols source (noa) */
                            z = zmat[off + (l - m1) + (nog - m1 + 1) * ((n - m1) + (noa - m1 + 1) * p)];
                        else /* atc(n,l): rows source (noa), cols target (nog) */
                            z = zmat[off + (n - m1) + (noa - m1 + 1) * ((l - m1) + (nog - m1 + 1) * p)];
                        s += z * ATT(m, n, p);
                    }
                    GTT(m, l, p) = s;
                }
            }
        }
        /* back rotation + phase */
        for (int p = 0; p < 2; p++)
            for (int n = 1; n <= nog; n++) {
                int n1 = imin(n, noa);
                for (int k = -n; k <= n; k++) {
                    zc s = 0;
                    for (int m = -n1; m <= n1; m++) s += GTT(m, n, p) * RM(m, n * (n + 1) + k);
                    tmp[k + n] = s;
                }
                for (int k = -n; k <= n; k++) g[amnaddr(k, n, nog) - 1 + p * blkg] = tmp[k + n] * phi[nmax - im * k];
            }
#undef ATT
#undef GTT
#undef RM
        free(tmp);
        free(att);
        free(gtt);
        free(zmat);
        free(phi);
        free(rot_mat);
    } else {
        int bs = nodrs * (nodrs + 2), bt = nodrt * (nodrt + 2);
        zc *gm = malloc(sizeof(zc) * bs * bt * 2);
        orc_gentranmatrix(nodrs, nodrt, rtran, ri, itype, gm);
        for (int p = 0; p < 2; p++) {
            for (int q = 0; q < blkg; q++) {
                zc s = 0;
                for (int t = 0; t < blka; t++) {
                    zc z = !sop ? gm[q + bt * (t + bs * p)] : gm[t + bt * (q + bs * p)];
                    s += z * a[t + p * blka];
                }
                g[q + p * blkg] = s;
            }
        }
        free(gm);
    }
    if (shift) mflip(nog, g);
    for (int q = 0; q < 2 * blkg; q++) gout[q] += g[q];
    free(a);
    free(g);
}

/* ------------------------------------------------------------------ */
/* Lorenz-Mie L/R 2x2 coefficients: mieoa (mstm-modules-33.f90:5438-5578)*/
/* for a non-chiral sphere in a real medium.                             */
/* ------------------------------------------------------------------ */
static void mie_core(double x, zc ms, double mm, int nstop, zc *anp, double *qext1, double *qsca_out)
{
    double xm = x * mm;
    zc xs = x * ms;
    double *psi = malloc(sizeof(double) * (nstop + 2));
    zc *xi = malloc(sizeof(zc) * (nstop + 2));
    zc *dn = malloc(sizeof(zc) * (nstop + 2));
    orc_riccati_psi(nstop + 1, xm, psi);
    orc_riccati_xi(nstop + 1, xm, xi);
    orc_logderiv(nstop + 1, creal(xs), cimag(xs), dn);
    double qsca = 0.0;
    zc rii[2][2] = {{mm, ms}, {mm, ms}}; /* rii(s, region): region 1 = medium, 2 = sphere */
    zc rib[2] = {mm, ms};
    for (int n = 1; n <= nstop; n++) {
        double psip = psi[n - 1] - (double)n * psi[n] / xm;
        zc xip = xi[n - 1] - (double)n * xi[n] / xm;
        zc g[2][2], a[2][2];
        for (int s = 0; s < 2; s++) {
            double ss = (s == 0) ? -1.0 : 1.0;
            for (int t = 0; t < 2; t++) {
                double st = (t == 0) ? -1.0 : 1.0;
                zc f = (ss * rib[0] + st * rib[1]) / (rii[s][1] * rii[t][0]);
                g[s][t] = f * (dn[n] * xi[n] - ss * st * xip);
                a[s][t] = f * (dn[n] * psi[n] - ss * st * psip);
            }
        }
        zc det = g[0][0] * g[1][1] - g[1][0] * g[0][1];
        zc gi[2][2] = {{g[1][1] / det, -g[0][1] / det}, {-g[1][0] / det, g[0][0] / det}};
        zc an[2][2];
        for (int s = 0; s < 2; s++)
            for (int t = 0; t < 2; t++) an[s][t] = -(gi[s][0] * a[0][t] + gi[s][1] * a[1][t]);
        if (anp)
            for (int s = 0; s < 2; s++)
                for (int t = 0; t < 2; t++) anp[s + 2 * t + 4 * (n - 1)] = an[s][t];
        double fn1 = 2 * n + 1;
        qext1[n - 1] = 0.0;
        for (int p = 0; p < 2; p++) {
            for (int q = 0; q < 2; q++) qsca += fn1 * cabs(an[p][q]) * cabs(an[p][q]);
            qext1[n - 1] -= fn1 * creal(an[p][p]);
        }
    }
    if (qsca_out) *qsca_out = qsca;
    free(psi);
    free(xi);
    free(dn);
}

int orc_mie_order(double x, double mre, double mim, double mmed, double qeps)
{
    zc m_sphere = mre + CI * mim;
    /* nstop = nint(x + 4 x**(1./3.)) + 5 with a real(4) exponent (:5461) */
    int nstop = (int)lround(x + 4.0 * pow(x, (double)(1.0f / 3.0f))) + 5;
    double *q1 = malloc(sizeof(double) * nstop);
    mie_core(x, m_sphere, mmed, nstop, NULL, q1, NULL);
    double qt = 0.0;
    for (int n = 0; n < nstop; n++) qt += q1[n];
    double q = 0.0;
    int n;
    for (n = 1; n <= nstop; n++) {
        q += q1[n - 1];
        if (fabs(1.0 - q / qt) < qeps || n == nstop) break;
    }
    free(q1);
    return n;
}

void orc_mie_coef(double x, double mre, double mim, double mmed, int nodr, zc *anp, double *qext, double *qsca)
{
    zc m_sphere = mre + CI * mim;
    double *q1 = malloc(sizeof(double) * nodr), qs;
    mie_core(x, m_sphere, mmed, nodr, anp, q1, &qs);
    double qe = 0.0;
    for (int n = 0; n < nodr; n++) qe += q1[n];
    if (qext) *qext = 2.0 / x / x * qe;
    if (qsca) *qsca = 2.0 / x / x * qs;
    free(q1);
}

/* ------------------------------------------------------------------ */
/* cluster state (spheredata + mie module arrays)                        */
/* ------------------------------------------------------------------ */
struct orc_cluster {
    int n, neqns, nmie, switch_order;
    double *pos, *rad;
    zc *ri, rimed;
    int *order, *offset, *mieoff;
    zc *an, *aninv;
    /* fft state */
    int fft_ready, cell_dim[3], node_order, nnb;
    double d_cell, cell_boundary[3];
    int *sphere_node, nbr[27][3];
    int *pairs, npairs;
};

orc_cluster *orc_cluster_new(int n, const double *pos, const double *radius, const zc *ri, double ri_medium,
                             double mie_eps, int switch_order)
{
    orc_cluster *c = calloc(1, sizeof(orc_cluster));
    c->n = n;
    c->switch_order = switch_order;
    c->pos = malloc(sizeof(double) * 3 * n);
    c->rad = malloc(sizeof(double) * n);
    c->ri = malloc(sizeof(zc) * n);
    memcpy(c->pos, pos, sizeof(double) * 3 * n);
    memcpy(c->rad, radius, sizeof(double) * n);
    memcpy(c->ri, ri, sizeof(zc) * n);
    c->rimed = ri_medium;
    c->order = malloc(sizeof(int) * n);
    c->offset = malloc(sizeof(int) * (n + 1));
    c->mieoff = malloc(sizeof(int) * (n + 1));
    int off = 0, moff = 0;
    for (int i = 0; i < n; i++) {
        c->order[i] = orc_mie_order(radius[i], creal(ri[i]), cimag(ri[i]), ri_medium, mie_eps);
        c->offset[i] = off;
        c->mieoff[i] = moff;
        off += 2 * c->order[i] * (c->order[i] + 2);
        moff += 4 * c->order[i];
    }
    c->offset[n] = off;
    c->mieoff[n] = moff;
    c->neqns = off;
    c->nmie = moff;
    c->an = malloc(sizeof(zc) * moff);
    c->aninv = malloc(sizeof(zc) * moff);
    for (int i = 0; i < n; i++) {
        orc_mie_coef(radius[i], creal(ri[i]), cimag(ri[i]), ri_medium, c->order[i], c->an + c->mieoff[i], NULL, NULL);
        for (int k = 0; k < c->order[i]; k++) {
            zc *a = c->an + c->mieoff[i] + 4 * k, *b = c->aninv + c->mieoff[i] + 4 * k;
            zc det = a[0] * a[3] - a[1] * a[2];
            b[0] = a[3] / det;
            b[3] = a[0] / det;
            b[1] = -a[1] / det;
            b[2] = -a[2] / det;
        }
    }
    return c;
}

void orc_cluster_free(orc_cluster *c)
{
    if (!c) return;
    free(c->pos);
    free(c->rad);
    free(c->ri);
    free(c->order);
    free(c->offset);
    free(c->mieoff);
    free(c->an);
    free(c->aninv);
    free(c->sphere_node);
    free(c->pairs);
    free(c);
}

int orc_cluster_neqns(const orc_cluster *c) { return c->neqns; }
void orc_cluster_orders(const orc_cluster *c, int *order, int *offset)
{
    memcpy(order, c->order, sizeof(int) * c->n);
    memcpy(offset, c->offset, sizeof(int) * (c->n + 1));
}
void orc_cluster_mie(const orc_cluster *c, zc *an) { memcpy(an, c->an, sizeof(zc) * c->nmie); }

/* multmiecoeffmult (mstm-modules-33.f90:5633-5751) for spheres without
 * internal expansions; inverse=1 applies an_inv_mie (onemiecoeffmult 'i'). */
void orc_mie_mult(const orc_cluster *c, int idir, int inverse, const zc *ain, zc *aout)
{
    const zc *tab = inverse ? c->aninv : c->an;
    for (int i = 0; i < c->n; i++) {
        int L = c->order[i], bl = L * (L + 2), o = c->offset[i];
        for (int n = 1; n <= L; n++) {
            const zc *t = tab + c->mieoff[i] + 4 * (n - 1);
            for (int m = -n; m <= n; m++) {
                int a0 = o + amnaddr(m, n, L) - 1;
                zc x0 = ain[a0], x1 = ain[a0 + bl];
                if (idir == 1) {
                    aout[a0] = t[0] * x0 + t[2] * x1;
                    aout[a0 + bl] = t[1] * x0 + t[3] * x1;
                } else {
                    aout[a0] = t[0] * x0 + t[1] * x1;
                    aout[a0 + bl] = t[2] * x0 + t[3] * x1;
                }
            }
        }
    }
}

/* external_to_external_expansion (mstm-modules-33.f90:6390-6525) */
void orc_external_to_external(const orc_cluster *c, int nrhs, const int *contran, const zc *ain, zc *gout)
{
    int ne = c->neqns;
    zc ri2[2] = {c->rimed, c->rimed};
    memset(gout, 0, sizeof(zc) * ne * nrhs);
    for (int i = 0; i < c->n - 1; i++)
        for (int j = i + 1; j < c->n; j++) {
            double rt[3] = {c->pos[3 * i] - c->pos[3 * j], c->pos[3 * i + 1] - c->pos[3 * j + 1],
                            c->pos[3 * i + 2] - c->pos[3 * j + 2]};
            int noi = c->order[i], noj = c->order[j];
            int rot = imin(noi, noj) >= c->switch_order;
            for (int r = 0; r < nrhs; r++) {
                orc_pair_translate(3, rt, ri2, noj, noi, 0, contran[r], rot, ain + r * ne + c->offset[j],
                                   gout + r * ne + c->offset[i]);
                orc_pair_translate(3, rt, ri2, noi, noj, 1, contran[r], rot, ain + r * ne + c->offset[i],
                                   gout + r * ne + c->offset[j]);
            }
        }
}

static void flip_all(const orc_cluster *c, zc *v)
{
    for (int i = 0; i < c->n; i++) mflip(c->order[i], v + c->offset[i]);
}

/* sphereinteraction (mstm-scatprops-26.f90:25-180) */
void orc_sphereinteraction(const orc_cluster *c, int nrhs, const int *contran, const int *miemult, int use_fft,
                           const zc *ain, zc *aout)
{
    int ne = c->neqns;
    zc *at = malloc(sizeof(zc) * ne * nrhs), *gt = malloc(sizeof(zc) * ne * nrhs);
    for (int r = 0; r < nrhs; r++) {
        if (contran[r]) {
            for (int q = 0; q < ne; q++) at[q + r * ne] = conj(ain[q + r * ne]);
            if (miemult[r]) orc_mie_mult(c, -1, 0, at + r * ne, at + r * ne);
        } else
            memcpy(at + r * ne, ain + r * ne, sizeof(zc) * ne);
    }
    if (use_fft) orc_fft_external_to_external(c, nrhs, contran, at, gt);
    else orc_external_to_external(c, nrhs, contran, at, gt);
    for (int r = 0; r < nrhs; r++) {
        if (!contran[r]) {
            if (miemult[r]) orc_mie_mult(c, 1, 0, gt + r * ne, aout + r * ne);
            else memcpy(aout + r * ne, gt + r * ne, sizeof(zc) * ne);
        } else
            for (int q = 0; q < ne; q++) aout[q + r * ne] = conj(gt[q + r * ne]);
    }
    free(at);
    free(gt);
}

/* genplanewavecoef (mstm-modules-33.f90:1453-1520, lr_tran) with the
 * zero-boundary phase of layerplanewavecoef (:3901-3965):
 * pmnp(neqns, 2), column = incident polarization.                      */
void orc_planewave(const orc_cluster *c, double alpha, double sinc, int dir, zc *pmnp)
{
    orc_planewave_ex(c, alpha, sinc, dir, NULL, pmnp);
}

/* the same with sphereplanewavecoef's excited_spheres (mstm-scatprops-26.f90:
 * 262-264): spheres with excited[i] == 0 keep zero blocks; NULL = all    */
void orc_planewave_ex(const orc_cluster *c, double alpha, double sinc, int dir, const int *excited, zc *pmnp)
{
    int ne = c->neqns;
    double rim = creal(c->rimed);
    double cb = sqrt((1.0 - sinc / rim) * (1.0 + sinc / rim)) * (3 - 2 * dir);
    int Lmax = 0;
    for (int i = 0; i < c->n; i++) Lmax = imax(Lmax, c->order[i]);
    double *drot = malloc(sizeof(double) * 3 * (Lmax + 1) * (Lmax + 1));
    orc_rotcoef(cb, 1, Lmax, drot);
    memset(pmnp, 0, sizeof(zc) * ne * 2);
    for (int i = 0; i < c->n; i++) {
        if (excited && !excited[i]) continue;
        int L = c->order[i], bl = L * (L + 2), o = c->offset[i];
        const double *x = c->pos + 3 * i;
        zc ph = cexp(CI * (sinc * (cos(alpha) * x[0] + sin(alpha) * x[1]) + rim * cb * x[2]));
        for (int n = 1; n <= L; n++) {
            double fnm = sqrt((double)(2 * n + 1) / 2.0) / 4.0;
            zc cin = 4.0 * cpow(CI, n + 1);
            for (int m = -n; m <= n; m++) {
                int mn = n * (n + 1) + m;
                double dm1 = drot[0 + 3 * mn], dp1 = drot[2 + 3 * mn];
                double t1 = -fnm * (-dm1 + dp1), t2 = -fnm * (dm1 + dp1);
                double tl[2] = {0.5 * (t1 + t2), 0.5 * (t1 - t2)};
                zc ea = cexp(-CI * (double)m * alpha);
                for (int p = 0; p < 2; p++) {
                    double sp = (p == 0) ? 1.0 : -1.0;
                    int a = o + amnaddr(m, n, L) - 1 + p * bl;
                    pmnp[a] = -cin * tl[p] * ea * ph;
                    pmnp[a + ne] = sp * CI * cin * tl[p] * ea * ph;
                }
            }
        }
    }
    free(drot);
}

static zc zdotc(int n, const zc *a, const zc *b)
{
    zc s = 0;
    for (int q = 0; q < n; q++) s += conj(a[q]) * b[q];
    return s;
}

/* cbicg (mstm-solver-8.f90:654-949), serial branch.  Returns the
 * iteration count (niter+1 when it never converged, as Fortran leaves the
 * do-variable).                                                          */
int orc_cbicg(const orc_cluster *c, int use_fft, int niter, double eps, const zc *pnp, zc *anp, double *errmax)
{
    int ne = c->neqns;
    *errmax = 0.0;
    if (creal(zdotc(ne, pnp, pnp)) == 0.0) return 0;
    if (niter == 0) {
        /* (:768-781): anp = the host-sphere part of the product with
         * skip_external_translation, zero for external spheres */
        memset(anp, 0, sizeof(zc) * ne);
        return 0;
    }
    if (niter < 0) {
        /* order of scattering (:785-825): cp = anp; per term cr = T H cp,
         * anp += cr, cp = cr; stop when |cr|^2/|pnp|^2 < eps */
        zc *cp = malloc(sizeof(zc) * ne), *cr = malloc(sizeof(zc) * ne);
        double enorm = creal(zdotc(ne, pnp, pnp));
        int f1[1] = {0}, t1[1] = {1}, iter;
        memcpy(cp, anp, sizeof(zc) * ne);
        for (iter = 1; iter <= -niter; iter++) {
            orc_sphereinteraction(c, 1, f1, t1, use_fft, cp, cr);
            for (int q = 0; q < ne; q++) anp[q] += cr[q];
            memcpy(cp, cr, sizeof(zc) * ne);
            double eerr = creal(zdotc(ne, cr, cr)) / enorm;
            if (eerr < eps) {
                *errmax = eerr;
                break;
            }
        }
        free(cp);
        free(cr);
        return iter;
    }
    zc *cr = calloc(ne, sizeof(zc)), *cp = malloc(sizeof(zc) * ne), *cw = malloc(sizeof(zc) * ne);
    zc *cq = malloc(sizeof(zc) * ne), *ctin = malloc(sizeof(zc) * 2 * ne), *ctout = malloc(sizeof(zc) * 2 * ne);
    double enorm = creal(zdotc(ne, pnp, pnp));
    int f1[1] = {0}, t1[1] = {1};
    orc_sphereinteraction(c, 1, f1, t1, use_fft, anp, cr);
    for (int q = 0; q < ne; q++) {
        cr[q] = pnp[q] - anp[q] + cr[q];
        cq[q] = conj(cr[q]);
        cw[q] = cq[q];
        cp[q] = cr[q];
    }
    zc csk = 0;
    for (int q = 0; q < ne; q++) csk += cr[q] * cr[q];
    int iter = 0;
    if (cabs(csk) == 0.0) goto done;
    int ct[2] = {0, 1}, mm[2] = {1, 1};
    for (iter = 1; iter <= niter; iter++) {
        memcpy(ctin, cp, sizeof(zc) * ne);
        memcpy(ctin + ne, cw, sizeof(zc) * ne);
        orc_sphereinteraction(c, 2, ct, mm, use_fft, ctin, ctout);
        zc *cap = ctout, *caw = ctout + ne;
        for (int q = 0; q < ne; q++) {
            cap[q] = cp[q] - cap[q];
            caw[q] = cw[q] - caw[q];
        }
        zc cak = zdotc(ne, cw, cap);
        if (cabs(cak) == 0.0) goto done;
        cak = csk / cak;
        for (int q = 0; q < ne; q++) {
            anp[q] += cak * cp[q];
            cr[q] -= cak * cap[q];
            cq[q] -= conj(cak) * caw[q];
        }
        zc csk2 = zdotc(ne, cq, cr);
        double eerr = creal(zdotc(ne, cr, cr)) / enorm;
        *errmax = eerr;
        if (eerr < eps || cabs(csk) == 0.0) goto done;
        zc cbk = csk2 / csk;
        csk = csk2;
        for (int q = 0; q < ne; q++) {
            cp[q] = cr[q] + cbk * cp[q];
            cw[q] = cq[q] + conj(cbk) * cw[q];
        }
    }
done:
    free(cr);
    free(cp);
    free(cw);
    free(cq);
    free(ctin);
    free(ctout);
    return iter;
}

/* lrsphereqeff + qefficiencyfactors (mstm-scatprops-26.f90:676-870),
 * npol = 2.  qeff[3*(k + 3*i)] layout = Fortran qeff(3,3,nsphere).     */
void orc_qeff(const orc_cluster *c, const zc *amnp, const zc *pmnp, double *qeff)
{
    int ne = c->neqns;
    zc *g = malloc(sizeof(zc) * ne * 2);
    orc_mie_mult(c, 1, 1, amnp, g);
    orc_mie_mult(c, 1, 1, amnp + ne, g + ne);
    for (int i = 0; i < c->n; i++) {
        int L = c->order[i], bl = L * (L + 2), o = c->offset[i];
        double xs = c->rad[i], rim = creal(c->rimed);
        double xr = xs * rim;
        double *psi = malloc(sizeof(double) * (L + 1));
        zc *xi = malloc(sizeof(zc) * (L + 1));
        orc_riccati_psi(L, xr, psi);
        orc_riccati_xi(L, xr, xi);
        double rib = rim;
        double qe[3] = {0}, qa[3] = {0}, qi[3] = {0};
        for (int n = 1; n <= L; n++) {
            double psip = psi[n - 1] - n * psi[n] / xr;
            zc xip = xi[n - 1] - n * xi[n] / xr;
            zc aam[2][2], ggm[2][2], agm[2][2];
            for (int s = 0; s < 2; s++) {
                double ss = s == 0 ? -1.0 : 1.0;
                for (int t = 0; t < 2; t++) {
                    double st = t == 0 ? -1.0 : 1.0;
                    zc cterm = CI * conj(1.0 / rim) / rim;
                    aam[s][t] = cterm * (xip * conj(xi[n]) - ss * st * xi[n] * conj(xip));
                    ggm[s][t] = cterm * (psip * psi[n] - ss * st * psi[n] * psip);
                    agm[s][t] = cterm * (xip * psi[n] - ss * st * xi[n] * psip);
                }
            }
            for (int m = -n; m <= n; m++) {
                int a0 = o + amnaddr(m, n, L) - 1;
                for (int p1 = 0; p1 < 2; p1++)
                    for (int p2 = 0; p2 < 2; p2++) {
                        int k = (p1 == p2) ? p1 : 2;
                        double cst = (p1 == p2) ? 1.0 : 0.5;
                        for (int s = 0; s < 2; s++)
                            for (int t = 0; t < 2; t++) {
                                zc as = amnp[a0 + s * bl + p1 * ne], at = amnp[a0 + t * bl + p2 * ne];
                                zc gs = g[a0 + s * bl + p1 * ne], gt = g[a0 + t * bl + p2 * ne];
                                zc fs = pmnp[a0 + s * bl + p1 * ne], ft = pmnp[a0 + t * bl + p2 * ne];
                                double sgst = ((s + t) & 1) ? -1.0 : 1.0;
                                qi[k] += creal(cst * ggm[s][t] * fs * conj(rib * ft));
                                qa[k] += creal(cst * (aam[s][t] * as * conj(rib * at) + ggm[s][t] * gs * conj(rib * gt) +
                                                      agm[s][t] * as * conj(gt) * (rib + sgst * rib)));
                                qe[k] += creal(cst * agm[s][t] * as * conj(ft) * (rib + sgst * rib));
                            }
                    }
            }
        }
        double f = 2.0 / xs / xs;
        double qeffi[3][3]; /* (quantity, k) */
        for (int k = 0; k < 3; k++) {
            qeffi[0][k] = qe[k] * f;
            qeffi[1][k] = qa[k] * f;
            qeffi[2][k] = qi[k] * f;
        }
        for (int q = 0; q < 3; q++) {
            qeff[q + 3 * (0 + 3 * i)] = 0.5 * (qeffi[q][0] + qeffi[q][1]);
            qeff[q + 3 * (1 + 3 * i)] = qeffi[q][0];
            qeff[q + 3 * (2 + 3 * i)] = qeffi[q][1];
        }
        free(psi);
        free(xi);
    }
    free(g);
}

/* qtotcalc (mstm-scatprops-26.f90:872-912) with the cross-section radius
 * of main_calling_program (mstm-input-37.f90:1204-1211): (sum a^3) raised
 * to the real(4) literal .333333.  qtot[3*k + q], k = unpol/par/perp.  */
void orc_qtot(const orc_cluster *c, const double *qeff, double *qtot)
{
    double v = 0.0;
    for (int i = 0; i < c->n; i++) v += c->rad[i] * c->rad[i] * c->rad[i];
    double xg = pow(v, (double)0.333333f);
    memset(qtot, 0, sizeof(double) * 9);
    for (int i = 0; i < c->n; i++)
        for (int k = 0; k < 3; k++)
            for (int q = 0; q < 3; q++) qtot[q + 3 * k] += qeff[q + 3 * (k + 3 * i)] * c->rad[i] * c->rad[i] / xg / xg;
    for (int k = 0; k < 3; k++) qtot[2 + 3 * k] = qtot[0 + 3 * k] + qtot[2 + 3 * k] - qtot[1 + 3 * k];
}

/* fixedorsoln (mstm-solver-8.f90:350-569), iterative branch, serial. */
int orc_fixedorsoln(const orc_cluster *c, int use_fft, double alpha, double sinc, int dir, double eps, int niter,
                    zc *amnp, double *qeff, double *maxerr, int *maxiter)
{
    return orc_fixedorsoln_ex(c, use_fft, alpha, sinc, dir, eps, niter, NULL, amnp, qeff, maxerr, maxiter);
}

/* the same with excited_spheres (NULL = all); niter = 0 solves nothing
 * (:513-522), niter < 0 is cbicg's order-of-scattering series        */
int orc_fixedorsoln_ex(const orc_cluster *c, int use_fft, double alpha, double sinc, int dir, double eps, int niter,
                       const int *excited, zc *amnp, double *qeff, double *maxerr, int *maxiter)
{
    int ne = c->neqns, istat = 0;
    zc *pmnp = malloc(sizeof(zc) * 2 * ne), *pan = malloc(sizeof(zc) * ne);
    orc_planewave_ex(c, alpha, sinc, dir, excited, pmnp);
    for (int p = 0; p < 2; p++) {
        /* istat, maxiter, maxerr are reset for each polarization (:471-475),
         * so they report the second solve only */
        istat = 0;
        *maxerr = 0.0;
        *maxiter = 0;
        orc_mie_mult(c, 1, 0, pmnp + p * ne, pan);
        memcpy(amnp + p * ne, pan, sizeof(zc) * ne);
        double serr = 0.0;
        int iter = niter != 0 ? orc_cbicg(c, use_fft, niter, eps, pan, amnp + p * ne, &serr) : 0;
        if (iter > *maxiter) *maxiter = iter;
        if (serr > *maxerr) *maxerr = serr;
        if (iter > niter || serr > eps) istat = 1;
    }
    if (qeff) orc_qeff(c, amnp, pmnp, qeff);
    free(pmnp);
    free(pan);
    return istat;
}

/* ------------------------------------------------------------------ */
/* FFT node grid (fft_translation-5.f90)                                */
/* ------------------------------------------------------------------ */
static int check235(int n)
{
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

/* node_selection (:654-825) for all spheres external (local_host = 0),
 * target box = sphere centre bounds (mstm-input-37.f90:1250-1251), and
 * node_order from mstm-input-37.f90:1256-1260.                          */
int orc_node_selection(orc_cluster *c, double fva, int node_order_in, int neighbor_model, int *cell_dim,
                       double *d_cell, int *node_order)
{
    return orc_node_selection_ex(c, &fva, NULL, NULL, 0, d_cell, node_order_in, neighbor_model, cell_dim,
                                 node_order);
}

/* the same with node_selection's optional arguments: target box (NULL:
 * sphere centre bounds), d_specified (then *d_cell is the cell size in);
 * fva in/out as in the reference (set to the fraction used)            */
int orc_node_selection_ex(orc_cluster *c, double *fva_io, const double *target_min, const double *target_max,
                          int d_specified, double *d_cell, int node_order_in, int neighbor_model, int *cell_dim,
                          int *node_order)
{
    double tmin[3] = {1e10, 1e10, 1e10}, tmax[3] = {-1e10, -1e10, -1e10};
    for (int i = 0; i < c->n; i++)
        for (int m = 0; m < 3; m++) {
            tmin[m] = fmin(tmin[m], c->pos[3 * i + m]);
            tmax[m] = fmax(tmax[m], c->pos[3 * i + m]);
        }
    for (int m = 0; m < 3; m++) {
        if (target_min) tmin[m] = target_min[m];
        if (target_max) tmax[m] = target_max[m];
    }
    const double fva = *fva_io;
    double amean = 0.0, svol = 0.0;
    for (int i = 0; i < c->n; i++) {
        amean += c->rad[i];
        svol += c->rad[i] * c->rad[i] * c->rad[i];
    }
    svol = pow(svol, 1.0 / 3.0);
    amean /= c->n;
    double tvol = 1.0;
    for (int m = 0; m < 3; m++) tvol *= fmax(tmax[m] - tmin[m], amean);
    double d = *d_cell;
    if (!d_specified) {
        double fv;
        if (fva <= 0.0) {
            fv = 4.0 * PI_R4 / 3.0 * svol * svol * svol / tvol;
            fv = fmin(fv, 1.0);
            fv = fmax(fv, 0.02);
        } else
            fv = fva;
        *fva_io = fv;
        d = pow(4.0 * PI_R4 / 3.0 / fv / (double)c->n, 1.0 / 3.0) * svol;
    }
    for (int m = 0; m < 3; m++) {
        int nd = (int)ceil((tmax[m] - tmin[m]) / d);
        if (nd < 1) return -1; /* correctn235 would not end on 0 */
        while (!check235(nd)) nd++;
        c->cell_dim[m] = nd;
    }
    d = 0.0;
    for (int m = 0; m < 3; m++) d = fmax(d, (tmax[m] - tmin[m]) / (double)c->cell_dim[m]);
    c->d_cell = d;
    for (int m = 0; m < 3; m++) c->cell_boundary[m] = tmin[m];
    c->node_order = (node_order_in <= 0) ? -node_order_in + (int)ceil(d) : node_order_in;
    free(c->sphere_node);
    c->sphere_node = malloc(sizeof(int) * 3 * c->n);
    for (int i = 0; i < c->n; i++)
        for (int m = 0; m < 3; m++) {
            int nd = (int)floor((c->pos[3 * i + m] - tmin[m]) / d) + 1;
            nd = imin(nd, c->cell_dim[m]);
            nd = imax(nd, 1);
            c->sphere_node[3 * i + m] = nd;
        }
    int nn = 0;
    for (int iz = -1; iz <= 1; iz++)
        for (int iy = -1; iy <= 1; iy++)
            for (int ix = -1; ix <= 1; ix++) {
                int ir = ix * ix + iy * iy + iz * iz;
                if (neighbor_model == 0 && ir > 0) continue;
                if (neighbor_model == 1 && ir > 1) continue;
                if (neighbor_model == 2 && ir > 2) continue;
                c->nbr[nn][0] = ix;
                c->nbr[nn][1] = iy;
                c->nbr[nn][2] = iz;
                nn++;
            }
    c->nnb = nn;
    /* near-field pair list: for sphere i, each stencil cell in order, members
     * in insertion (index) order, keep j > i (:794-823) */
    free(c->pairs);
    c->pairs = malloc(sizeof(int) * 2 * (size_t)c->n * 64 + 16);
    int cap = c->n * 64, np = 0;
    for (int i = 0; i < c->n; i++) {
        for (int q = 0; q < nn; q++) {
            int cx = c->sphere_node[3 * i] - c->nbr[q][0], cy = c->sphere_node[3 * i + 1] - c->nbr[q][1],
                cz = c->sphere_node[3 * i + 2] - c->nbr[q][2];
            if (cx < 1 || cy < 1 || cz < 1 || cx > c->cell_dim[0] || cy > c->cell_dim[1] || cz > c->cell_dim[2]) continue;
            for (int j = 0; j < c->n; j++) {
                if (c->sphere_node[3 * j] != cx || c->sphere_node[3 * j + 1] != cy || c->sphere_node[3 * j + 2] != cz)
                    continue;
                if (j > i) {
                    if (np >= cap) {
                        cap *= 2;
                        c->pairs = realloc(c->pairs, sizeof(int) * 2 * cap);
                    }
                    c->pairs[2 * np] = i;
                    c->pairs[2 * np + 1] = j;
                    np++;
                }
            }
        }
    }
    c->npairs = np;
    c->fft_ready = 1;
    memcpy(cell_dim, c->cell_dim, sizeof(int) * 3);
    *d_cell = c->d_cell;
    *node_order = c->node_order;
    return 0;
}

int orc_local_pairs(const orc_cluster *c, int *pairs)
{
    if (pairs) memcpy(pairs, c->pairs, sizeof(int) * 2 * c->npairs);
    return c->npairs;
}

static int ncells_of(const orc_cluster *c) { return c->cell_dim[0] * c->cell_dim[1] * c->cell_dim[2]; }

/* local_sphere_to_node_translation (:482-652), sphere_to_node branch.
 * anode(Nx,Ny,Nz, 2*nblk) for one rhs.                                 */
void orc_node_scatter(const orc_cluster *c, const zc *asphere, zc *anode)
{
    int LG = c->node_order, nb2 = 2 * LG * (LG + 2), nc = ncells_of(c);
    zc ri2[2] = {c->rimed, c->rimed};
    zc *t = malloc(sizeof(zc) * nb2);
    memset(anode, 0, sizeof(zc) * nc * nb2);
    for (int i = 0; i < c->n; i++) {
        const int *nd = c->sphere_node + 3 * i;
        double rt[3];
        for (int m = 0; m < 3; m++) rt[m] = c->d_cell * ((double)nd[m] - 0.5) - c->pos[3 * i + m] + c->cell_boundary[m];
        memset(t, 0, sizeof(zc) * nb2);
        orc_pair_translate(1, rt, ri2, c->order[i], LG, 0, 0, imin(c->order[i], LG) >= c->switch_order,
                           asphere + c->offset[i], t);
        int cell = (nd[0] - 1) + c->cell_dim[0] * ((nd[1] - 1) + c->cell_dim[1] * (nd[2] - 1));
        for (int q = 0; q < nb2; q++) anode[cell + nc * q] += t[q];
    }
    free(t);
}

void orc_node_gather(const orc_cluster *c, const zc *anode, zc *asphere)
{
    int LG = c->node_order, nb2 = 2 * LG * (LG + 2), nc = ncells_of(c);
    zc ri2[2] = {c->rimed, c->rimed};
    zc *t = malloc(sizeof(zc) * nb2);
    memset(asphere, 0, sizeof(zc) * c->neqns);
    for (int i = 0; i < c->n; i++) {
        const int *nd = c->sphere_node + 3 * i;
        double rt[3];
        for (int m = 0; m < 3; m++) rt[m] = c->d_cell * ((double)nd[m] - 0.5) - c->pos[3 * i + m] + c->cell_boundary[m];
        int cell = (nd[0] - 1) + c->cell_dim[0] * ((nd[1] - 1) + c->cell_dim[1] * (nd[2] - 1));
        for (int q = 0; q < nb2; q++) t[q] = anode[cell + nc * q];
        orc_pair_translate(1, rt, ri2, LG, c->order[i], 1, 0, imin(c->order[i], LG) >= c->switch_order, t,
                           asphere + c->offset[i]);
    }
    free(t);
}

/* fft_node_to_node_translation (:827-899) restated as the direct sum it
 * computes: g(i) = sum_j H(d (i-j)) a(j) over node pairs whose offset lies
 * outside the neighbour stencil (the hole, :929-937), per mode.         */
void orc_node_convolution(const orc_cluster *c, const zc *anode, zc *gnode)
{
    int LG = c->node_order, nb = LG * (LG + 2), nc = ncells_of(c);
    int Nx = c->cell_dim[0], Ny = c->cell_dim[1], Nz = c->cell_dim[2];
    zc ri2[2] = {c->rimed, c->rimed};
    memset(gnode, 0, sizeof(zc) * nc * 2 * nb);
    zc *h = malloc(sizeof(zc) * nb * nb * 2);
    for (int dz = -(Nz - 1); dz <= Nz - 1; dz++)
        for (int dy = -(Ny - 1); dy <= Ny - 1; dy++)
            for (int dx = -(Nx - 1); dx <= Nx - 1; dx++) {
                int inhole = 0;
                for (int q = 0; q < c->nnb; q++)
                    if (iabs_(dx) == c->nbr[q][0] && iabs_(dy) == c->nbr[q][1] && iabs_(dz) == c->nbr[q][2]) inhole = 1;
                if (inhole) continue;
                double xp[3] = {c->d_cell * dx, c->d_cell * dy, c->d_cell * dz};
                orc_gentranmatrix(LG, LG, xp, ri2, 3, h);
                for (int iz = imax(0, dz); iz < imin(Nz, Nz + dz); iz++)
                    for (int iy = imax(0, dy); iy < imin(Ny, Ny + dy); iy++)
                        for (int ix = imax(0, dx); ix < imin(Nx, Nx + dx); ix++) {
                            int ci = ix + Nx * (iy + Ny * iz);
                            int cj = (ix - dx) + Nx * ((iy - dy) + Ny * (iz - dz));
                            for (int p = 0; p < 2; p++)
                                for (int n = 0; n < nb; n++) {
                                    zc a = anode[cj + nc * (n + nb * p)];
                                    if (a == 0.0) continue;
                                    for (int l = 0; l < nb; l++) gnode[ci + nc * (l + nb * p)] += h[l + nb * (n + nb * p)] * a;
                                }
                        }
            }
    free(h);
}

/* fft_external_to_external_expansion (:72-336), one process. */
void orc_fft_external_to_external(const orc_cluster *c, int nrhs, const int *contran, const zc *ain, zc *gout)
{
    int ne = c->neqns, LG = c->node_order, nb2 = 2 * LG * (LG + 2), nc = ncells_of(c);
    zc ri2[2] = {c->rimed, c->rimed};
    zc *at = malloc(sizeof(zc) * ne), *an = malloc(sizeof(zc) * nc * nb2), *gn = malloc(sizeof(zc) * nc * nb2);
    zc *gt = malloc(sizeof(zc) * ne);
    for (int r = 0; r < nrhs; r++) {
        memcpy(at, ain + r * ne, sizeof(zc) * ne);
        if (contran[r]) flip_all(c, at);
        orc_node_scatter(c, at, an);
        orc_node_convolution(c, an, gn);
        orc_node_gather(c, gn, gt);
        for (int q = 0; q < c->npairs; q++) {
            int i = c->pairs[2 * q], j = c->pairs[2 * q + 1];
            double rt[3] = {c->pos[3 * i] - c->pos[3 * j], c->pos[3 * i + 1] - c->pos[3 * j + 1],
                            c->pos[3 * i + 2] - c->pos[3 * j + 2]};
            int rot = imin(c->order[i], c->order[j]) >= c->switch_order;
            orc_pair_translate(3, rt, ri2, c->order[j], c->order[i], 0, 0, rot, at + c->offset[j], gt + c->offset[i]);
            orc_pair_translate(3, rt, ri2, c->order[i], c->order[j], 1, 0, rot, at + c->offset[i], gt + c->offset[j]);
        }
        if (contran[r]) flip_all(c, gt);
        memcpy(gout + r * ne, gt, sizeof(zc) * ne);
    }
    free(at);
    free(an);
    free(gn);
    free(gt);
}

/* ------------------------------------------------------------------ */
/* Post-solve observables of a single-origin run                        */
/* (mstm-input-37.f90:1159-1197 and 1503-1535).                         */
/* ------------------------------------------------------------------ */

/* tranordertest (mstm-modules-33.f90:1896-1942): the smallest degree n at
 * which the regular translation of a degree-lmax, m = 1 wave over r has
 * gathered all but eps of its norm (capped at 200, at least lmax).  The
 * medium index is real here, so xi_l = psi_l(z)/z i^l with real z.  psi
 * is evaluated once to degree 200 + lmax (the reference re-evaluates it
 * per n; both are the same functions).                                  */
int orc_tranorder(double r, double rimed, int lmax, double eps)
{
    const int nlim = 200;
    if (r == 0.0) return lmax;
    double z = r * rimed;
    int nw = nlim + lmax;
    double *psi = malloc(sizeof(double) * (nw + 1));
    orc_riccati_psi(nw, z, psi);
    double sum = 0.0;
    int n;
    for (n = 1; n <= nlim; n++) {
        int l = lmax;
        zc c = fnr(2 * n + 1) * fnr(2 * l + 1) * cpow(CI, n - l);
        zc a = 0.0, b = 0.0;
        for (int w = iabs_(n - l); w <= n + l; w++) {
            double v = orc_cg(n, -1, l, 1, w);
            zc xw = psi[w] / z * cpow(CI, w);
            if ((n + l + w) % 2 == 0) a += v * v * xw;
            else b += v * v * xw;
        }
        a *= c;
        b *= c;
        sum += creal(a * conj(a) + b * conj(b));
        if (fabs(1.0 - sum) < eps) break;
    }
    free(psi);
    int nmax = imin(n, nlim);
    return imax(nmax, lmax);
}

/* translation orders about origin r0 and the common-origin order T
 * (mstm-input-37.f90:1159-1197 with set_t_matrix_order = .true.):
 * T = max(min(max_mie_order, maxt), max_i min(tranorder_i, maxt)).     */
int orc_translation_orders(const orc_cluster *c, const double *r0, double eps, int maxt, int *torder)
{
    int lmx = 0;
    for (int i = 0; i < c->n; i++) lmx = imax(lmx, c->order[i]);
    int T = imin(lmx, maxt);
    for (int i = 0; i < c->n; i++) {
        double d[3] = {c->pos[3 * i] - r0[0], c->pos[3 * i + 1] - r0[1], c->pos[3 * i + 2] - r0[2]};
        double rt = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        torder[i] = imin(orc_tranorder(rt, creal(c->rimed), c->order[i], eps), maxt);
        T = imax(T, torder[i]);
    }
    return T;
}

/* merge_to_common_origin (mstm-scatprops-26.f90:447-559): every sphere's
 * solution (L/R modes) is carried to r0 by the regular translation of
 * order min(T, torder_i) and summed; the result is stored in the E/M form
 * a0(.,.,1) = aL + aR, a0(.,.,2) = aL - aR, layout (0:T+1, T, 2) per rhs.
 * amnp holds nrhs columns of neqns, amnp0 nrhs columns of 2T(T+2).      */
void orc_merge(const orc_cluster *c, int T, const int *torder, const double *r0, int nrhs, const zc *amnp, zc *amnp0)
{
    int ne = c->neqns, b0 = T * (T + 2);
    zc ri[2] = {c->rimed, c->rimed};
    memset(amnp0, 0, sizeof(zc) * 2 * b0 * nrhs);
    for (int i = 0; i < c->n; i++) {
        int L = c->order[i], Ti = imin(T, torder[i]), bt = Ti * (Ti + 2);
        double rt[3] = {r0[0] - c->pos[3 * i], r0[1] - c->pos[3 * i + 1], r0[2] - c->pos[3 * i + 2]};
        zc *t = malloc(sizeof(zc) * 2 * bt);
        for (int r = 0; r < nrhs; r++) {
            memset(t, 0, sizeof(zc) * 2 * bt);
            orc_pair_translate(1, rt, ri, L, Ti, 0, 0, imax(L, Ti) >= c->switch_order, amnp + (size_t)r * ne + c->offset[i],
                               t);
            zc *o = amnp0 + (size_t)r * 2 * b0;
            for (int n = 1; n <= Ti; n++)
                for (int m = -n; m <= n; m++) {
                    zc a1 = t[amnaddr(m, n, Ti) - 1], a2 = t[amnaddr(m, n, Ti) - 1 + bt];
                    o[amnaddr(m, n, T) - 1] += a1 + a2;
                    o[amnaddr(m, n, T) - 1 + b0] += a1 - a2;
                }
        }
        free(t);
    }
}

/* amplitude_to_scattering_matrix (mstm-scatprops-26.f90:1410-1436);
 * sm is the Fortran sm(4,4), column-major.                              */
static void amp_to_sm(const zc *sa, double *sm)
{
    zc sp[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) sp[i][j] = sa[i] * conj(sa[j]);
#define SM(i, j) sm[((i) - 1) + 4 * ((j) - 1)]
#define SP(i, j) sp[(i) - 1][(j) - 1]
    SM(1, 1) = creal(SP(1, 1) + SP(2, 2) + SP(3, 3) + SP(4, 4));
    SM(1, 2) = creal(-SP(1, 1) + SP(2, 2) - SP(3, 3) + SP(4, 4));
    SM(2, 1) = creal(-SP(1, 1) + SP(2, 2) + SP(3, 3) - SP(4, 4));
    SM(2, 2) = creal(SP(1, 1) + SP(2, 2) - SP(3, 3) - SP(4, 4));
    SM(3, 3) = creal(2.0 * (SP(1, 2) + SP(3, 4)));
    SM(3, 4) = 2.0 * cimag(SP(2, 1) + SP(4, 3));
    SM(4, 3) = 2.0 * cimag(SP(1, 2) - SP(3, 4));
    SM(4, 4) = creal(2.0 * (SP(1, 2) - SP(3, 4)));
    SM(1, 3) = creal(2.0 * (SP(2, 3) + SP(1, 4)));
    SM(3, 1) = creal(2.0 * (SP(2, 4) + SP(1, 3)));
    SM(1, 4) = -2.0 * cimag(SP(2, 3) - SP(1, 4));
    SM(4, 1) = -2.0 * cimag(SP(4, 2) + SP(1, 3));
    SM(2, 3) = creal(-2.0 * (SP(2, 3) - SP(1, 4)));
    SM(3, 2) = creal(-2.0 * (SP(2, 4) - SP(1, 3)));
    SM(2, 4) = -2.0 * cimag(SP(2, 3) + SP(1, 4));
    SM(4, 2) = -2.0 * cimag(SP(4, 2) - SP(1, 3));
#undef SM
#undef SP
}

/* scatteringmatrix (mstm-scatprops-26.f90:1498-1573) with rotate_plane =
 * .false., normalize_s11 = .false.: amplitude matrix from the common-origin
 * E/M coefficients a0 (0:T+1, T, 2, 2) at (ct, phi) with the normalised
 * vector harmonics of taufunc (mstm-modules-33.f90:1324-1344), then the
 * 16 Mueller elements / (4 pi) / medium index.  pi is the reference's
 * widened real(4) value.                                                */
void orc_scatmat(int T, const zc *a0, double ct, double phi, double rimed, double *sm, zc *sa_out)
{
    int b0 = T * (T + 2);
    double *dc = malloc(sizeof(double) * 3 * (T + 1) * (T + 1));
    orc_rotcoef(ct, 1, T, dc);
#define DR(k, mn) dc[(k) + 1 + 3 * (mn)]
    zc ephi = cos(phi) + CI * sin(phi);
    zc *ephim = malloc(sizeof(zc) * (2 * T + 1));
    ephim[T] = 1.0;
    for (int m = 1; m <= T; m++) {
        ephim[T + m] = ephi * ephim[T + m - 1];
        ephim[T - m] = conj(ephim[T + m]);
    }
    zc sa[4] = {0, 0, 0, 0};
    for (int n = 1; n <= T; n++) {
        zc cin = cpow(-CI, n);
        double fnm = sqrt((double)(2 * n + 1) / 2.0) / 4.0;
        for (int m = -n; m <= n; m++) {
            int mn = n * (n + 1) + m;
            double tau[2] = {-fnm * (-DR(-1, mn) + DR(1, mn)), -fnm * (DR(-1, mn) + DR(1, mn))};
            int ad = amnaddr(m, n, T) - 1;
            for (int p = 0; p < 2; p++) {
                zc a = a0[ad + p * b0], b = -a0[ad + p * b0 + 2 * b0];
                sa[0] += cin * tau[1 - p] * b * ephim[T + m];
                sa[1] += CI * cin * tau[p] * a * ephim[T + m];
                sa[2] += CI * cin * tau[p] * b * ephim[T + m];
                sa[3] += cin * tau[1 - p] * a * ephim[T + m];
            }
        }
    }
#undef DR
    double qsca = 1.0 / PI_R4;
    for (int k = 0; k < 4; k++) sa[k] *= 4.0 / sqrt(2.0 * qsca);
    amp_to_sm(sa, sm);
    for (int k = 0; k < 16; k++) sm[k] = sm[k] / (4.0 * PI_R4) / rimed;
    if (sa_out)
        for (int k = 0; k < 4; k++) sa_out[k] = sa[k];
    free(dc);
    free(ephim);
}

/* fosmexpansion (mstm-scatprops-26.f90:1820-1951): generalized-spherical-
 * function expansion of the azimuthally averaged scattering matrix.  Each
 * output is a Fortran real(8) (4,4,0:2T) array.  s00(1,1,1)/(3 s00(1,1,0))
 * is the asymmetry parameter; 4 pi s00(1,1,0) is the scattering cross
 * section in units of k^-2.                                             */
void orc_fosm(int T, const zc *amnp, double rimed, double *s00, double *s02, double *sp22, double *sm22)
{
    int b0 = T * (T + 2), W = 2 * T + 1, ld = 2 * T + 5;
    zc *A1 = calloc((size_t)ld * (T + 1) * 2, sizeof(zc)), *A2 = calloc((size_t)ld * (T + 1) * 2, sizeof(zc));
#define AA(A, m, n, p) A[((m) + T + 2) + ld * ((n) + (T + 1) * (p))]
    for (int n = 1; n <= T; n++)
        for (int p = 0; p < 2; p++)
            for (int m = -n; m <= n; m++) {
                AA(A1, m, n, p) = amnp[amnaddr(m, n, T) - 1 + p * b0];
                AA(A2, m, n, p) = amnp[amnaddr(m, n, T) - 1 + p * b0 + 2 * b0];
            }
    size_t nall = (size_t)16 * W;
    memset(s00, 0, sizeof(double) * nall);
    memset(s02, 0, sizeof(double) * nall);
    memset(sp22, 0, sizeof(double) * nall);
    memset(sm22, 0, sizeof(double) * nall);
#define S(a, i, j, w) a[((i) - 1) + 4 * (((j) - 1) + 4 * (w))]
    double *cm1p1 = malloc(sizeof(double) * (W + 1)), *cm1m1 = malloc(sizeof(double) * (W + 1));
    double *cmm = malloc(sizeof(double) * (W + 1)), *cmm2 = malloc(sizeof(double) * (W + 1)),
           *cmp2 = malloc(sizeof(double) * (W + 1));
    for (int n = 1; n <= T; n++)
        for (int l = 1; l <= T; l++) {
            int wmin = iabs_(n - l), wmax = n + l;
            zc fnl = sqrt((double)((2 * n + 1) * (2 * l + 1))) * cpow(CI, l - n);
            for (int w = wmin; w <= wmax; w++) {
                cm1p1[w] = orc_cg(n, -1, l, 1, w);
                cm1m1[w] = orc_cg(n, -1, l, -1, w);
            }
            int mm = imin(n, l + 2);
            for (int m = -mm; m <= mm; m++) {
                double m1m = sgn1(iabs_(m));
                for (int w = wmin; w <= wmax; w++) {
                    cmm[w] = (iabs_(m) <= l) ? orc_cg(n, -m, l, m, w) : 0.0;
                    cmm2[w] = (iabs_(m - 2) <= l) ? orc_cg(n, -m, l, m - 2, w) : 0.0;
                    cmp2[w] = (iabs_(m + 2) <= l) ? orc_cg(n, -m, l, m + 2, w) : 0.0;
                }
                for (int p = 0; p < 2; p++)
                    for (int q = 0; q < 2; q++) {
                        double m1mq = sgn1(iabs_(m + q + 1));
                        double m1mnpl = sgn1(iabs_(m + n + p + 1 + l));
                        zc a1122 = AA(A1, m, n, p) * conj(AA(A1, m, l, q)) + AA(A2, m, n, p) * conj(AA(A2, m, l, q));
                        zc a2112 = AA(A2, m, n, p) * conj(AA(A1, m, l, q)) - AA(A1, m, n, p) * conj(AA(A2, m, l, q));
                        zc a1p2 = (AA(A1, m, n, p) + CI * AA(A2, m, n, p)) *
                                  conj(AA(A1, m - 2, l, q) - CI * AA(A2, m - 2, l, q));
                        zc a1m2 = (AA(A1, m, n, p) - CI * AA(A2, m, n, p)) *
                                  conj(AA(A1, m + 2, l, q) + CI * AA(A2, m + 2, l, q));
                        for (int w = wmin; w <= wmax; w++) {
                            double m1w = sgn1(w);
                            int even = ((n + l + w + p + q + 2) % 2) == 0;
                            if (even) S(s00, 1, 1, w) += creal(-(m1m * fnl * a1122 * cm1p1[w] * cmm[w]) / 2.0);
                            else S(s00, 4, 4, w) += creal(CI / 2.0 * m1m * fnl * a2112 * cm1p1[w] * cmm[w]);
                            if (w < 2) continue;
                            if (even) {
                                S(s02, 2, 1, w) += creal(-(m1mq * a1122 * fnl * cm1m1[w] * cmm[w]) / 2.0);
                                S(s02, 1, 2, w) += creal(-(m1m * a1p2 * fnl * cm1p1[w] * cmm2[w]) / 4.0);
                                S(s02, 1, 2, w) += creal(-(m1m * a1m2 * fnl * cm1p1[w] * cmp2[w]) / 4.0);
                            } else {
                                S(s02, 3, 4, w) += cimag(-CI / 2.0 * m1mq * a2112 * fnl * cm1m1[w] * cmm[w]);
                                S(s02, 4, 3, w) += cimag(m1m * a1p2 * fnl * cm1p1[w] * cmm2[w]) / 4.0;
                                S(s02, 4, 3, w) -= cimag(m1m * a1m2 * fnl * cm1p1[w] * cmp2[w]) / 4.0;
                            }
                            S(sm22, 2, 2, w) += creal(-(m1mnpl * m1w * a1p2 * fnl * cm1m1[w] * cmm2[w]) / 8.0);
                            S(sp22, 2, 2, w) += creal(-(m1mq * a1p2 * fnl * cm1m1[w] * cmm2[w]) / 8.0);
                            S(sm22, 2, 2, w) += creal(-(m1mq * a1m2 * fnl * cm1m1[w] * cmp2[w]) / 8.0);
                            S(sp22, 2, 2, w) += creal(-(m1mnpl * m1w * a1m2 * fnl * cm1m1[w] * cmp2[w]) / 8.0);
                        }
                    }
            }
        }
    /* sign patch and scale of the reference (:1930-1948) */
    for (int w = 0; w < W; w++) {
        for (int i = 3; i <= 4; i++)
            for (int j = 1; j <= i; j++) {
                S(s00, j, i, w) = -S(s00, j, i, w);
                S(s02, j, i, w) = -S(s02, j, i, w);
                S(sm22, j, i, w) = -S(sm22, j, i, w);
                S(sp22, j, i, w) = -S(sp22, j, i, w);
            }
        S(sm22, 3, 3, w) = -S(sm22, 2, 2, w);
        S(sp22, 3, 3, w) = S(sp22, 2, 2, w);
    }
    for (size_t k = 0; k < nall; k++) {
        s00[k] *= 0.5 / rimed;
        s02[k] *= 0.5 / rimed;
        sm22[k] *= 0.5 / rimed;
        sp22[k] *= 0.5 / rimed;
    }
#undef S
#undef AA
    free(A1);
    free(A2);
    free(cm1p1);
    free(cm1m1);
    free(cmm);
    free(cmm2);
    free(cmp2);
}

/* distribute_from_common_origin (mstm-scatprops-26.f90:561-669), vswf_type
 * 1, one column: the common-origin expansion a0 (order l, layout
 * (0:l+1, l, 2)) is carried to every sphere along r_i - r0, order l -> L_i,
 * rotation path where max(L_i, l) >= translation_switch_order (:642).    */
void orc_distribute(const orc_cluster *c, int l, const double *r0, const zc *a0, zc *amnp)
{
    zc ri[2] = {c->rimed, c->rimed};
    memset(amnp, 0, sizeof(zc) * c->neqns);
    for (int i = 0; i < c->n; i++) {
        int L = c->order[i];
        double rt[3] = {c->pos[3 * i] - r0[0], c->pos[3 * i + 1] - r0[1], c->pos[3 * i + 2] - r0[2]};
        orc_pair_translate(1, rt, ri, l, L, 0, 0, imax(L, l) >= c->switch_order, a0, amnp + c->offset[i]);
    }
}

/* tmatrix_solution (mstm-solver-8.f90:22-332), serial, iterative branch.
 * For l = 1..T and (k, q) in the reference's order (k = -l..l, q = 1, 2):
 * the unit TE/TM excitation of (k, l, q) at the origin (:169-175) is
 * distributed to the spheres, multiplied by the Mie matrices, solved with
 * cbicg from x = b, and merged back to the origin at order l (target order
 * min(l, torder_i)).  Outputs:
 *   tmat   the T-matrix file body in its write order (:243-256): per
 *          column, rows n = 1..l, m = -n..n, p = 1, 2 (TE/TM form);
 *   sqeff  [sphere][3] sums of the per-sphere efficiencies over all
 *          columns (qefficiencyfactors with npol = 1: ext, abs, sca);
 *   qteff  [3] the T-matrix efficiency sums of :233-237;
 *   qlog   [l][4] per order: cumulative max iterations, qeff(1), qeff(2),
 *          converr (:265-276).
 * Stops after order l when converr < conveps (conveps > 0, :322).
 * Returns the order reached; *istat = 1 if a solve missed eps or niter.   */
int orc_tmatrix(const orc_cluster *c, int T, const int *torder, const double *r0, double vol_radius, double eps,
                int niter, double conveps, zc *tmat, double *sqeff, double *qteff, double *qlog, int *maxiter,
                double *maxerr, int *istat)
{
    int ne = c->neqns;
    zc *dist = malloc(sizeof(zc) * 2 * ne), *b = malloc(sizeof(zc) * 2 * ne), *x = malloc(sizeof(zc) * 2 * ne);
    zc *a0 = malloc(sizeof(zc) * 2 * T * (T + 2)), *u = malloc(sizeof(zc) * 2 * T * (T + 2));
    zc *am0 = malloc(sizeof(zc) * 2 * T * (T + 2));
    double *dq = malloc(sizeof(double) * 9 * c->n);
    double qeffold = 0.0, vr2 = vol_radius * vol_radius;
    size_t nt = 0;
    int l;
    *istat = 0;
    *maxiter = 0;
    *maxerr = 0.0;
    memset(sqeff, 0, sizeof(double) * 3 * c->n);
    qteff[0] = qteff[1] = qteff[2] = 0.0;
    for (l = 1; l <= T; l++) {
        int b0 = l * (l + 2);
        for (int k = -l; k <= l; k++) {
            /* both q of one k: their solves are independent, done one by one */
            for (int q = 1; q <= 2; q++) {
                memset(u, 0, sizeof(zc) * 2 * b0);
                u[amnaddr(k, l, l) - 1] = 0.5;
                u[amnaddr(k, l, l) - 1 + b0] = -0.5 * ((q & 1) ? -1.0 : 1.0);
                orc_distribute(c, l, r0, u, dist + (size_t)(q - 1) * ne);
                orc_mie_mult(c, 1, 0, dist + (size_t)(q - 1) * ne, b + (size_t)(q - 1) * ne);
                memcpy(x + (size_t)(q - 1) * ne, b + (size_t)(q - 1) * ne, sizeof(zc) * ne);
                double err = 0.0;
                int it = orc_cbicg(c, 0, niter, eps, b + (size_t)(q - 1) * ne, x + (size_t)(q - 1) * ne, &err);
                if (it > *maxiter) *maxiter = it;
                if (err > *maxerr) *maxerr = err;
                if (it > niter || err > eps) *istat = 1;
            }
            orc_qeff(c, x, dist, dq);
            for (int q = 1; q <= 2; q++) {
                for (int i = 0; i < c->n; i++)
                    for (int e = 0; e < 3; e++) sqeff[3 * i + e] += dq[9 * i + 3 * q + e];
                orc_merge(c, l, torder, r0, 1, x + (size_t)(q - 1) * ne, am0);
                int ka = amnaddr(k, l, l) - 1 + (q - 1) * b0;
                double d1 = -2.0 / vr2 * creal(am0[ka]), d3 = 0.0;
                for (int e = 0; e < 2 * b0; e++) d3 += creal(am0[e] * conj(am0[e]));
                d3 *= 2.0 / vr2;
                qteff[0] += d1;
                qteff[1] += d1 - d3;
                qteff[2] += d3;
                for (int n = 1; n <= l; n++)
                    for (int m = -n; m <= n; m++)
                        for (int p = 0; p < 2; p++) tmat[nt++] = am0[amnaddr(m, n, l) - 1 + p * b0];
            }
        }
        double qe[2] = {0.0, 0.0};
        for (int i = 0; i < c->n; i++)
            for (int e = 0; e < 2; e++) qe[e] += sqeff[3 * i + e] * (c->rad[i] * c->rad[i]) / vr2;
        double converr = qe[0] - qeffold;
        qeffold = qe[0];
        qlog[4 * (l - 1)] = *maxiter;
        qlog[4 * (l - 1) + 1] = qe[0];
        qlog[4 * (l - 1) + 2] = qe[1];
        qlog[4 * (l - 1) + 3] = converr;
        if (conveps > 0.0 && converr < conveps) break;
    }
    free(dist);
    free(b);
    free(x);
    free(a0);
    free(u);
    free(am0);
    free(dq);
    return l <= T ? l : T;
}

/* the expansion-coefficient stage of ranorientscatmatrix (:2336-2472):
 * dm/dmcf [u+T+1][ku][n][p][n1][p1] (u = -T-1..T+1) -> sm, smcf.          */
void orc_ranorient_coef(int T, const zc *dm, const zc *dmcf, double rimed, double *sm, double *smcf)
{
    const int W = 2 * T;
#define DMI(A, u, ku, n, p, n1, p1) \
    A[(((((size_t)((u) + T + 1) * 3 + (ku) - 1) * T + (n) - 1) * 2 + (p) - 1) * T + (n1) - 1) * 2 + (p1) - 1]
    size_t nfm = (size_t)3 * T * 2 * T * 2;
    zc *fm = malloc(sizeof(zc) * nfm), *fmcf = malloc(sizeof(zc) * nfm);
#define FM(A, k, n, p, n1, p1) A[((((size_t)(k) - 1) * T + (n) - 1) * 2 + (p) - 1) * T * 2 + ((n1) - 1) * 2 + (p1) - 1]
    /* aw(0:2,-1:1,w), bw likewise, cw(w), dw(w) */
    zc *aw = calloc((size_t)9 * (W + 1), sizeof(zc)), *bw = calloc((size_t)9 * (W + 1), sizeof(zc));
    zc *awcf = calloc((size_t)9 * (W + 1), sizeof(zc)), *bwcf = calloc((size_t)9 * (W + 1), sizeof(zc));
    zc *cw = calloc(W + 1, sizeof(zc)), *dw = calloc(W + 1, sizeof(zc));
    zc *cwcf = calloc(W + 1, sizeof(zc)), *dwcf = calloc(W + 1, sizeof(zc));
#define AW(A, i, k, w) A[((w) * 3 + (k) + 1) * 3 + (i)]
    for (int w = 0; w <= W; w++) {
        memset(fm, 0, sizeof(zc) * nfm);
        memset(fmcf, 0, sizeof(zc) * nfm);
        const double s2w = sqrt((double)(w + w + 1));
        for (int u = -T - 1; u <= T + 1; u++) {
            for (int k = -1; k <= 1; k += 2) {
                int m = u + k, ik = (k + 1) / 2 + 1, ns = imax(1, iabs_(m));
                for (int n = ns; n <= T; n++)
                    for (int n1 = ns; n1 <= T; n1++) {
                        if (n + n1 < w || iabs_(n - n1) > w) continue;
                        double fl = -((n & 1) ? -1.0 : 1.0) * orc_cg(n, m, w, 0, n1) * s2w / sqrt((double)(n1 + n1 + 1));
                        for (int p = 1; p <= 2; p++)
                            for (int p1 = 1; p1 <= 2; p1++) {
                                FM(fm, ik, n, p, n1, p1) += DMI(dm, u, ik, n, p, n1, p1) * fl;
                                FM(fmcf, ik, n, p, n1, p1) += DMI(dmcf, u, ik, n, p, n1, p1) * fl;
                            }
                    }
            }
            if (w < 2) continue;
            int m = u + 1, m1 = u - 1, ns = imax(1, iabs_(m)), n1s = imax(1, iabs_(m1));
            for (int n = ns; n <= T; n++)
                for (int n1 = n1s; n1 <= T; n1++) {
                    if (n + n1 < w || iabs_(n - n1) > w) continue;
                    double fl = -((n & 1) ? -1.0 : 1.0) * orc_cg(n, m, w, -2, n1) * s2w / sqrt((double)(n1 + n1 + 1));
                    for (int p = 1; p <= 2; p++)
                        for (int p1 = 1; p1 <= 2; p1++) {
                            FM(fm, 3, n, p, n1, p1) += DMI(dm, u, 3, n, p, n1, p1) * fl;
                            FM(fmcf, 3, n, p, n1, p1) += DMI(dmcf, u, 3, n, p, n1, p1) * fl;
                        }
                }
        }
        for (int n = 1; n <= T; n++) {
            int n1s = imax(1, iabs_(n - w)), n1e = imin(T, n + w);
            double in = (n & 1) ? -1.0 : 1.0;
            for (int n1 = n1s; n1 <= n1e; n1++) {
                double fl = 2.0 * in * orc_cg(n, 1, w, 0, n1) * s2w / sqrt((double)(n1 + n1 + 1));
                int i = (n + n1 - w) % 2 + 1;
                for (int p = 1; p <= 2; p++) {
                    int p1 = (2 - i) * p + (i - 1) * (3 - p);
                    for (int k = -1; k <= 1; k += 2) {
                        int ik = (k + 1) / 2 + 1;
                        AW(aw, 0, k, w) += FM(fm, ik, n, p, n1, p1) * fl;
                        AW(bw, 0, k, w) += FM(fm, ik, n, p, n1, 3 - p1) * fl;
                        AW(awcf, 0, k, w) += FM(fmcf, ik, n, p, n1, p1) * fl;
                        AW(bwcf, 0, k, w) += FM(fmcf, ik, n, p, n1, 3 - p1) * fl;
                    }
                    AW(bw, 2, 0, w) += FM(fm, 3, n, p, n1, 3 - p1) * fl;
                    AW(aw, 2, 0, w) += FM(fm, 3, n, p, n1, p1) * fl;
                    AW(bwcf, 2, 0, w) += FM(fmcf, 3, n, p, n1, 3 - p1) * fl;
                    AW(awcf, 2, 0, w) += FM(fmcf, 3, n, p, n1, p1) * fl;
                }
            }
            if (w < 2) continue;
            for (int n1 = n1s; n1 <= n1e; n1++) {
                double vc = orc_cg(n, 1, w, -2, n1);
                double fl = 2.0 * in * vc * s2w / sqrt((double)(n1 + n1 + 1));
                int i = (n + n1 - w) % 2 + 1;
                for (int p = 1; p <= 2; p++) {
                    int p1 = (2 - i) * p + (i - 1) * (3 - p);
                    double sp1 = (p1 & 1) ? -1.0 : 1.0, sq1 = ((3 - p1) & 1) ? -1.0 : 1.0;
                    for (int k = -1; k <= 1; k += 2) {
                        int ik = (k + 1) / 2 + 1;
                        AW(aw, 2, k, w) += FM(fm, ik, n, p, n1, p1) * fl * sp1;
                        AW(bw, 2, k, w) += FM(fm, ik, n, p, n1, 3 - p1) * fl * sq1;
                        AW(awcf, 2, k, w) += FM(fmcf, ik, n, p, n1, p1) * fl * sp1;
                        AW(bwcf, 2, k, w) += FM(fmcf, ik, n, p, n1, 3 - p1) * fl * sq1;
                    }
                }
                double fl2 = 2.0 * (((n1 + w) & 1) ? -1.0 : 1.0) * vc * s2w / sqrt((double)(n1 + n1 + 1));
                for (int p = 1; p <= 2; p++)
                    for (int p1 = 1; p1 <= 2; p1++) {
                        double sp1 = (p1 & 1) ? -1.0 : 1.0, sp = (p & 1) ? -1.0 : 1.0;
                        cw[w] += FM(fm, 3, n, p, n1, p1) * fl * sp1;
                        dw[w] += FM(fm, 3, n, p, n1, p1) * fl2 * sp;
                        cwcf[w] += FM(fmcf, 3, n, p, n1, p1) * fl * sp1;
                        dwcf[w] += FM(fmcf, 3, n, p, n1, p1) * fl2 * sp;
                    }
            }
        }
    }
    for (int pass = 0; pass < 2; pass++) {
        const zc *a = pass ? awcf : aw, *b = pass ? bwcf : bw, *c = pass ? cwcf : cw, *d = pass ? dwcf : dw;
        double *S = pass ? smcf : sm;
        for (int n = 0; n <= W; n++) {
#define S_(i, j) S[((i) - 1) + 4 * ((j) - 1) + 16 * n]
            S_(1, 1) = creal(AW(a, 0, -1, n) + AW(a, 0, 1, n));
            S_(1, 2) = creal(AW(a, 2, -1, n) + AW(a, 2, 1, n));
            S_(1, 3) = 2.0 * cimag(AW(a, 2, 0, n));
            S_(1, 4) = creal(AW(a, 0, 1, n) - AW(a, 0, -1, n));
            S_(2, 2) = creal(d[n]);
            S_(2, 3) = cimag(d[n]);
            S_(2, 4) = creal(AW(a, 2, 1, n) - AW(a, 2, -1, n));
            S_(3, 1) = -cimag(AW(b, 2, -1, n) + AW(b, 2, 1, n));
            S_(3, 2) = cimag(c[n]);
            S_(3, 3) = creal(c[n]);
            S_(3, 4) = cimag(AW(b, 2, -1, n) - AW(b, 2, 1, n));
            S_(4, 1) = creal(AW(b, 0, -1, n) + AW(b, 0, 1, n));
            S_(4, 2) = 2.0 * creal(AW(b, 2, 0, n));
            S_(4, 4) = creal(AW(b, 0, 1, n) - AW(b, 0, -1, n));
            for (int e = 0; e < 16; e++) S[16 * n + e] = S[16 * n + e] / rimed / 2.0;
#undef S_
        }
    }
    free(fm);
    free(fmcf);
    free(aw);
    free(bw);
    free(awcf);
    free(bwcf);
    free(cw);
    free(dw);
    free(cwcf);
    free(dwcf);
#undef DMI
#undef FM
#undef AW
}

/* ranorientscatmatrix (mstm-scatprops-26.f90:1963-2530): the expansion
 * coefficients sm(4,4,0:2T) / smcf(4,4,0:2T) (Fortran order, column-major
 * per w) of the orientation-averaged scattering matrix and of its coherent
 * part, from the T-matrix file body of tmatrix_solution (order T, TE/TM).
 * As the reference: the file is read into complex(4) and its rows n > l
 * filled by reciprocity (:2066-2076); the Gaussian-beam weights (beam width
 * cbeam, 0 = plane wave) go into the same complex(4) array (:2143-2169);
 * the rest is FP64.  Vector-coupling values C(j1 m1; j2 m2 | J m1+m2) come
 * from orc_cg (vcfunc, mstm-modules-33.f90:950-991, gives the same values by
 * recurrence).                                                            */
typedef float complex fzc;
static inline zc ipow(int e)
{
    int r = ((e % 4) + 4) % 4;
    return r == 0 ? 1.0 : (r == 1 ? I : (r == 2 ? -1.0 : -I));
}
void orc_ranorient(int T, const zc *tmat, double rimed, double cbeam, double *sm, double *smcf)
{
    const int nblk = T * (T + 2), W = 2 * T;
    /* tc(p, mn, q, kl), 1-based mn/kl -> [((kl-1)*2 + q-1)*nblk + mn-1][p-1] */
    fzc *tc = calloc((size_t)4 * nblk * nblk, sizeof(fzc));
#define TC(p, mn, q, kl) tc[((((size_t)(kl) - 1) * 2 + (q) - 1) * nblk + (mn) - 1) * 2 + (p) - 1]
    size_t it = 0;
    for (int l = 1; l <= T; l++)
        for (int k = -l; k <= l; k++) {
            int kl = l * (l + 1) + k, klm = l * (l + 1) - k;
            for (int q = 1; q <= 2; q++)
                for (int n = 1; n <= l; n++)
                    for (int m = -n; m <= n; m++) {
                        int mn = n * (n + 1) + m, mnm = n * (n + 1) - m;
                        for (int p = 1; p <= 2; p++) {
                            zc v = tmat[it++];
                            TC(p, mn, q, kl) = (float)creal(v) + (float)cimag(v) * I;
                        }
                        if (n < l) {
                            float ikm = ((m + k) & 1) ? -1.0f : 1.0f;
                            for (int p = 1; p <= 2; p++) TC(q, klm, p, mnm) = TC(p, mn, q, kl) * ikm;
                        }
                    }
        }
    /* Gaussian-beam modified T matrix */
    zc *pp = calloc((size_t)(T + 1) * 4, sizeof(zc));
#define PP(n, a, b) pp[(n) * 4 + ((a) - 1) * 2 + (b) - 1]
    for (int n = 1; n <= T; n++) {
        double gbn = exp(-pow(((double)n + .5) * cbeam, 2.0));
        zc cin = ipow(n + 1);
        PP(n, 1, 1) = -.5 * cin * sqrt((double)(n + n + 1)) * gbn;
        PP(n, 2, 1) = -PP(n, 1, 1);
        PP(n, 1, 2) = -PP(n, 1, 1);
        PP(n, 2, 2) = PP(n, 2, 1);
    }
    for (int n = 1; n <= T; n++)
        for (int m = -n; m <= n; m++) {
            int mn = n * (n + 1) + m;
            for (int p = 1; p <= 2; p++)
                for (int l = 1; l <= T; l++)
                    for (int k = -l; k <= l; k++) {
                        int kl = l * (l + 1) + k;
                        zc a = TC(p, mn, 1, kl);
                        TC(p, mn, 1, kl) = (fzc)(TC(p, mn, 1, kl) * PP(l, 1, 1) + TC(p, mn, 2, kl) * PP(l, 1, 2));
                        TC(p, mn, 2, kl) = (fzc)(a * PP(l, 2, 1) + TC(p, mn, 2, kl) * PP(l, 2, 2));
                    }
        }
    /* dm(u, ku, n, p, n1, p1), u = -T-1..T+1 */
    const int NU = 2 * T + 3;
    size_t ndm = (size_t)NU * 3 * T * 2 * T * 2;
    zc *dm = calloc(ndm, sizeof(zc)), *dmcf = calloc(ndm, sizeof(zc));
#define DM(A, u, ku, n, p, n1, p1) \
    A[(((((size_t)((u) + T + 1) * 3 + (ku) - 1) * T + (n) - 1) * 2 + (p) - 1) * T + (n1) - 1) * 2 + (p1) - 1]
    zc *bm = malloc(sizeof(zc) * 2 * nblk * 2), *bmcf = malloc(sizeof(zc) * 2 * nblk * 2);
#define BM(A, k, mn, p) A[(((k) - 1) * nblk + (mn) - 1) * 2 + (p) - 1]
    zc *am = malloc(sizeof(zc) * 2 * (T + 2) * 2), *amcf = malloc(sizeof(zc) * 2 * (T + 2) * 2);
#define AM(A, k, l, p) A[(((k) - 1) * (T + 2) + (l)) * 2 + (p) - 1]
    for (int w = 0; w <= W; w++)
        for (int v = -w; v <= w; v++) {
            memset(bm, 0, sizeof(zc) * 4 * nblk);
            memset(bmcf, 0, sizeof(zc) * 4 * nblk);
            for (int n = 1; n <= T; n++) {
                int nn1 = n * (n + 1);
                memset(am, 0, sizeof(zc) * 4 * (T + 2));
                memset(amcf, 0, sizeof(zc) * 4 * (T + 2));
                for (int t = -n; t <= n; t++) {
                    int tn = nn1 + t, lmax = imin(T, w + n);
                    for (int l = imax(1, imax(iabs_(v - t), iabs_(n - w))); l <= lmax; l++) {
                        double vc = orc_cg(w, v, n, -t, l);
                        int tvl = l * (l + 1) + t - v;
                        for (int k = 1; k <= 2; k++)
                            for (int p = 1; p <= 2; p++) {
                                AM(am, k, l, p) += vc * (zc)TC(p, tn, k, tvl);
                                if (l == n && v == 0) {
                                    zc tct = 0.0;
                                    for (int kt = -n; kt <= n; kt++) tct += (zc)TC(p, nn1 + kt, k, nn1 + kt);
                                    AM(amcf, k, l, p) += vc * tct / (double)(n + n + 1);
                                }
                            }
                    }
                }
                for (int m = -n; m <= n; m++) {
                    int mn = nn1 + m;
                    for (int k = 1; k <= 2; k++) {
                        int u = m - (-3 + 2 * k);
                        if (iabs_(u) > w) continue;
                        int lmax = imin(T, w + n);
                        for (int l = imax(1, iabs_(w - n)); l <= lmax; l++) {
                            double fl = -((l & 1) ? -1.0 : 1.0) * orc_cg(w, -u, n, m, l) / (double)(l + l + 1);
                            for (int p = 1; p <= 2; p++) {
                                BM(bm, k, mn, p) += AM(am, k, l, p) * fl;
                                if (v == 0) BM(bmcf, k, mn, p) += AM(amcf, k, l, p) * fl;
                            }
                        }
                    }
                }
            }
            for (int u = -imin(w, T + 1); u <= imin(w, T + 1); u++)
                for (int ku = 1; ku <= 3; ku++) {
                    int k = ku == 1 ? -1 : 1, k1 = ku == 2 ? 1 : -1;
                    int m = u + k, ns = imax(1, iabs_(m)), ik = (k + 1) / 2 + 1, ik1 = (k1 + 1) / 2 + 1, m1 = u + k1;
                    for (int n = ns; n <= T; n++) {
                        int nu = n * (n + 1) + m;
                        int n1s = imax(1, imax(iabs_(m1), n - W)), n1e = imin(T, n + W);
                        for (int n1 = n1s; n1 <= n1e; n1++) {
                            zc cin = ipow(n - n1);
                            int nu1 = n1 * (n1 + 1) + m1;
                            double fl = -sqrt((double)(n + n + 1)) * sqrt((double)(n1 + n1 + 1)) * (double)(w + w + 1);
                            for (int p = 1; p <= 2; p++)
                                for (int p1 = 1; p1 <= 2; p1++) {
                                    DM(dm, u, ku, n, p, n1, p1) += BM(bm, ik, nu, p) * cin * fl * conj(BM(bm, ik1, nu1, p1));
                                    if (v == 0)
                                        DM(dmcf, u, ku, n, p, n1, p1) +=
                                            BM(bmcf, ik, nu, p) * cin * fl * conj(BM(bmcf, ik1, nu1, p1));
                                }
                        }
                    }
                }
        }
    orc_ranorient_coef(T, dm, dmcf, rimed, sm, smcf);
    free(tc);
    free(pp);
    free(dm);
    free(dmcf);
    free(bm);
    free(bmcf);
    free(am);
    free(amcf);
#undef TC
#undef PP
#undef BM
#undef AM
}
