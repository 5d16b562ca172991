// specfun.hpp -- FP64 special functions that generate translation tables.
//
// Written from the mathematics; each routine names the reference routine
// whose outputs it reproduces (mstm-modules-33.f90 in the MSTM tree).
#pragma once
#include "common.hpp"

// Spherical Hankel h_w(x) = xi_w(x)/x, w = 0..nw, real x > 0, by upward
// recurrence h_w = (2w-1)/x h_{w-1} - h_{w-2} (the recurrence gentranmatrix
// uses, mstm-modules-33.f90:1838-1842).  y_w dominates, so the error stays
// relative to |h_w|.
MSTM_HD void sph_hankel_up(int nw, double x, cplx<double> *h)
{
    double c = cos(x), s = sin(x);
    // h0 = -i e^{ix}/x ; h1 = -e^{ix}(i+x)/x^2
    h[0] = cmk(s / x, -c / x);
    if (nw >= 1) {
        cplx<double> e = cmk(c, s), t = cmk(x, 1.0);
        cplx<double> v = cmul(e, t);
        h[1] = cmk(-v.re / (x * x), -v.im / (x * x));
    }
    for (int w = 2; w <= nw; w++) {
        double f = (double)(2 * w - 1) / x;
        h[w] = cmk(f * h[w - 1].re - h[w - 2].re, f * h[w - 1].im - h[w - 2].im);
    }
}

// Spherical Bessel j_w(x), w = 0..nw, real x >= 0, by Miller's downward
// recurrence normalised on the larger of j_0, j_1 (upward when x >= nw).
// Same functions as cricbessel(...)/x (mstm-modules-33.f90:262-271).
MSTM_HD void sph_bessel_j(int nw, double x, double *j)
{
    if (x < 1e-12) {
        j[0] = 1.0;
        for (int w = 1; w <= nw; w++) j[w] = 0.0;
        return;
    }
    double s = sin(x), c = cos(x);
    double j0 = s / x, j1 = s / (x * x) - c / x;
    if (x >= (double)nw) {
        j[0] = j0;
        if (nw >= 1) j[1] = j1;
        for (int w = 1; w < nw; w++) j[w + 1] = (double)(2 * w + 1) / x * j[w] - j[w - 1];
        return;
    }
    int ns = nw + 30 + (int)(2.0 * x);
    double fp1 = 0.0, f = 1e-290;
    for (int k = ns; k >= 1; k--) {
        double fm1 = (double)(2 * k + 1) / x * f - fp1;  // f_{k-1}
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp1 *= 1e-250;
            for (int q = k; q <= nw; q++) j[q] *= 1e-250;
        }
        if (k - 1 <= nw) j[k - 1] = f;
    }
    double sc = (fabs(j0) > fabs(j1)) ? j0 / j[0] : j1 / j[1];
    for (int w = 0; w <= nw; w++) j[w] *= sc;
}

// Constants of the diamond recurrence (axialtrancoefinit,
// mstm-modules-33.f90:1711-1720); zero outside l >= max(1,|m|).
MSTM_HD double ax_fnm1(int m, int l)
{
    if (l < imx(1, iab(m))) return 0.0;
    return sqrt((double)(l - 1)) * sqrt((double)(l + 1)) * sqrt((double)(l - m)) * sqrt((double)(l + m)) /
           sqrt((double)(2 * l - 1)) / sqrt((double)(2 * l + 1)) / (double)l;
}
MSTM_HD double ax_fn(int m, int l)
{
    if (l < imx(1, iab(m))) return 0.0;
    return (double)m / (double)l / (double)(l + 1);
}
MSTM_HD double ax_fnp1(int m, int l)
{
    if (l < imx(1, iab(m))) return 0.0;
    return sqrt((double)l) * sqrt((double)(l + 2)) * sqrt((double)(l + 1 - m)) * sqrt((double)(l + 1 + m)) /
           sqrt((double)(2 * l + 1)) / sqrt((double)(2 * l + 3)) / (double)(l + 1);
}

// flat index of vcc(n, l, m, w) for orders <= LM
MSTM_HD int vcc_index(int n, int l, int m, int w, int LM)
{
    return ((n * (LM + 1) + l) * (2 * LM + 1) + (m + LM)) * (2 * LM + 1) + w;
}
