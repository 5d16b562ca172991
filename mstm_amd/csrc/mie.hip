// mie.hip -- FP64 single-sphere physics on the device: Lorenz-Mie L/R
// matrices and truncation orders, T-matrix application, plane-wave
// right-hand sides and per-sphere efficiency sums.
#include "common.hpp"
#include "kernels.hpp"
#include "specfun.hpp"

#define MIE_NMAX 64

// Riccati psi_n(x) = x j_n(x), xi_n(x) = x h_n(x), n = 0..nmax, real x.
__device__ void riccati_real(int nmax, double x, double *psi, cplx<double> *xi)
{
    double j[MIE_NMAX + 2];
    sph_bessel_j(nmax, x, j);
    double c0 = -cos(x), c1 = -cos(x) / x - sin(x);  // chi_n = x y_n
    for (int n = 0; n <= nmax; n++) psi[n] = x * j[n];
    xi[0] = cmk(psi[0], c0);
    if (nmax >= 1) xi[1] = cmk(psi[1], c1);
    for (int n = 1; n < nmax; n++) {
        double c2 = (double)(2 * n + 1) / x * c1 - c0;
        xi[n + 1] = cmk(psi[n + 1], c2);
        c0 = c1;
        c1 = c2;
    }
}

// log-derivative D_n(z) = psi_n'(z)/psi_n(z), downward recurrence
__device__ void logderiv(int nmax, cplx<double> z, cplx<double> *d)
{
    int ns = imx(nmax, (int)sqrt(z.re * z.re + z.im * z.im)) + 30;
    cplx<double> dk = cmk(0.0, 0.0);
    for (int k = ns; k >= 1; k--) {
        cplx<double> kz = cdiv(cmk((double)k, 0.0), z);
        cplx<double> dkm1 = csub(kz, cdiv(cmk(1.0, 0.0), cadd(dk, kz)));
        if (k <= nmax) d[k] = dk;
        dk = dkm1;
    }
    d[0] = dk;
}

// 2x2 L/R Mie matrix of degree n (mieoa, mstm-modules-33.f90:5483-5524,
// non-chiral sphere, real medium mm): a = -G^{-1} A with
//   G(s,t) = f(s,t) (D_n(x ms) xi_n(x mm)  - s_s s_t xi_n'(x mm))
//   A(s,t) = f(s,t) (D_n(x ms) psi_n(x mm) - s_s s_t psi_n'(x mm))
//   f(s,t) = (s_s mm + s_t ms) / (ms mm),  s_1 = -1, s_2 = +1.
__device__ void mie_degree(int n, double xm, cplx<double> ms, double mm, const double *psi, const cplx<double> *xi,
                           const cplx<double> *dn, cplx<double> *a4)
{
    double psip = psi[n - 1] - (double)n * psi[n] / xm;
    cplx<double> xip = csub(xi[n - 1], cscale(xi[n], (double)n / xm));
    cplx<double> g[2][2], A[2][2];
    cplx<double> msmm = cscale(ms, mm);
    for (int s = 0; s < 2; s++) {
        double ss = s == 0 ? -1.0 : 1.0;
        for (int t = 0; t < 2; t++) {
            double st = t == 0 ? -1.0 : 1.0;
            cplx<double> f = cdiv(cadd(cmk(ss * mm, 0.0), cscale(ms, st)), msmm);
            g[s][t] = cmul(f, csub(cmul(dn[n], xi[n]), cscale(xip, ss * st)));
            A[s][t] = cmul(f, csub(cscale(dn[n], psi[n]), cmk(ss * st * psip, 0.0)));
        }
    }
    cplx<double> det = csub(cmul(g[0][0], g[1][1]), cmul(g[1][0], g[0][1]));
    cplx<double> gi[2][2] = {{cdiv(g[1][1], det), cdiv(cscale(g[0][1], -1.0), det)},
                             {cdiv(cscale(g[1][0], -1.0), det), cdiv(g[0][0], det)}};
    for (int s = 0; s < 2; s++)
        for (int t = 0; t < 2; t++) {
            cplx<double> v = cadd(cmul(gi[s][0], A[0][t]), cmul(gi[s][1], A[1][t]));
            a4[s + 2 * t] = cscale(v, -1.0);
        }
}

// One thread per sphere.  order[i] from the Q_ext convergence rule of
// mieoa (:5461-5470, 5563-5571); an[i*ldan + 4(n-1) + p + 2q] for n <= order.
__global__ void mie_kernel(const double *__restrict__ x, const double *__restrict__ mre, const double *__restrict__ mim,
                           double mm, int nsph, double qeps, int *__restrict__ order, cplx<double> *__restrict__ an,
                           int ldan)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsph) return;
    double xs = x[i];
    double xm = xs * mm;
    cplx<double> ms = cmk(mre[i], mim[i]);
    // nstop = nint(x + 4 x**(1./3.)) + 5, the exponent being a real(4) literal
    int nstop = (int)llround(xs + 4.0 * pow(xs, (double)(1.0f / 3.0f))) + 5;
    if (nstop > MIE_NMAX - 2) nstop = MIE_NMAX - 2;
    double psi[MIE_NMAX + 2];
    cplx<double> xi[MIE_NMAX + 2], dn[MIE_NMAX + 2];
    riccati_real(nstop + 1, xm, psi, xi);
    logderiv(nstop + 1, cscale(ms, xs), dn);
    double q1[MIE_NMAX], qt = 0.0;
    cplx<double> a4[4];
    for (int n = 1; n <= nstop; n++) {
        mie_degree(n, xm, ms, mm, psi, xi, dn, a4);
        q1[n - 1] = -(double)(2 * n + 1) * (a4[0].re + a4[3].re);
        qt += q1[n - 1];
    }
    int L = nstop;
    if (qeps > 0.0) {
        double q = 0.0;
        for (L = 1; L <= nstop; L++) {
            q += q1[L - 1];
            if (fabs(1.0 - q / qt) < qeps || L == nstop) break;
        }
    } else if (qeps < 0.0) {
        L = (int)ceil(-qeps);
    }
    order[i] = L;
    if (an == nullptr) return;
    for (int n = 1; n <= imn(L, ldan / 4); n++) {
        mie_degree(n, xm, ms, mm, psi, xi, dn, a4);
        for (int q = 0; q < 4; q++) an[(int64_t)i * ldan + 4 * (n - 1) + q] = a4[q];
    }
}

hipError_t launch_mie(const double *x, const double *mre, const double *mim, double mmed, int nsph, double qeps,
                      int *order, double *an, int ldan, hipStream_t st)
{
    int bs = 64;
    hipLaunchKernelGGL(mie_kernel, dim3((nsph + bs - 1) / bs), dim3(bs), 0, st, x, mre, mim, mmed, nsph, qeps, order,
                       reinterpret_cast<cplx<double> *>(an), ldan);
    return hipGetLastError();
}

// T-matrix application (multmiecoeffmult / onemiecoeffmult 'i',
// mstm-modules-33.f90:5585-5751): idir = 1 a(p) = sum_q T(p,q) g(q),
// otherwise the transpose; inverse = 1 uses T^{-1}.  One thread per (m,n).
__global__ void mie_mult_kernel(const cplx<double> *__restrict__ ain, cplx<double> *__restrict__ aout, int nsph,
                                const int *__restrict__ order, const int *__restrict__ offset,
                                const int *__restrict__ mieoff, const cplx<double> *__restrict__ an, int idir,
                                int inverse, int Lmax)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int NBM = Lmax * (Lmax + 2);
    if (q >= (int64_t)nsph * NBM) return;
    int i = (int)(q / NBM), e = (int)(q % NBM);
    int L = order[i], NB = L * (L + 2);
    if (e >= NB) return;
    int row = e % (L + 2), colm = e / (L + 2) + 1, n = (row <= colm) ? colm : row - 1;
    const cplx<double> *t = an + mieoff[i] + 4 * (n - 1);
    cplx<double> t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
    if (inverse) {
        cplx<double> det = csub(cmul(t0, t3), cmul(t1, t2));
        cplx<double> u0 = cdiv(t3, det), u3 = cdiv(t0, det), u1 = cdiv(cscale(t1, -1.0), det),
                     u2 = cdiv(cscale(t2, -1.0), det);
        t0 = u0;
        t1 = u1;
        t2 = u2;
        t3 = u3;
    }
    if (idir != 1) {
        cplx<double> tt = t1;
        t1 = t2;
        t2 = tt;
    }
    int64_t o = offset[i] + e;
    cplx<double> x0 = ain[o], x1 = ain[o + NB];
    aout[o] = cadd(cmul(t0, x0), cmul(t2, x1));
    aout[o + NB] = cadd(cmul(t1, x0), cmul(t3, x1));
}

hipError_t launch_mie_mult(const double *ain, double *aout, int nsph, const int *order, const int *offset,
                           const int *mieoff, const double *an, int idir, int inverse, int Lmax, hipStream_t st)
{
    int64_t n = (int64_t)nsph * Lmax * (Lmax + 2);
    int bs = 256;
    hipLaunchKernelGGL(mie_mult_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st,
                       reinterpret_cast<const cplx<double> *>(ain), reinterpret_cast<cplx<double> *>(aout), nsph, order,
                       offset, mieoff, reinterpret_cast<const cplx<double> *>(an), idir, inverse, Lmax);
    return hipGetLastError();
}

// Plane-wave VSWF coefficients (genplanewavecoef with lr_tran, :1453-1520)
// times the zero-boundary phase of layerplanewavecoef (:3901-3965), one
// thread per sphere.  pmnp(:,pol) columns of length ld (complex).
// excited (may be null): spheres with excited[i] == 0 get zero blocks, as
// sphereplanewavecoef's excited_spheres (mstm-scatprops-26.f90:262-264).
__global__ void planewave_kernel(cplx<double> *__restrict__ pmnp, int ld, int nsph, const double *__restrict__ pos,
                                 const int *__restrict__ order, const int *__restrict__ offset, double alpha,
                                 double sinc, int dir, double rimed, const int *__restrict__ excited)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsph) return;
    int L = order[i], NB = L * (L + 2);
    if (excited && !excited[i]) {
        cplx<double> *q0 = pmnp + offset[i], *q1 = pmnp + (int64_t)ld + offset[i];
        for (int k = 0; k < 2 * NB; k++) q0[k] = q1[k] = cmk(0.0, 0.0);
        return;
    }
    // dir = 1: incidence from below (cos beta > 0), dir = 2: from above (:3934)
    double cb = sqrt((1.0 - sinc / rimed) * (1.0 + sinc / rimed)) * (double)(3 - 2 * dir);
    const double *x = pos + 3 * i;
    double ph = sinc * (cos(alpha) * x[0] + sin(alpha) * x[1]) + rimed * cb * x[2];
    cplx<double> eph = cmk(cos(ph), sin(ph));
    // d^n_{k m}(beta) for k = -1, 0, 1 by the rotcoef recurrences
    double sbe = (fabs(cb) >= 1.0) ? 0.0 : sqrt(fabs((1.0 + cb) * (1.0 - cb)));
    double cbe2 = 0.5 * (1.0 + cb), sbe2 = 0.5 * (1.0 - cb);
    double dk0[2 * MIE_NMAX + 3], dk01[2 * MIE_NMAX + 3], r0[2 * MIE_NMAX + 1], r1[2 * MIE_NMAX + 1];
    const int o0 = MIE_NMAX + 1;
    for (int q = 0; q < 2 * MIE_NMAX + 3; q++) dk0[q] = dk01[q] = 0.0;
    dk0[o0] = 1.0;
    int sgn = 1;
    double sben = 1.0, bc = 1.0;
    cplx<double> *p0 = pmnp + offset[i], *p1 = pmnp + (int64_t)ld + offset[i];
    for (int n = 1; n <= L; n++) {
        sgn = -sgn;
        sben = sben * sbe / 2.0;
        if (sben < 1e-30) sben = 0.0;
        bc = bc * sqrt((double)(2 * n) * (double)(2 * n - 1)) / (double)n;
        dk0[o0 + n] = sgn * sben * bc;
        dk0[o0 - n] = sgn * dk0[o0 + n];
        dk01[o0 + n] = dk01[o0 - n] = 0.0;
        for (int k = -n + 1; k <= n - 1; k++) {
            double dkt = dk01[o0 + k];
            dk01[o0 + k] = dk0[o0 + k];
            dk0[o0 + k] = (cb * (double)(2 * n - 1) * dk01[o0 + k] - sqrt((double)(n - k - 1)) *
                                                                         sqrt((double)(n + k - 1)) * dkt) /
                          (sqrt((double)(n + k)) * sqrt((double)(n - k)));
        }
        for (int k = -n; k <= n; k++) r0[k + n] = dk0[o0 + k];
        // row m = 1 from row 0
        double fmn = 1.0 / sqrt((double)n) / sqrt((double)(n + 1));
        double dkm0 = 0.0, dkm1;
        for (int k = -n; k <= n; k++) {
            dkm1 = dkm0;
            dkm0 = r0[k + n];
            double dkn1 = (k == n) ? 0.0 : r0[k + 1 + n];
            r1[k + n] = (sqrt((double)(n + k)) * sqrt((double)(n - k + 1)) * cbe2 * dkm1 -
                         sqrt((double)(n - k)) * sqrt((double)(n + k + 1)) * sbe2 * dkn1 - (double)k * sbe * r0[k + n]) *
                        fmn;
        }
        double fnm = sqrt((double)(2 * n + 1) / 2.0) / 4.0;
        // cin = 4 i^(n+1)
        int ph4 = (n + 1) & 3;
        cplx<double> cin = ph4 == 0 ? cmk(4.0, 0.0) : ph4 == 1 ? cmk(0.0, 4.0) : ph4 == 2 ? cmk(-4.0, 0.0) : cmk(0.0, -4.0);
        for (int m = -n; m <= n; m++) {
            double dp1 = r1[m + n];
            // d(-1, m) = d(1, -m) (-1)^(m+1)
            double dm1 = r1[-m + n] * (((m + 1) & 1) ? -1.0 : 1.0);
            double t1 = -fnm * (-dm1 + dp1), t2 = -fnm * (dm1 + dp1);
            double tl[2] = {0.5 * (t1 + t2), 0.5 * (t1 - t2)};
            cplx<double> ea = cmk(cos((double)m * alpha), -sin((double)m * alpha));
            cplx<double> base = cmul(cmul(cin, ea), eph);
            int a = amn(m, n, L);
            for (int p = 0; p < 2; p++) {
                double sp = p == 0 ? 1.0 : -1.0;
                cplx<double> v = cscale(base, -tl[p]);
                p0[a + p * NB] = v;
                // sp * i * cin * tl * ...
                cplx<double> w = cscale(base, sp * tl[p]);
                p1[a + p * NB] = cmk(-w.im, w.re);
            }
        }
    }
}

hipError_t launch_planewave(double *pmnp, int ld, int nsph, const double *pos, const int *order, const int *offset,
                            double alpha, double sinc, int dir, double rimed, hipStream_t st, const int *excited)
{
    int bs = 64;
    hipLaunchKernelGGL(planewave_kernel, dim3((nsph + bs - 1) / bs), dim3(bs), 0, st,
                       reinterpret_cast<cplx<double> *>(pmnp), ld / 2, nsph, pos, order, offset, alpha, sinc, dir,
                       rimed, excited);
    return hipGetLastError();
}

// Per-sphere efficiencies (lrsphereqeff + qefficiencyfactors,
// mstm-scatprops-26.f90:676-870) for spheres in a real medium, two
// polarizations; g = T^{-1} a (onemiecoeffmult 'i').  qeff[i*9 + 3*k + q],
// k = (unpolarized, pol 1, pol 2), q = (ext, abs, sca(=qi)).
__global__ void qeff_kernel(const cplx<double> *__restrict__ amnp, const cplx<double> *__restrict__ pmnp, int ld,
                            int nsph, const double *__restrict__ radius, const int *__restrict__ order,
                            const int *__restrict__ offset, const int *__restrict__ mieoff,
                            const cplx<double> *__restrict__ an, double rim, double *__restrict__ qeff)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsph) return;
    int L = order[i], NB = L * (L + 2);
    double xs = radius[i], xr = xs * rim;
    double psi[MIE_NMAX + 2];
    cplx<double> xi[MIE_NMAX + 2];
    riccati_real(L, xr, psi, xi);
    double qe[3] = {0, 0, 0}, qa[3] = {0, 0, 0}, qi[3] = {0, 0, 0};
    double rib = rim;
    for (int n = 1; n <= L; n++) {
        double psip = psi[n - 1] - n * psi[n] / xr;
        cplx<double> xip = csub(xi[n - 1], cscale(xi[n], (double)n / xr));
        cplx<double> aam[2][2], agm[2][2];
        double ggm[2][2];
        double cterm = 1.0 / (rim * rim);  // i * conj(1/rim) / rim = i cterm
        for (int s = 0; s < 2; s++) {
            double ss = s == 0 ? -1.0 : 1.0;
            for (int t = 0; t < 2; t++) {
                double st = t == 0 ? -1.0 : 1.0;
                cplx<double> u = csub(cmul(xip, cconj(xi[n])), cscale(cmul(xi[n], cconj(xip)), ss * st));
                aam[s][t] = cmk(-cterm * u.im, cterm * u.re);
                double gv = psip * psi[n] - ss * st * psi[n] * psip;  // real
                ggm[s][t] = gv;                                        // times i cterm below
                cplx<double> w = csub(cscale(xip, psi[n]), cscale(xi[n], ss * st * psip));
                agm[s][t] = cmk(-cterm * w.im, cterm * w.re);
            }
        }
        const cplx<double> *tt = an + mieoff[i] + 4 * (n - 1);
        cplx<double> det = csub(cmul(tt[0], tt[3]), cmul(tt[1], tt[2]));
        cplx<double> ti[4] = {cdiv(tt[3], det), cdiv(cscale(tt[1], -1.0), det), cdiv(cscale(tt[2], -1.0), det),
                              cdiv(tt[0], det)};
        for (int m = -n; m <= n; m++) {
            int a0 = offset[i] + amn(m, n, L);
            cplx<double> A[2][2], G[2][2], F[2][2];  // [pol][mode]
            for (int pol = 0; pol < 2; pol++) {
                const cplx<double> *av = amnp + (int64_t)pol * ld, *fv = pmnp + (int64_t)pol * ld;
                A[pol][0] = av[a0];
                A[pol][1] = av[a0 + NB];
                F[pol][0] = fv[a0];
                F[pol][1] = fv[a0 + NB];
                G[pol][0] = cadd(cmul(ti[0], A[pol][0]), cmul(ti[2], A[pol][1]));
                G[pol][1] = cadd(cmul(ti[1], A[pol][0]), cmul(ti[3], A[pol][1]));
            }
            for (int p1 = 0; p1 < 2; p1++)
                for (int p2 = 0; p2 < 2; p2++) {
                    int k = (p1 == p2) ? p1 : 2;
                    double cst = (p1 == p2) ? 1.0 : 0.5;
                    for (int s = 0; s < 2; s++)
                        for (int t = 0; t < 2; t++) {
                            double fac = (((s + t) & 1) ? 0.0 : 2.0 * rib);  // conj(rib)+(-1)^(s+t) rib
                            // qi: i cterm ggm * fs * conj(rib ft)
                            cplx<double> v = cmul(F[p1][s], cconj(F[p2][t]));
                            qi[k] += cst * (-cterm * ggm[s][t] * rib * v.im);
                            cplx<double> a1 = cmul(aam[s][t], cmul(A[p1][s], cconj(A[p2][t])));
                            cplx<double> g1 = cmul(G[p1][s], cconj(G[p2][t]));
                            cplx<double> a2 = cmul(agm[s][t], cmul(A[p1][s], cconj(G[p2][t])));
                            qa[k] += cst * (a1.re * rib + (-cterm * ggm[s][t] * rib * g1.im) + a2.re * fac);
                            cplx<double> e2 = cmul(agm[s][t], cmul(A[p1][s], cconj(F[p2][t])));
                            qe[k] += cst * e2.re * fac;
                        }
                }
        }
    }
    double f = 2.0 / xs / xs;
    double qq[3][3];
    for (int k = 0; k < 3; k++) {
        qq[0][k] = qe[k] * f;
        qq[1][k] = qa[k] * f;
        qq[2][k] = qi[k] * f;
    }
    for (int q = 0; q < 3; q++) {
        qeff[i * 9 + 0 * 3 + q] = 0.5 * (qq[q][0] + qq[q][1]);
        qeff[i * 9 + 1 * 3 + q] = qq[q][0];
        qeff[i * 9 + 2 * 3 + q] = qq[q][1];
    }
}

hipError_t launch_qeff(const double *amnp, const double *pmnp, int ld, int nsph, const double *radius, const int *order,
                       const int *offset, const int *mieoff, const double *an, double rimed, double *qeff,
                       hipStream_t st)
{
    int bs = 64;
    hipLaunchKernelGGL(qeff_kernel, dim3((nsph + bs - 1) / bs), dim3(bs), 0, st,
                       reinterpret_cast<const cplx<double> *>(amnp), reinterpret_cast<const cplx<double> *>(pmnp),
                       ld / 2, nsph, radius, order, offset, mieoff, reinterpret_cast<const cplx<double> *>(an), rimed,
                       qeff);
    return hipGetLastError();
}
