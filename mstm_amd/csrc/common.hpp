// common.hpp -- shared device/host helpers for the MI355X MSTM hot path.
//
// Coefficient layout (matches MSTM's dense "model 2" packing so the Fortran
// host can pass its arrays straight through, mstm-modules-33.f90:2351-2377):
// a sphere of order L owns 2*L*(L+2) complex values, mode p (L/R helicity)
// major; inside a mode block (m,n) sits at
//     m >= 0 : (n-1)(L+2) + m          m < 0 : (-m-1)(L+2) + n + 1
// Inside kernels vectors are re-indexed "naturally" as nk = n(n+1)+k-1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MSTM_MAX_L 10  // largest sphere / node order with a compiled tile kernel
// largest order of a generated translation table (build_slot's per-thread
// arrays are sized by it): sphere orders of the pair tables and the node order
// of the FFT path's sphere<->node tables
#define MSTM_TAB_L 12
#define MSTM_SLOTS 16  // pair slots per tile step (and spheres per block)

#define MSTM_HD __host__ __device__ __forceinline__

template <typename T> struct cplx {
    T re, im;
};
template <typename T> MSTM_HD cplx<T> cmk(T r, T i) { return cplx<T>{r, i}; }
template <typename T> MSTM_HD cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> MSTM_HD cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> MSTM_HD cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> MSTM_HD cplx<T> cscale(cplx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T> MSTM_HD cplx<T> cconj(cplx<T> a) { return {a.re, -a.im}; }
// a += b*c
template <typename T> MSTM_HD void cfma(cplx<T> &a, cplx<T> b, cplx<T> c)
{
    a.re = fma(b.re, c.re, a.re);
    a.re = fma(-b.im, c.im, a.re);
    a.im = fma(b.re, c.im, a.im);
    a.im = fma(b.im, c.re, a.im);
}
// a += s*c (real s)
template <typename T> MSTM_HD void crfma(cplx<T> &a, T s, cplx<T> c)
{
    a.re = fma(s, c.re, a.re);
    a.im = fma(s, c.im, a.im);
}
template <typename T> MSTM_HD cplx<T> cdiv(cplx<T> a, cplx<T> b)
{
    T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

MSTM_HD constexpr int imx(int a, int b) { return a > b ? a : b; }
MSTM_HD constexpr int imn(int a, int b) { return a < b ? a : b; }
MSTM_HD constexpr int iab(int a) { return a < 0 ? -a : a; }

// MSTM model-2 address inside one mode block, 0-based
MSTM_HD constexpr int amn(int m, int n, int L) { return m >= 0 ? (n - 1) * (L + 2) + m : (-m - 1) * (L + 2) + n + 1; }
// natural index n(n+1)+k-1
MSTM_HD constexpr int nk(int n, int k) { return n * (n + 1) + k - 1; }
// m of the b-th axial block in the order 0,-1,1,-2,2,...
MSTM_HD constexpr int mblock_m(int b) { return b == 0 ? 0 : ((b & 1) ? -((b + 1) >> 1) : (b >> 1)); }

// ---------------------------------------------------------------------------
// Pair-table layout for a directed translation target order LT <- source LS
// (tables are shared by both directions of an unordered pair).  For each
// axial block m (order 0,-1,1,...,-nmin,nmin), m1 = max(1,|m|):
//   ROT  d^n(m,k), n = m1..max(LT,LS), k = -n..n           (reals)
//   AX   mode p=0,1: A(t,s), t = m1..LT fastest, s = m1..LS (complex)
// each section padded to the vector width VW (reals); then PHI (cos,sin).
// A(t,s) is the axial coefficient act(t,s,p) of axialtrancoefrecurrence
// (mstm-modules-33.f90:1566-1676) built with nmax = LT, lmax = LS.
// ---------------------------------------------------------------------------
MSTM_HD constexpr int pad_to(int x, int v) { return (x + v - 1) / v * v; }
MSTM_HD constexpr int rot_len(int m, int LT, int LS)
{
    int m1 = imx(1, iab(m)), s = 0;
    for (int n = m1; n <= imx(LT, LS); n++) s += 2 * n + 1;
    return s;
}
MSTM_HD constexpr int ax_len(int m, int LT, int LS)  // reals, one mode
{
    int m1 = imx(1, iab(m));
    return 2 * (LT - m1 + 1) * (LS - m1 + 1);
}
MSTM_HD constexpr int sec_len(int m, int LT, int LS, int VW)
{
    return pad_to(rot_len(m, LT, LS), VW) + 2 * pad_to(ax_len(m, LT, LS), VW);
}
MSTM_HD constexpr int sec_off(int b, int LT, int LS, int VW)
{
    int o = 0;
    for (int q = 0; q < b; q++) o += sec_len(mblock_m(q), LT, LS, VW);
    return o;
}
MSTM_HD constexpr int nmblocks(int LT, int LS) { return 2 * imn(LT, LS) + 1; }
MSTM_HD constexpr int phi_off(int LT, int LS, int VW) { return sec_off(nmblocks(LT, LS), LT, LS, VW); }
MSTM_HD constexpr int slot_len(int LT, int LS, int VW) { return phi_off(LT, LS, VW) + VW; }
// ---------------------------------------------------------------------------
// Symmetric pair-table layout of the direct tiles (LT = LS = L).  Two
// symmetries halve the table:
//   d^n(-m,-k) = (-1)^{m+k} d^n(m,k)      (Wigner d)
//   A_{-m}(t,s) of mode p = A_m(t,s) of the other mode (axial coefficients;
//   for m = 0 both modes coincide).
// Only blocks m = 0..L are stored:
//   ROT  degree by degree, n = 1..L; degree n holds the rows of blocks
//        m = 0..n: the half-sum and half-difference rows
//          E(n,0) = d^n(m,0),  E(n,k) = (d^n(m,k) + (-1)^k d^n(m,-k)) / 2,
//          O(n,k) = (d^n(m,k) - (-1)^k d^n(m,-k)) / 2        (k = 1..n)
//        stored E(n,0..n) then O(n,1..n); for m = 0 only E (= d^n(0,k));
//        the section is padded once, to VW reals
//   AX   per block m = 0..L: A_m(t,s), t fastest, for mode 0 and (m > 0)
//        mode 1, each padded to VW reals
// then PHI (cos, sin).  With u_k = x_k + (-1)^k x_{-k},
// v_k = x_k - (-1)^k x_{-k} (k > 0), u_0 = x_0, the rotation
// y_m = sum_k d(m,k) x_k of blocks m and -m becomes
//   Y+ = sum_{k>=0} E_k u_k,  Y- = sum_{k>0} O_k v_k,
//   y_m = Y+ + Y-,  y_{-m} = (-1)^m (Y+ - Y-)
// and the back rotation accumulates P_k = sum E_k (z_m + (-1)^m z_{-m}),
// Q_k = sum O_k (z_m - (-1)^m z_{-m}), from which U_k = P_k + Q_k,
// U_{-k} = (-1)^k (P_k - Q_k), U_0 = P_0.  Both rotations of a degree read
// the same contiguous rows, so a tile step runs degree by degree (forward),
// block by block (axial), degree by degree (back).
// ---------------------------------------------------------------------------
MSTM_HD constexpr int sdeg_len(int n) { return (n + 1) + n * (2 * n + 1); }
MSTM_HD constexpr int sdeg_off(int n)
{
    int s = 0;
    for (int q = 1; q < n; q++) s += sdeg_len(q);
    return s;
}
// row of block m (0 <= m <= n) of degree n inside the ROT section
MSTM_HD constexpr int srow_off(int m, int n) { return sdeg_off(n) + (m == 0 ? 0 : (n + 1) + (m - 1) * (2 * n + 1)); }
MSTM_HD constexpr int srow_len(int m, int n) { return m == 0 ? n + 1 : 2 * n + 1; }
MSTM_HD constexpr int srot_tot(int L) { return sdeg_off(L + 1); }
MSTM_HD constexpr int sax_len(int m, int L) { return 2 * (L - imx(1, m) + 1) * (L - imx(1, m) + 1); }  // one mode
MSTM_HD constexpr int sax_modes(int m) { return m == 0 ? 1 : 2; }
MSTM_HD constexpr int sax_off(int m, int L, int VW)
{
    int o = pad_to(srot_tot(L), VW);
    for (int q = 0; q < m; q++) o += sax_modes(q) * pad_to(sax_len(q, L), VW);
    return o;
}
MSTM_HD constexpr int sphi_off(int L, int VW) { return sax_off(L + 1, L, VW); }
MSTM_HD constexpr int sslot_len(int L, int VW) { return sphi_off(L, VW) + VW; }
// the direct tiles use the symmetric layout, staged in LDS, up to order 5
// (FP32) / 4 (FP64); larger orders keep the full layout, whose unstaged form
// folds the output phase into every term and so needs no P/Q accumulators
MSTM_HD constexpr bool direct_sym(int L, int elem)
{
    return L <= (elem == 4 ? 5 : 4) && (int64_t)sslot_len(L, 16 / elem) * MSTM_SLOTS * elem <= 65536;
}
MSTM_HD constexpr int direct_slot_len(int L, int elem)
{
    return direct_sym(L, elem) ? sslot_len(L, 16 / elem) : slot_len(L, L, 16 / elem);
}

// interleaved position of element e (multiple-of-VW chunks) of slot k
MSTM_HD constexpr int64_t ilv(int64_t e, int k, int VW) { return ((e / VW) * MSTM_SLOTS + k) * VW + e % VW; }
