// tables.hip -- FP64 generation of rotate/axial/phase translation tables.
//
// One thread builds the table of one pair slot: Wigner-d rows (the
// recurrence of rotcoef, mstm-modules-33.f90:1098-1175), axial translation
// coefficients (seed sums + diamond recurrence of axialtrancoefrecurrence,
// :1566-1676) and exp(i phi) (cartosphere :1986-2004).  Arithmetic is FP64;
// the store type T is double (plain path) or float (mixed path).
#include "common.hpp"
#include "specfun.hpp"
#include "kernels.hpp"

// Where build_slot's values go: one table of store type T, or the FP64 and
// FP32 tables of the same slot at once (the same FP64 values, rounded for
// the second).  off(VW) is an element's offset in a layout of VW elements
// per 16 bytes; the slot is interleaved at k.
template <typename T> struct OneTable {
    T *__restrict__ grp;
    int k;
    template <typename F> __device__ __forceinline__ void put(F off, double v) const
    {
        constexpr int VW = 16 / (int)sizeof(T);
        grp[ilv(off(VW), k, VW)] = (T)v;
    }
};
struct BothTables {
    double *__restrict__ g64;
    float *__restrict__ g32;
    int k;
    template <typename F> __device__ __forceinline__ void put(F off, double v) const
    {
        g64[ilv(off(2), k, 2)] = v;
        g32[ilv(off(4), k, 4)] = (float)v;
    }
};

// Table of one directed translation (target order LT <- source order LS)
// along rtran = r_target - r_source, written interleaved for its slot.
// LMX >= LT, LS bounds the work arrays; the direct tables pass their order
// as LMX and as LT = LS, so that every loop has constant bounds and the
// arrays can live in registers instead of scratch memory.
template <int LMX, typename Sink>
__device__ __forceinline__ void build_slot(const Sink &out, int LT, int LS, int itype, const double *rt, double rimed,
                                           const double *__restrict__ vcc, int LM, bool sym = false)
{
    const int nmax = imx(LT, LS), nmin = imn(LT, LS);
    double r = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    double ct = rt[2] / r;
    double rho = sqrt(rt[0] * rt[0] + rt[1] * rt[1]);
    double cph = 1.0, sph = 0.0;
    if (!(rt[0] == 0.0 && rt[1] == 0.0)) {
        cph = rt[0] / rho;
        sph = rt[1] / rho;
    }
    // ---- Wigner d rows, degree by degree --------------------------------
    double sinb = (fabs(ct) >= 1.0) ? 0.0 : sqrt(fabs((1.0 + ct) * (1.0 - ct)));
    double cos2h = 0.5 * (1.0 + ct), sin2h = 0.5 * (1.0 - ct);
    double row_cur[2 * LMX + 3], row_prev[2 * LMX + 3];
    double rowp[2 * LMX + 1], rowc[2 * LMX + 1];
    const int mid = LMX + 1;
    for (int q = 0; q < 2 * LMX + 3; q++) row_cur[q] = row_prev[q] = 0.0;
    row_cur[mid] = 1.0;
    int sgn_n = 1;
    double sinpow = 1.0, binom = 1.0;  // binom = bcof(n,n) = sqrt((2n)!/(n!)^2)
    for (int n = 1; n <= nmax; n++) {
        sgn_n = -sgn_n;
        sinpow = sinpow * sinb / 2.0;
        if (sinpow < 1e-30) sinpow = 0.0;
        binom = binom * sqrt((double)(2 * n) * (double)(2 * n - 1)) / (double)n;
        row_cur[mid + n] = sgn_n * sinpow * binom;
        row_cur[mid - n] = sgn_n * row_cur[mid + n];
        row_prev[mid + n] = 0.0;
        row_prev[mid - n] = 0.0;
        for (int kk = -n + 1; kk <= n - 1; kk++) {
            double older = row_prev[mid + kk];
            row_prev[mid + kk] = row_cur[mid + kk];
            row_cur[mid + kk] = (ct * (double)(2 * n - 1) * row_prev[mid + kk] -
                                 sqrt((double)(n - kk - 1)) * sqrt((double)(n + kk - 1)) * older) /
                                (sqrt((double)(n + kk)) * sqrt((double)(n - kk)));
        }
        int mtop = imn(n, nmin);
        for (int kk = -n; kk <= n; kk++) rowp[kk + n] = row_cur[mid + kk];
        // write row m = 0
        auto put_row = [&](int m, const double *row, double sgnm) {
            int b = (m == 0) ? 0 : (m < 0 ? 2 * (-m) - 1 : 2 * m);
            int m1 = imx(1, iab(m));
            int r0 = 0;
            for (int nn = m1; nn < n; nn++) r0 += 2 * nn + 1;
            for (int kk = -n; kk <= n; kk++) {
                double v = (sgnm == 0.0) ? row[kk + n] : row[-kk + n] * sgnm * ((kk & 1) ? -1.0 : 1.0);
                out.put([&](int VW) { return sec_off(b, LT, LS, VW) + r0 + kk + n; }, v);
            }
        };
        // symmetric layout (LT = LS): half-sum / half-difference rows of block m >= 0
        auto put_sym = [&](int m, const double *row) {
            const int e0 = srow_off(m, n);
            for (int kk = 0; kk <= n; kk++) {
                double sg = (kk & 1) ? -1.0 : 1.0;
                double e = (kk == 0) ? row[n] : 0.5 * (row[kk + n] + sg * row[-kk + n]);
                out.put([&](int) { return e0 + kk; }, e);
                if (m > 0 && kk > 0) out.put([&](int) { return e0 + n + kk; }, 0.5 * (row[kk + n] - sg * row[-kk + n]));
            }
        };
        if (sym) put_sym(0, rowp);
        else put_row(0, rowp, 0.0);
        int sgn_m = 1;
        for (int m = 1; m <= mtop; m++) {
            sgn_m = -sgn_m;
            double rnorm = 1.0 / sqrt((double)(n - m + 1)) / sqrt((double)(n + m));
            double d_mid = 0.0, d_lo;
            for (int kk = -n; kk <= n; kk++) {
                d_lo = d_mid;
                d_mid = rowp[kk + n];
                double d_hi = (kk == n) ? 0.0 : rowp[kk + 1 + n];
                rowc[kk + n] = (sqrt((double)(n + kk)) * sqrt((double)(n - kk + 1)) * cos2h * d_lo -
                                sqrt((double)(n - kk)) * sqrt((double)(n + kk + 1)) * sin2h * d_hi -
                                (double)kk * sinb * rowp[kk + n]) *
                               rnorm;
            }
            if (sym) {
                put_sym(m, rowc);
            } else {
                put_row(m, rowc, 0.0);
                // d(-m, -k) = d(m, k) (-1)^(k+m)
                put_row(-m, rowc, (double)sgn_m);
            }
            for (int kk = -n; kk <= n; kk++) rowp[kk + n] = rowc[kk + n];
        }
    }
    // ---- axial coefficients ---------------------------------------------
    cplx<double> hw[2 * LMX + 1];
    const int nw = LT + LS;
    double z = r * rimed;
    if (itype == 3) {
        sph_hankel_up(nw, z, hw);
    } else {
        double jw[2 * LMX + 1];
        sph_bessel_j(nw, z, jw);
        for (int w = 0; w <= nw; w++) hw[w] = cmk(jw[w], 0.0);
    }
    cplx<double> act[(LMX + 1) * (LMX + 1)][2];
#define ACT(n, l) act[(n) * (LMX + 1) + (l)]
    for (int b = 0; b < 2 * nmin + 1; b++) {
        int m = mblock_m(b);
        if (sym && m < 0) continue;  // A_{-m} of one mode is A_m of the other
        int m1 = imx(1, iab(m));
        // seed: row n = m1 over all l, column l = LS over n > m1
        for (int pass = 0; pass < 2; pass++) {
            int cnt = pass == 0 ? (LS - m1 + 1) : (LT - m1);
            for (int q = 0; q < cnt; q++) {
                int n = pass == 0 ? m1 : m1 + 1 + q;
                int l = pass == 0 ? m1 + q : LS;
                int wmin = iab(n - l), wmax = n + l;
                cplx<double> se = cmk(0.0, 0.0), so = cmk(0.0, 0.0);
                for (int w = wmin; w <= wmax; w += 2) crfma(se, vcc[vcc_index(n, l, m, w, LM)], hw[w]);
                for (int w = wmin + 1; w <= wmax - 1; w += 2) crfma(so, vcc[vcc_index(n, l, m, w, LM)], hw[w]);
                // mode 1: se + i so ; mode 2: se - i so
                ACT(n, l)[0] = cmk(se.re - so.im, se.im + so.re);
                ACT(n, l)[1] = cmk(se.re + so.im, se.im - so.re);
            }
        }
        if (m1 < nmin) {
            for (int n = m1; n <= LT - 1; n++) {
                double fpn = ax_fnp1(m, n), fnn = ax_fn(m, n), fmn = ax_fnm1(m, n);
                for (int p = 0; p < 2; p++) {
                    double sp = (p == 0) ? 1.0 : -1.0;
                    for (int l = m1; l <= LS - 1; l++) {
                        cplx<double> v = cscale(ACT(n, l + 1)[p], -ax_fnp1(m, l));
                        double f = sp * (ax_fn(m, l) - fnn);
                        cplx<double> a = ACT(n, l)[p];
                        v = cadd(v, cmk(-f * a.im, f * a.re));  // f * i * a
                        if (l >= m1 + 1) v = cadd(v, cscale(ACT(n, l - 1)[p], ax_fnm1(m, l)));
                        if (n > m1) v = cadd(v, cscale(ACT(n - 1, l)[p], fmn));
                        ACT(n + 1, l)[p] = cscale(v, 1.0 / fpn);
                    }
                }
            }
        }
        // block start and mode stride in a layout of VW elements per 16 bytes
        auto ax0 = [&](int VW) {
            return sym ? sax_off(m, LT, VW) : sec_off(b, LT, LS, VW) + pad_to(rot_len(m, LT, LS), VW);
        };
        auto axp = [&](int VW) { return sym ? pad_to(sax_len(m, LT), VW) : pad_to(ax_len(m, LT, LS), VW); };
        for (int p = 0; p < (sym ? sax_modes(m) : 2); p++)
            for (int l = m1; l <= LS; l++)
                for (int n = m1; n <= LT; n++) {
                    const int e = 2 * ((n - m1) + (LT - m1 + 1) * (l - m1));
                    out.put([&](int VW) { return ax0(VW) + p * axp(VW) + e; }, ACT(n, l)[p].re);
                    out.put([&](int VW) { return ax0(VW) + p * axp(VW) + e + 1; }, ACT(n, l)[p].im);
                }
    }
#undef ACT
    auto ep = [&](int VW) { return sym ? sphi_off(LT, VW) : phi_off(LT, LS, VW); };
    out.put(ep, cph);
    out.put([&](int VW) { return ep(VW) + 1; }, sph);
}

// Direct-path tables: slot q = (tile*S + step)*S + k (S = MSTM_SLOTS) holds
// the pair target = blockI*S + k, source = blockJ*S + (k+step)%S (diagonal
// tiles have blockJ = blockI); self pairs and padding spheres get zeros.
// LMX: the order when it is one of the compiled ones (1..6), else
// MSTM_TAB_L with the order L taken at run time
template <typename T, int LMX>
__global__ void direct_tables_kernel(T *__restrict__ tab, const int2 *__restrict__ tiles, int ntiles, int L_, int nsph,
                                     const double *__restrict__ pos, double rimed, const double *__restrict__ vcc,
                                     int LM, int64_t slot_elems)
{
    const int L = LMX < MSTM_TAB_L ? LMX : L_;
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)ntiles * MSTM_SLOTS * MSTM_SLOTS) return;
    int k = (int)(q % MSTM_SLOTS);
    int64_t g = q / MSTM_SLOTS;
    int step = (int)(g % MSTM_SLOTS);
    int tile = (int)(g / MSTM_SLOTS);
    int2 bt = tiles[tile];
    int it = bt.x * MSTM_SLOTS + k;
    int is = bt.y * MSTM_SLOTS + (k + step) % MSTM_SLOTS;
    T *grp = tab + g * slot_elems * MSTM_SLOTS;
    const int VW = 16 / (int)sizeof(T);
    if (it >= nsph || is >= nsph || it == is) {
        for (int e = 0; e < slot_elems; e++) grp[ilv(e, k, VW)] = (T)0;
        return;
    }
    double rt[3] = {pos[3 * it] - pos[3 * is], pos[3 * it + 1] - pos[3 * is + 1], pos[3 * it + 2] - pos[3 * is + 2]};
    build_slot<LMX>(OneTable<T>{grp, k}, L, L, 3, rt, rimed, vcc, LM, direct_sym(L, (int)sizeof(T)));
}

// Both direct tables, FP64 and FP32, in one pass (orders with the same
// layout kind at both precisions: symmetric up to 4, full from 6): each
// slot's FP64 values are worked out once and stored twice.
template <int LMX>
__global__ void direct_tables_both_kernel(double *__restrict__ tab64, float *__restrict__ tab32,
                                          const int2 *__restrict__ tiles, int ntiles, int L_, int nsph,
                                          const double *__restrict__ pos, double rimed,
                                          const double *__restrict__ vcc, int LM, int64_t se64, int64_t se32)
{
    const int L = LMX < MSTM_TAB_L ? LMX : L_;
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)ntiles * MSTM_SLOTS * MSTM_SLOTS) return;
    int k = (int)(q % MSTM_SLOTS);
    int64_t g = q / MSTM_SLOTS;
    int step = (int)(g % MSTM_SLOTS);
    int tile = (int)(g / MSTM_SLOTS);
    int2 bt = tiles[tile];
    int it = bt.x * MSTM_SLOTS + k;
    int is = bt.y * MSTM_SLOTS + (k + step) % MSTM_SLOTS;
    double *g64 = tab64 + g * se64 * MSTM_SLOTS;
    float *g32 = tab32 + g * se32 * MSTM_SLOTS;
    if (it >= nsph || is >= nsph || it == is) {
        for (int e = 0; e < se64; e++) g64[ilv(e, k, 2)] = 0.0;
        for (int e = 0; e < se32; e++) g32[ilv(e, k, 4)] = 0.0f;
        return;
    }
    double rt[3] = {pos[3 * it] - pos[3 * is], pos[3 * it + 1] - pos[3 * is + 1], pos[3 * it + 2] - pos[3 * is + 2]};
    build_slot<LMX>(BothTables{g64, g32, k}, L, L, 3, rt, rimed, vcc, LM, direct_sym(L, 8));
}

template <typename T>
hipError_t launch_direct_tables(T *tab, const int2 *tiles, int ntiles, int L, int nsph, const double *pos, double rimed,
                                const double *vcc, int LM, hipStream_t st)
{
    if (L < 1 || L > MSTM_TAB_L) return hipErrorInvalidValue;  // build_slot's arrays
    int64_t se = direct_slot_len(L, (int)sizeof(T));
    int64_t nq = (int64_t)ntiles * MSTM_SLOTS * MSTM_SLOTS;
    int bs = 128;
    int64_t nb = (nq + bs - 1) / bs;
    switch (L) {
#define DT(LL)                                                                                                         \
    case LL:                                                                                                           \
        hipLaunchKernelGGL((direct_tables_kernel<T, LL>), dim3((unsigned)nb), dim3(bs), 0, st, tab, tiles, ntiles, L,   \
                           nsph, pos, rimed, vcc, LM, se);                                                             \
        break;
        DT(1) DT(2) DT(3) DT(4) DT(5) DT(6)
#undef DT
    default:
        hipLaunchKernelGGL((direct_tables_kernel<T, MSTM_TAB_L>), dim3((unsigned)nb), dim3(bs), 0, st, tab, tiles,
                           ntiles, L, nsph, pos, rimed, vcc, LM, se);
    }
    return hipGetLastError();
}
template hipError_t launch_direct_tables<float>(float *, const int2 *, int, int, int, const double *, double,
                                                const double *, int, hipStream_t);
template hipError_t launch_direct_tables<double>(double *, const int2 *, int, int, int, const double *, double,
                                                 const double *, int, hipStream_t);

bool direct_tables_both_ok(int L) { return L >= 1 && L <= MSTM_TAB_L && direct_sym(L, 8) == direct_sym(L, 4); }

hipError_t launch_direct_tables_both(double *tab64, float *tab32, const int2 *tiles, int ntiles, int L, int nsph,
                                     const double *pos, double rimed, const double *vcc, int LM, hipStream_t st)
{
    if (!direct_tables_both_ok(L)) return hipErrorInvalidValue;
    const int64_t nq = (int64_t)ntiles * MSTM_SLOTS * MSTM_SLOTS;
    const int bs = 128;
    const dim3 g((unsigned)((nq + bs - 1) / bs));
    const int64_t s64 = direct_slot_len(L, 8), s32 = direct_slot_len(L, 4);
    switch (L) {
#define DT(LL)                                                                                                         \
    case LL:                                                                                                           \
        hipLaunchKernelGGL((direct_tables_both_kernel<LL>), g, dim3(bs), 0, st, tab64, tab32, tiles, ntiles, L, nsph,   \
                           pos, rimed, vcc, LM, s64, s32);                                                             \
        break;
        DT(1) DT(2) DT(3) DT(4) DT(6)
#undef DT
    default:
        hipLaunchKernelGGL((direct_tables_both_kernel<MSTM_TAB_L>), g, dim3(bs), 0, st, tab64, tab32, tiles, ntiles, L,
                           nsph, pos, rimed, vcc, LM, s64, s32);
    }
    return hipGetLastError();
}

// General slot list (FFT path: sphere<->node J tables and near-field H
// tables).  For slot q: target/source orders from (LT,LS), rtran given.
template <typename T>
__global__ void list_tables_kernel(T *__restrict__ tab, const double *__restrict__ rtran, int nslots, int LT, int LS,
                                   int itype, double rimed, const double *__restrict__ vcc, int LM, int64_t slot_elems)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslots) return;
    int k = q % MSTM_SLOTS;
    int64_t g = q / MSTM_SLOTS;
    T *grp = tab + g * slot_elems * MSTM_SLOTS;
    const int VW = 16 / (int)sizeof(T);
    const double *rt = rtran + 3 * (int64_t)q;
    double r2 = rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2];
    if (r2 < 1e-12) {  // zero translation handled by the caller (identity / none)
        for (int e = 0; e < slot_elems; e++) grp[ilv(e, k, VW)] = (T)0;
        return;
    }
    build_slot<MSTM_TAB_L>(OneTable<T>{grp, k}, LT, LS, itype, rt, rimed, vcc, LM);
}

template <typename T>
hipError_t launch_list_tables(T *tab, const double *rtran, int nslots, int LT, int LS, int itype, double rimed,
                              const double *vcc, int LM, hipStream_t st)
{
    if (LT < 1 || LS < 1 || LT > MSTM_TAB_L || LS > MSTM_TAB_L) return hipErrorInvalidValue;  // build_slot's arrays
    const int VW = 16 / (int)sizeof(T);
    int64_t se = slot_len(LT, LS, VW);
    int bs = 128;
    int nb = (nslots + bs - 1) / bs;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(list_tables_kernel<T>, dim3(nb), dim3(bs), 0, st, tab, rtran, nslots, LT, LS, itype, rimed, vcc,
                       LM, se);
    return hipGetLastError();
}
template hipError_t launch_list_tables<float>(float *, const double *, int, int, int, int, double, const double *, int,
                                              hipStream_t);
template hipError_t launch_list_tables<double>(double *, const double *, int, int, int, int, double, const double *,
                                               int, hipStream_t);
