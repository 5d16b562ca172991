// lu.hip -- dense interaction matrix and LU solution on MI355X.
//
// Replaces the direct solution branch of MSTM:
//   general_interaction_matrix   mstm-modules-33.f90:5854-5985
//       A = I - T H, column by column (multmiecoeffmult of each column of H);
//       here the blocks come straight from the pair tables when all spheres
//       have one order, else every column is the FP64 operator applied to a
//       unit vector;
//   lu_decomposition             mstm-solver-8.f90:951-1020
//       Crout LU with implicit row scaling: the pivot of column j is the row
//       maximizing vv(i) |a(i,j)| (vv(i) = 1 / max_k |a(i,k)| of the input
//       rows, moved with the row swaps; ties go to the last row), a zero
//       pivot becomes 1e-20.  Here the same pivot rule drives a blocked
//       factorization (panels of LU_NB columns, left-looking inside a group
//       of LU_GROUP panels, then the group's row swaps, block row and one
//       rank-LU_GROUP*LU_NB trailing update on the FP64 matrix cores);
//   lu_backsubstitution          mstm-solver-8.f90:1022-1052
//   direct_solver                mstm-solver-8.f90:571-642
//       solve, then niter refinement steps r = A x - p, x -= LU^{-1} r.
// Matrices are complex FP64, column-major with leading dimension n.
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

#define LU_NB 32
#define LU_GROUP 4  // panels per trailing update
#define LU_PANEL_THREADS 1024
#define LU_PANEL_ONE_WG_ROWS 512  // longer panels: pivot/pick/update launches per column
#define LU_SOLVE_THREADS 1024

namespace {

typedef cplx<double> zc;

__device__ inline double zabs(zc a) { return hypot(a.re, a.im); }

// unit vectors e_{j0+c} (c < ncol) in columns of 2*ld doubles, zero elsewhere
__global__ void lu_unit_kernel(double *x, int64_t ld, int n, int j0, int ncol)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)ncol * n) return;
    const int c = (int)(q / n), i = (int)(q - (int64_t)c * n);
    x[2 * ld * c + 2 * i] = (i == j0 + c) ? 1.0 : 0.0;
    x[2 * ld * c + 2 * i + 1] = 0.0;
}

// A(:, j0+c) = e_{j0+c} - (T H e_{j0+c}); th columns of 2*ld doubles
__global__ void lu_store_kernel(zc *A, int n, int j0, int ncol, const double *th, int64_t ld)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)ncol * n) return;
    const int c = (int)(q / n), i = (int)(q - (int64_t)c * n);
    const double *t = th + 2 * ld * c + 2 * i;
    A[(int64_t)(j0 + c) * n + i] = cmk((i == j0 + c ? 1.0 : 0.0) - t[0], -t[1]);
}

// vv(i) = 1 / max_k |a(i,k)|; a zero row sets *err
// 64 rows per workgroup; the four waves take every fourth column each
__global__ __launch_bounds__(256) void lu_rowscale_kernel(const zc *A, int n, double *vv, int *err)
{
    __shared__ double sm[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = blockIdx.x * 64 + lane;
    double m = 0.0;
    if (i < n) {
#pragma unroll 8
        for (int k = w; k < n; k += 4) m = fmax(m, zabs(A[(int64_t)k * n + i]));
    }
    sm[w][lane] = m;
    __syncthreads();
    if (w != 0 || i >= n) return;
    m = fmax(fmax(sm[0][lane], sm[1][lane]), fmax(sm[2][lane], sm[3][lane]));
    if (m == 0.0) {
        err[0] = 1;  // same value from every zero row
        vv[i] = 0.0;
    } else {
        vv[i] = 1.0 / m;
    }
}

// Factor columns c0..c0+nb-1 (rows c0..n-1) with the scaled pivot rule; the
// row swaps are applied inside the panel only and recorded in piv.
__global__ __launch_bounds__(LU_PANEL_THREADS) void lu_panel_kernel(zc *A, int n, int c0, int nb, double *vv,
                                                                    int *piv)
{
    __shared__ double sval[LU_PANEL_THREADS / 64];
    __shared__ int sidx[LU_PANEL_THREADS / 64];
    __shared__ int imax_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int jj = 0; jj < nb; jj++) {
        const int j = c0 + jj;
        zc *colj = A + (int64_t)j * n;
        // pivot search: max vv(i)|a(i,j)|, ties to the larger row (as '.ge.')
        double bv = -1.0;
        int bi = -1;
        for (int i = j + tid; i < n; i += LU_PANEL_THREADS) {
            const double d = vv[i] * zabs(colj[i]);
            if (d >= bv) {
                bv = d;
                bi = i;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(bv, off, 64);
            const int oi = __shfl_down(bi, off, 64);
            if (ov > bv || (ov == bv && oi > bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            sval[wv] = bv;
            sidx[wv] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double v = sval[0];
            int ix = sidx[0];
            for (int w = 1; w < LU_PANEL_THREADS / 64; w++)
                if (sval[w] > v || (sval[w] == v && sidx[w] > ix)) {
                    v = sval[w];
                    ix = sidx[w];
                }
            if (ix < j) ix = j;  // no finite candidate: keep row j
            imax_s = ix;
            piv[j] = ix;
            if (ix != j) vv[ix] = vv[j];
        }
        __syncthreads();
        const int im = imax_s;
        // swap rows j and im inside the panel
        if (im != j)
            for (int k = c0 + tid; k < c0 + nb; k += LU_PANEL_THREADS) {
                zc *ck = A + (int64_t)k * n;
                const zc t = ck[im];
                ck[im] = ck[j];
                ck[j] = t;
            }
        __syncthreads();
        if (tid == 0 && zabs(colj[j]) == 0.0) colj[j] = cmk(1e-20, 0.0);
        __syncthreads();
        const zc p = colj[j];
        const double den = p.re * p.re + p.im * p.im;
        const zc rcp = cmk(p.re / den, -p.im / den);  // 1 / a(j,j)
        for (int i = j + 1 + tid; i < n; i += LU_PANEL_THREADS) colj[i] = cmul(colj[i], rcp);
        __syncthreads();
        // rank-1 update of the panel's remaining columns
        for (int k = j + 1; k < c0 + nb; k++) {
            zc *ck = A + (int64_t)k * n;
            const zc u = ck[j];
            for (int i = j + 1 + tid; i < n; i += LU_PANEL_THREADS) ck[i] = csub(ck[i], cmul(colj[i], u));
        }
        __syncthreads();
    }
}

// The same panel factorization spread over the chip, per column j: (1) block
// partial maxima of vv(i)|a(i,j)| over rows >= j (a launch of its own only
// for a panel's first column: the update of column j-1 leaves them for the
// others), (2) one wave picks the pivot (ties to the larger row), swaps rows
// j and imax across the panel and keeps 1/a(j,j), (3) scaling of column j
// and the rank-1 update of the panel's remaining columns, one thread per row.
#define LU_PIV_BLOCKS 256
__global__ __launch_bounds__(256) void lu_pivot_kernel(const zc *A, int n, int j, const double *vv, double *pval,
                                                       int *pidx)
{
    __shared__ double sv[4];
    __shared__ int si[4];
    const zc *colj = A + (int64_t)j * n;
    double bv = -1.0;
    int bi = -1;
    for (int i = j + blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x) {
        const double d = vv[i] * zabs(colj[i]);
        if (d >= bv) {
            bv = d;
            bi = i;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64);
        const int oi = __shfl_down(bi, off, 64);
        if (ov > bv || (ov == bv && oi > bi)) {
            bv = ov;
            bi = oi;
        }
    }
    if ((threadIdx.x & 63) == 0) {
        sv[threadIdx.x >> 6] = bv;
        si[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++)
            if (sv[w] > bv || (sv[w] == bv && si[w] > bi)) {
                bv = sv[w];
                bi = si[w];
            }
        pval[blockIdx.x] = bv;
        pidx[blockIdx.x] = bi;
    }
}

__global__ __launch_bounds__(64) void lu_pick_kernel(zc *A, int n, int c0, int nb, int j, double *vv, int *piv,
                                                     const double *pval, const int *pidx, int nblk, zc *rcp)
{
    const int t = threadIdx.x;
    double v = -1.0;
    int ix = -1;  // no candidate
    for (int b = t; b < nblk; b += 64) {
        const double pv = pval[b];
        const int pi = pidx[b];
        if (pi >= 0 && (ix < 0 || pv > v || (pv == v && pi > ix))) {
            v = pv;
            ix = pi;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(v, off, 64);
        const int oi = __shfl_down(ix, off, 64);
        if (oi >= 0 && (ix < 0 || ov > v || (ov == v && oi > ix))) {
            v = ov;
            ix = oi;
        }
    }
    ix = __shfl(ix, 0, 64);
    if (ix < j) ix = j;  // no finite candidate: keep row j
    if (t == 0) {
        piv[j] = ix;
        if (ix != j) vv[ix] = vv[j];
    }
    // rows j and ix swap across the panel (nb <= 64 columns, one per lane);
    // the lane of column j keeps the pivot and its reciprocal
    const int k = c0 + t;
    if (t < nb) {
        zc *ck = A + (int64_t)k * n;
        zc a = ck[ix];
        if (ix != j) {
            ck[ix] = ck[j];
            ck[j] = a;
        }
        if (k == j) {
            if (zabs(a) == 0.0) {
                a = cmk(1e-20, 0.0);
                ck[j] = a;
            }
            const double den = a.re * a.re + a.im * a.im;
            rcp[0] = cmk(a.re / den, -a.im / den);
        }
    }
}

// When column j+1 is in the panel, the update also leaves the pivot
// partials of column j+1 over its rows (one per workgroup, as the pivot
// kernel would): the search of the next column needs no launch of its own.
__global__ __launch_bounds__(256) void lu_colupdate_kernel(zc *A, int n, int c0, int nb, int j, const zc *rcp,
                                                           const double *vv, double *pval, int *pidx)
{
    __shared__ double sv[4];
    __shared__ int si[4];
    __shared__ zc urow[LU_NB];  // row j of the columns after j in the panel
    const int i = j + 1 + blockIdx.x * 256 + threadIdx.x, nk = c0 + nb - j - 1;
    const bool ahead = nk > 0, mine = i < n;
    // every load of the row goes out before the barrier and before any store
    // (no thread here writes row j)
    zc x[LU_NB - 1], l = cmk(0.0, 0.0);
    if (mine) {
#pragma unroll
        for (int t = 0; t < LU_NB - 1; t++)
            if (t < nk) x[t] = A[(int64_t)(j + 1 + t) * n + i];
        l = cmul(A[(int64_t)j * n + i], rcp[0]);
    }
    if (threadIdx.x < nk) urow[threadIdx.x] = A[(int64_t)(j + 1 + threadIdx.x) * n + j];
    __syncthreads();
    double bv = -1.0;
    int bi = -1;
    if (mine) {
        A[(int64_t)j * n + i] = l;
#pragma unroll
        for (int t = 0; t < LU_NB - 1; t++)
            if (t < nk) {
                x[t] = csub(x[t], cmul(l, urow[t]));
                A[(int64_t)(j + 1 + t) * n + i] = x[t];
            }
        if (ahead) {
            bv = vv[i] * zabs(x[0]);
            bi = i;
        }
    }
    if (!ahead) return;
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64);
        const int oi = __shfl_down(bi, off, 64);
        if (ov > bv || (ov == bv && oi > bi)) {
            bv = ov;
            bi = oi;
        }
    }
    if ((threadIdx.x & 63) == 0) {
        sv[threadIdx.x >> 6] = bv;
        si[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++)
            if (sv[w] > bv || (sv[w] == bv && si[w] > bi)) {
                bv = sv[w];
                bi = si[w];
            }
        pval[blockIdx.x] = bv;
        pidx[blockIdx.x] = bi;
    }
}

// the panel's row swaps on columns [k0, k1)
__global__ void lu_swap_kernel(zc *A, int n, int c0, int nb, const int *piv, int k0, int k1)
{
    const int k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    zc *ck = A + (int64_t)k * n;
    for (int jj = 0; jj < nb; jj++) {
        const int j = c0 + jj, p = piv[j];
        if (p != j) {
            const zc t = ck[p];
            ck[p] = ck[j];
            ck[j] = t;
        }
    }
}

// block row U12 = L11^{-1} A12 for the columns [kb, ke) (L11 unit lower,
// LU_NB x LU_NB at (c0, c0): only full panels have a block row to their
// right), one column per thread
__global__ __launch_bounds__(256) void lu_trsm_kernel(zc *A, int n, int c0, int kb, int ke)
{
    __shared__ zc l11[LU_NB][LU_NB];
    for (int e = threadIdx.x; e < LU_NB * LU_NB; e += blockDim.x) {
        const int r = e % LU_NB, c = e / LU_NB;
        l11[r][c] = A[(int64_t)(c0 + c) * n + c0 + r];
    }
    __syncthreads();
    const int k = kb + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ke) return;
    zc *ck = A + (int64_t)k * n + c0;
    zc x[LU_NB];
#pragma unroll
    for (int r = 0; r < LU_NB; r++) x[r] = ck[r];
#pragma unroll
    for (int r = 1; r < LU_NB; r++) {
        zc s = x[r];
#pragma unroll
        for (int q = 0; q < r; q++) s = csub(s, cmul(l11[r][q], x[q]));
        x[r] = s;
    }
#pragma unroll
    for (int r = 0; r < LU_NB; r++) ck[r] = x[r];
}

// update A[rs:re, kb:ke] -= A[rs:re, c0:c0+kd] A[c0:c0+kd, kb:ke] (the
// trailing update A22 -= L21 U12) on the FP64 matrix cores: a 64x64 complex
// output tile per workgroup, each of the 4 waves a 32x32 quarter as 2x2
// v_mfma_f64_16x16x4f64 blocks with real and imaginary accumulators (four
// real products per complex one: re += Lr Ur - Li Ui, im += Lr Ui + Li Ur).
// The instruction forms D[i][j] (i = lane/16 + 4*item, j = lane%16) from
// A[i][k] and B[k][j] held as lane (i or j) + 16*k; it is applied to the
// transposed product U12^T L21^T so that j runs down the rows of A22 and a
// wave's read-modify-write of its block is 16 consecutive rows per column.
typedef double lu_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void lu_gemm_kernel(zc *A, int n, int c0, int kd, int rs, int re, int kb, int ke)
{
    constexpr int TK = 16;
    __shared__ zc Ls[TK][64 + 1];
    __shared__ zc Us[TK][64 + 1];
    const int r0 = rs + blockIdx.x * 64, k0 = kb + blockIdx.y * 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32, li = lane & 15, lk = lane >> 4;
    lu_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            cr[a][b] = lu_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = lu_d4{0.0, 0.0, 0.0, 0.0};
        }
    // the next K chunk is fetched into registers while this one is used (the
    // block updated never overlaps the L and U panels read)
    constexpr int PF = TK * 64 / 256;
    zc pl[PF], pu[PF];
    auto fetch = [&](int kk) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + 256 * u, q = e / 64, rr = e % 64;
            const int row = r0 + rr, col = k0 + rr;
            pl[u] = (row < re && kk + q < kd) ? A[(int64_t)(c0 + kk + q) * n + row] : cmk(0.0, 0.0);
            pu[u] = (col < ke && kk + q < kd) ? A[(int64_t)col * n + c0 + kk + q] : cmk(0.0, 0.0);
        }
    };
    fetch(0);
    for (int kk = 0; kk < kd; kk += TK) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + 256 * u;
            Ls[e / 64][e % 64] = pl[u];
            Us[e / 64][e % 64] = pu[u];
        }
        __syncthreads();
        if (kk + TK < kd) fetch(kk + TK);
#pragma unroll
        for (int q = 0; q < TK; q += 4) {
            double ur[2], ui[2], lr[2], lm[2], lim[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {  // a: column block (MFMA rows i)
                const zc u = Us[q + lk][wc + 16 * a + li];
                ur[a] = u.re;
                ui[a] = u.im;
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {  // b: row block (MFMA columns j)
                const zc l = Ls[q + lk][wr + 16 * b + li];
                lr[b] = l.re;
                lm[b] = l.im;
                lim[b] = -l.im;
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], lr[b], cr[a][b], 0, 0, 0);
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[a], lim[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], lm[b], ci[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[a], lr[b], ci[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int col = k0 + wc + 16 * a + lk + 4 * t, row = r0 + wr + 16 * b + li;
                if (row < re && col < ke) {
                    zc &o = A[(int64_t)col * n + row];
                    o = cmk(o.re - cr[a][b][t], o.im - ci[a][b][t]);
                }
            }
}

// LU x = P b for nrhs columns of b (2*ld doubles apart), one workgroup:
// the row swaps in order, forward substitution with unit L, back
// substitution with U (lu_backsubstitution)
__global__ __launch_bounds__(LU_SOLVE_THREADS) void lu_solve_kernel(const zc *LU, int n, const int *piv, double *b,
                                                                    int64_t ld, int nrhs)
{
    const int tid = threadIdx.x;
    if (tid < nrhs) {
        zc *x = reinterpret_cast<zc *>(b + 2 * ld * tid);
        for (int i = 0; i < n; i++) {
            const int p = piv[i];
            if (p != i) {
                const zc t = x[p];
                x[p] = x[i];
                x[i] = t;
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < n; j++) {
        const zc *cj = LU + (int64_t)j * n;
        for (int r = 0; r < nrhs; r++) {
            zc *x = reinterpret_cast<zc *>(b + 2 * ld * r);
            const zc xj = x[j];
            for (int i = j + 1 + tid; i < n; i += LU_SOLVE_THREADS) x[i] = csub(x[i], cmul(cj[i], xj));
        }
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; j--) {
        const zc *cj = LU + (int64_t)j * n;
        const zc d = cj[j];
        for (int r = 0; r < nrhs; r++) {
            zc *x = reinterpret_cast<zc *>(b + 2 * ld * r);
            const zc xj = cdiv(x[j], d);
            for (int i = tid; i < j; i += LU_SOLVE_THREADS) x[i] = csub(x[i], cmul(cj[i], xj));
            __syncthreads();
            if (tid == 0) x[j] = xj;
        }
        __syncthreads();
    }
}

// Blocked substitution on LU_NB-row blocks: lower = unit L forward over the
// block rows [r0, r0+nb), else U backward.
//
// One diagonal block (nb <= LU_NB rows) of the substitution: the triangle in
// LDS, one wave per two right-hand sides (lane = 32 * (rhs & 1) + row), the
// solved values passed between lanes by shuffles, so no barrier per column.
__global__ __launch_bounds__(128) void lu_tri_block_kernel(const zc *LU, int n, int r0, int nb, double *b, int64_t ld,
                                                           int nrhs, int lower)
{
    __shared__ zc T[LU_NB][LU_NB + 1];
    for (int e = threadIdx.x; e < LU_NB * LU_NB; e += 128) {
        const int ti = e % LU_NB, tj = e / LU_NB;
        T[ti][tj] = (ti < nb && tj < nb) ? LU[(int64_t)(r0 + tj) * n + r0 + ti] : cmk(0.0, 0.0);
    }
    const int lane = threadIdx.x & 63, i = lane & 31, base = lane & 32, r = (threadIdx.x >> 6) * 2 + (lane >> 5);
    const bool act = r < nrhs && i < nb;
    zc *x = reinterpret_cast<zc *>(b + 2 * ld * (r < nrhs ? r : 0));
    zc xi = act ? x[r0 + i] : cmk(0.0, 0.0);
    __syncthreads();
    if (lower) {  // unit lower
        for (int j = 0; j < nb; j++) {
            const zc xj = cmk(__shfl(xi.re, base + j, 64), __shfl(xi.im, base + j, 64));
            if (i > j) xi = csub(xi, cmul(T[i][j], xj));
        }
    } else {
        for (int j = nb - 1; j >= 0; j--) {
            zc xj = cmk(__shfl(xi.re, base + j, 64), __shfl(xi.im, base + j, 64));
            xj = cdiv(xj, T[j][j]);
            if (i < j)
                xi = csub(xi, cmul(T[i][j], xj));
            else if (i == j)
                xi = xj;
        }
    }
    if (act) x[r0 + i] = xi;
}

// The other rows take the block's solved values: each LU element is loaded
// once for every right-hand side (nrhs <= 4), the block's values sit in LDS.
__global__ __launch_bounds__(256) void lu_tri_update_kernel(const zc *LU, int n, int r0, int nb, double *b, int64_t ld,
                                                            int nrhs, int lower)
{
    __shared__ zc xs[4][LU_NB];
    for (int e = threadIdx.x; e < 4 * LU_NB; e += 256) {
        const int r = e / LU_NB, j = e % LU_NB;
        xs[r][j] = (r < nrhs && j < nb) ? reinterpret_cast<const zc *>(b + 2 * ld * r)[r0 + j] : cmk(0.0, 0.0);
    }
    __syncthreads();
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = lower ? r0 + nb + i0 : i0;
    if (lower ? i >= n : i >= r0) return;
    zc s[4];
#pragma unroll
    for (int r = 0; r < 4; r++) s[r] = r < nrhs ? reinterpret_cast<const zc *>(b + 2 * ld * r)[i] : cmk(0.0, 0.0);
    zc lij[LU_NB];  // the row's loads all in flight at once
#pragma unroll
    for (int j = 0; j < LU_NB; j++)
        if (j < nb) lij[j] = LU[(int64_t)(r0 + j) * n + i];
#pragma unroll
    for (int j = 0; j < LU_NB; j++)
        if (j < nb)
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < nrhs) s[r] = csub(s[r], cmul(lij[j], xs[r][j]));
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (r < nrhs) reinterpret_cast<zc *>(b + 2 * ld * r)[i] = s[r];
}

__global__ void lu_permute_kernel(const int *piv, int n, double *b, int64_t ld, int nrhs)
{
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    zc *x = reinterpret_cast<zc *>(b + 2 * ld * r);
    for (int i = 0; i < n; i++) {
        const int p = piv[i];
        if (p != i) {
            const zc t = x[p];
            x[p] = x[i];
            x[i] = t;
        }
    }
}

// r = A x - p (one thread per row)
__global__ void lu_resid_kernel(const zc *A, int n, const double *x, const double *p, double *r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const zc *xv = reinterpret_cast<const zc *>(x);
    zc s = cmk(0.0, 0.0);
    for (int k = 0; k < n; k++) cfma(s, A[(int64_t)k * n + i], xv[k]);
    r[2 * i] = s.re - p[2 * i];
    r[2 * i + 1] = s.im - p[2 * i + 1];
}

// unpack a model-2 address e of a mode block of order L into (m, n)
__device__ inline void lu_unpack(int e, int L, int &m, int &n)
{
    const int row = e % (L + 2), colm = e / (L + 2) + 1;
    if (row <= colm) {
        m = row;
        n = colm;
    } else {
        m = -colm;
        n = row - 1;
    }
}

// Direct build of A = I - T H for spheres of one order L (offsets 2 NB i):
// one workgroup per unordered pair (i, j) of the chunk, from its FP64 list
// table (target i, source j, full layout).  Element of the forward block
//   H_ij[(k,t),(k',q)] = e^{-ik phi} e^{ik' phi} sum_m d^t(m,k) A_m(t,q) d^q(m,k')
// and of the reverse block (the shift form of the same table)
//   H_ji[(k,t),(k',q)] = (-1)^{k+k'} e^{-ik phi} e^{ik' phi} sum_m d^t(m,-k) A_m(q,t) d^q(m,-k');
// H is diagonal in the mode, and A takes -T_target(t)[p][mode] H for both
// target modes p (multmiecoeffmult of the column, mstm-modules-33.f90:5976-5981).
__global__ __launch_bounds__(256) void lu_expand_kernel(const double *__restrict__ tab, int slot_elems,
                                                        const int2 *__restrict__ pairs, int L, int n,
                                                        const zc *__restrict__ an, zc *__restrict__ A)
{
    extern __shared__ double sh[];
    const int q = blockIdx.x;
    const double *grp = tab + (int64_t)(q / MSTM_SLOTS) * slot_elems * MSTM_SLOTS;
    const int ks = q % MSTM_SLOTS;
    for (int e = threadIdx.x; e < slot_elems; e += blockDim.x) sh[e] = grp[ilv(e, ks, 2)];
    __syncthreads();
    const int2 pr = pairs[q];
    const int NB = L * (L + 2);
    const int po = phi_off(L, L, 2);
    const double phi = atan2(sh[po + 1], sh[po]);
    for (int e = threadIdx.x; e < 4 * NB * NB; e += blockDim.x) {
        const int sc = e % NB, tr = (e / NB) % NB, mode = (e / (NB * NB)) & 1, dir = e / (2 * NB * NB);
        int kt, t, kq, qd;
        lu_unpack(tr, L, kt, t);
        lu_unpack(sc, L, kq, qd);
        const int mm = imn(t, qd);
        double sr = 0.0, si = 0.0;
        for (int m = -mm; m <= mm; m++) {
            const int b = (m == 0) ? 0 : (m < 0 ? -2 * m - 1 : 2 * m);
            const int m1 = imx(1, iab(m));
            const int e0 = sec_off(b, L, L, 2);
            const int ND = L - m1 + 1;
            const int ea = e0 + pad_to(rot_len(m, L, L), 2) + mode * pad_to(ax_len(m, L, L), 2);
            double dt, ds;
            int ia;
            if (!dir) {
                dt = sh[e0 + t * t - m1 * m1 + kt + t];
                ds = sh[e0 + qd * qd - m1 * m1 + kq + qd];
                ia = ea + 2 * ((t - m1) + ND * (qd - m1));
            } else {
                dt = sh[e0 + t * t - m1 * m1 - kt + t];
                ds = sh[e0 + qd * qd - m1 * m1 - kq + qd];
                ia = ea + 2 * ((qd - m1) + ND * (t - m1));
            }
            const double f = dt * ds;
            sr = fma(f, sh[ia], sr);
            si = fma(f, sh[ia + 1], si);
        }
        double sn, cs;
        sincos((double)(kq - kt) * phi, &sn, &cs);
        zc h = cmk(sr * cs - si * sn, sr * sn + si * cs);
        if (dir && ((kt + kq) & 1)) h = cmk(-h.re, -h.im);
        const int it = dir ? pr.y : pr.x, js = dir ? pr.x : pr.y;
        const zc *tt = an + ((int64_t)it * L + t - 1) * 4;  // T[p][q] = tt[2q + p]
        const int64_t col = (int64_t)js * 2 * NB + mode * NB + sc;
        for (int p = 0; p < 2; p++) {
            const zc v = cmul(tt[2 * mode + p], h);
            A[col * n + (int64_t)it * 2 * NB + p * NB + tr] = cmk(-v.re, -v.im);
        }
    }
}

__global__ void lu_identity_kernel(zc *A, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(int64_t)i * n + i] = cmk(1.0, 0.0);
}

}  // namespace

hipError_t launch_lu_expand(const double *tab, int slot_elems, const int2 *pairs, int npairs, int L, int n,
                            const double *an, double *A, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(lu_expand_kernel, dim3(npairs), dim3(256), sizeof(double) * slot_elems, st, tab, slot_elems,
                       pairs, L, n, reinterpret_cast<const zc *>(an), reinterpret_cast<zc *>(A));
    return hipGetLastError();
}

hipError_t launch_lu_identity(double *A, int n, hipStream_t st)
{
    hipLaunchKernelGGL(lu_identity_kernel, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<zc *>(A), n);
    return hipGetLastError();
}

hipError_t launch_lu_unit(double *x, int64_t ld, int n, int j0, int ncol, hipStream_t st)
{
    const int64_t tot = (int64_t)ncol * n;
    hipLaunchKernelGGL(lu_unit_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, ld, n, j0, ncol);
    return hipGetLastError();
}

hipError_t launch_lu_store(double *A, int n, int j0, int ncol, const double *th, int64_t ld, hipStream_t st)
{
    const int64_t tot = (int64_t)ncol * n;
    hipLaunchKernelGGL(lu_store_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<zc *>(A), n, j0, ncol, th, ld);
    return hipGetLastError();
}

// scratch: pval[LU_PIV_BLOCKS] doubles, rcp one complex, pidx[LU_PIV_BLOCKS] ints
size_t lu_factor_scratch_bytes() { return LU_PIV_BLOCKS * (sizeof(double) + sizeof(int)) + 2 * sizeof(double); }

hipError_t launch_lu_factor(double *LUm, int n, double *vv, int *piv, int *err, void *scratch, hipStream_t st)
{
    if (n < 1 || n > LU_PIV_BLOCKS * 256) return hipErrorInvalidValue;
    zc *A = reinterpret_cast<zc *>(LUm);
    double *pval = reinterpret_cast<double *>(scratch);
    zc *rcp = reinterpret_cast<zc *>(pval + LU_PIV_BLOCKS);
    int *pidx = reinterpret_cast<int *>(pval + LU_PIV_BLOCKS + 2);
    hipLaunchKernelGGL(lu_rowscale_kernel, dim3((n + 63) / 64), dim3(256), 0, st, A, n, vv, err);
    hipError_t e = hipGetLastError();
    // one panel of nb columns at c0: pivot search, row swaps and elimination
    // inside the panel's own columns
    auto panel = [&](int c0, int nb) {
        if (n - c0 <= LU_PANEL_ONE_WG_ROWS) {
            // short panels: one workgroup does all its columns
            hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_PANEL_THREADS), 0, st, A, n, c0, nb, vv, piv);
            return;
        }
        for (int jj = 0; jj < nb; jj++) {
            const int j = c0 + jj, rows = n - j;
            // partials of column j: its own search for the panel's first
            // column, else those the update of column j-1 left (one per
            // workgroup of that launch, (rows + 255) / 256 <= LU_PIV_BLOCKS)
            const int nbk = jj == 0 ? imn(LU_PIV_BLOCKS, (rows + 255) / 256) : (rows + 255) / 256;
            if (jj == 0) hipLaunchKernelGGL(lu_pivot_kernel, dim3(nbk), dim3(256), 0, st, A, n, j, vv, pval, pidx);
            hipLaunchKernelGGL(lu_pick_kernel, dim3(1), dim3(64), 0, st, A, n, c0, nb, j, vv, piv, pval, pidx, nbk,
                               rcp);
            if (rows > 1)
                hipLaunchKernelGGL(lu_colupdate_kernel, dim3((rows - 1 + 255) / 256), dim3(256), 0, st, A, n, c0, nb,
                                   j, rcp, vv, pval, pidx);
        }
    };
    auto swap = [&](int p0, int np, int k0, int k1) {  // pivots [p0, p0+np) on columns [k0, k1)
        if (k1 > k0)
            hipLaunchKernelGGL(lu_swap_kernel, dim3((k1 - k0 + 255) / 256), dim3(256), 0, st, A, n, p0, np, piv, k0,
                               k1);
    };
    auto trsm = [&](int c0, int k0, int k1) {
        if (k1 > k0) hipLaunchKernelGGL(lu_trsm_kernel, dim3((k1 - k0 + 255) / 256), dim3(256), 0, st, A, n, c0, k0, k1);
    };
    auto gemm = [&](int c0, int kd, int rs, int re, int k0, int k1) {
        if (re > rs && k1 > k0)
            hipLaunchKernelGGL(lu_gemm_kernel, dim3((re - rs + 63) / 64, (k1 - k0 + 63) / 64), dim3(256), 0, st, A, n,
                               c0, kd, rs, re, k0, k1);
    };
    // LU_GROUP panels per step, so that the update of the trailing matrix is
    // of rank LU_GROUP LU_NB: each panel of the group is first brought up to
    // date with the ones before it (their row swaps, its part of the block
    // row, one update of its own columns) and factored; then the rest of the
    // matrix takes the group's row swaps, its block row and one update
    for (int c0 = 0; c0 < n && e == hipSuccess; c0 += LU_GROUP * LU_NB) {
        const int cE = imn(n, c0 + LU_GROUP * LU_NB);
        for (int cg = c0; cg < cE; cg += LU_NB) {
            const int ce = imn(cE, cg + LU_NB);
            if (cg > c0) {  // the panels before this one are full
                swap(c0, cg - c0, cg, ce);
                for (int ch = c0; ch < cg; ch += LU_NB) {
                    trsm(ch, cg, ce);
                    gemm(ch, LU_NB, ch + LU_NB, cg, cg, ce);
                }
                gemm(c0, cg - c0, cg, n, cg, ce);
            }
            panel(cg, ce - cg);
            if (cg > c0) swap(cg, ce - cg, c0, cg);
        }
        swap(c0, cE - c0, 0, c0);
        if (cE < n) {  // every panel of the group is full here
            swap(c0, cE - c0, cE, n);
            for (int ch = c0; ch < cE; ch += LU_NB) {
                trsm(ch, cE, n);
                gemm(ch, LU_NB, ch + LU_NB, cE, cE, n);
            }
            gemm(c0, cE - c0, cE, n, cE, n);
        }
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_lu_solve(const double *LUm, int n, const int *piv, double *b, int64_t ld, int nrhs, hipStream_t st)
{
    if (nrhs < 1 || nrhs > 4) return hipErrorInvalidValue;
    const zc *LU = reinterpret_cast<const zc *>(LUm);
    if (n <= 2048) {  // small systems: one workgroup does it all
        hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(LU_SOLVE_THREADS), 0, st, LU, n, piv, b, ld, nrhs);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(lu_permute_kernel, dim3(1), dim3(64), 0, st, piv, n, b, ld, nrhs);
    for (int r0 = 0; r0 < n; r0 += LU_NB) {
        const int nb = n - r0 < LU_NB ? n - r0 : LU_NB, rest = n - r0 - nb;
        hipLaunchKernelGGL(lu_tri_block_kernel, dim3(1), dim3(128), 0, st, LU, n, r0, nb, b, ld, nrhs, 1);
        if (rest > 0)
            hipLaunchKernelGGL(lu_tri_update_kernel, dim3((rest + 255) / 256), dim3(256), 0, st, LU, n, r0, nb, b, ld,
                               nrhs, 1);
    }
    const int nblk = (n + LU_NB - 1) / LU_NB;
    for (int bk = nblk - 1; bk >= 0; bk--) {
        const int r0 = bk * LU_NB, nb = n - r0 < LU_NB ? n - r0 : LU_NB;
        hipLaunchKernelGGL(lu_tri_block_kernel, dim3(1), dim3(128), 0, st, LU, n, r0, nb, b, ld, nrhs, 0);
        if (r0 > 0)
            hipLaunchKernelGGL(lu_tri_update_kernel, dim3((r0 + 255) / 256), dim3(256), 0, st, LU, n, r0, nb, b, ld,
                               nrhs, 0);
    }
    return hipGetLastError();
}

hipError_t launch_lu_resid(const double *A, int n, const double *x, const double *p, double *r, hipStream_t st)
{
    hipLaunchKernelGGL(lu_resid_kernel, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<const zc *>(A), n,
                       x, p, r);
    return hipGetLastError();
}
