// lu.hip -- dense interaction matrix and LU solution on MI355X.
//
// Replaces the direct solution branch of MSTM:
//   general_interaction_matrix   mstm-modules-33.f90:5854-5985
//       A = I - T H, column by column (multmiecoeffmult of each column of H);
//       here every column is the FP64 operator applied to a unit vector;
//   lu_decomposition             mstm-solver-8.f90:951-1020
//       Crout LU with implicit row scaling: the pivot of column j is the row
//       maximizing vv(i) |a(i,j)| (vv(i) = 1 / max_k |a(i,k)| of the input
//       rows, moved with the row swaps; ties go to the last row), a zero
//       pivot becomes 1e-20.  Here the same pivot rule drives a blocked
//       right-looking factorization (panels of LU_NB columns, row swaps,
//       triangular solve of the block row, trailing update by GEMM);
//   lu_backsubstitution          mstm-solver-8.f90:1022-1052
//   direct_solver                mstm-solver-8.f90:571-642
//       solve, then niter refinement steps r = A x - p, x -= LU^{-1} r.
// Matrices are complex FP64, column-major with leading dimension n.
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

#define LU_NB 32
#define LU_PANEL_THREADS 1024
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
__global__ void lu_rowscale_kernel(const zc *A, int n, double *vv, int *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m = 0.0;
    for (int k = 0; k < n; k++) m = fmax(m, zabs(A[(int64_t)k * n + i]));
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

// block row U12 = L11^{-1} A12 (L11 unit lower, LU_NB x LU_NB: only full
// panels have a block row to their right), one column per thread
__global__ __launch_bounds__(256) void lu_trsm_kernel(zc *A, int n, int c0)
{
    __shared__ zc l11[LU_NB][LU_NB];
    for (int e = threadIdx.x; e < LU_NB * LU_NB; e += blockDim.x) {
        const int r = e % LU_NB, c = e / LU_NB;
        l11[r][c] = A[(int64_t)(c0 + c) * n + c0 + r];
    }
    __syncthreads();
    const int k = c0 + LU_NB + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
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

// trailing update A22 -= L21 U12: 64x64 output tiles, 256 threads x 4x4,
// K = nb staged in LDS in halves of 16
__global__ __launch_bounds__(256) void lu_gemm_kernel(zc *A, int n, int c0, int nb)
{
    constexpr int TK = 16;
    __shared__ zc Ls[TK][64 + 1];
    __shared__ zc Us[TK][64 + 1];
    const int r0 = c0 + nb + blockIdx.x * 64, k0 = c0 + nb + blockIdx.y * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    zc acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = cmk(0.0, 0.0);
    for (int kk = 0; kk < nb; kk += TK) {
        for (int e = threadIdx.x; e < TK * 64; e += 256) {
            const int q = e / 64, rr = e % 64;
            const int row = r0 + rr, col = k0 + rr;
            Ls[q][rr] = (row < n && kk + q < nb) ? A[(int64_t)(c0 + kk + q) * n + row] : cmk(0.0, 0.0);
            Us[q][rr] = (col < n && kk + q < nb) ? A[(int64_t)col * n + c0 + kk + q] : cmk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TK; q++) {
            zc lv[4], uv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) lv[a] = Ls[q][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; b++) uv[b] = Us[q][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) cfma(acc[a][b], lv[a], uv[b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int row = r0 + ty + 16 * a, col = k0 + tx + 16 * b;
            if (row < n && col < n) {
                zc &o = A[(int64_t)col * n + row];
                o = csub(o, acc[a][b]);
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

}  // namespace

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

hipError_t launch_lu_factor(double *LUm, int n, double *vv, int *piv, int *err, hipStream_t st)
{
    zc *A = reinterpret_cast<zc *>(LUm);
    hipLaunchKernelGGL(lu_rowscale_kernel, dim3((n + 255) / 256), dim3(256), 0, st, A, n, vv, err);
    hipError_t e = hipGetLastError();
    for (int c0 = 0; c0 < n && e == hipSuccess; c0 += LU_NB) {
        const int nb = n - c0 < LU_NB ? n - c0 : LU_NB;
        hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_PANEL_THREADS), 0, st, A, n, c0, nb, vv, piv);
        if (c0 > 0) hipLaunchKernelGGL(lu_swap_kernel, dim3((c0 + 255) / 256), dim3(256), 0, st, A, n, c0, nb, piv, 0, c0);
        const int rest = n - c0 - nb;
        if (rest > 0) {
            hipLaunchKernelGGL(lu_swap_kernel, dim3((rest + 255) / 256), dim3(256), 0, st, A, n, c0, nb, piv, c0 + nb, n);
            hipLaunchKernelGGL(lu_trsm_kernel, dim3((rest + 255) / 256), dim3(256), 0, st, A, n, c0);
            const int g = (rest + 63) / 64;
            hipLaunchKernelGGL(lu_gemm_kernel, dim3(g, g), dim3(256), 0, st, A, n, c0, nb);
        }
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_lu_solve(const double *LUm, int n, const int *piv, double *b, int64_t ld, int nrhs, hipStream_t st)
{
    if (nrhs < 1 || nrhs > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(LU_SOLVE_THREADS), 0, st, reinterpret_cast<const zc *>(LUm), n,
                       piv, b, ld, nrhs);
    return hipGetLastError();
}

hipError_t launch_lu_resid(const double *A, int n, const double *x, const double *p, double *r, hipStream_t st)
{
    hipLaunchKernelGGL(lu_resid_kernel, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<const zc *>(A), n,
                       x, p, r);
    return hipGetLastError();
}
