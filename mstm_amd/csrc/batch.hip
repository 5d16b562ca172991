// batch.hip -- many clusters x many right-hand sides: one solve for a list of
// independent clusters (uneven sphere counts), each with the same list of
// incidence directions, both polarizations.
//
// MSTM solves one cluster and one right-hand side pair at a time
// (fixedorsoln, mstm-solver-8.f90:350-569, called per incidence; the
// T-matrix path loops over 2(2l+1) right-hand sides, :22-332).  Here the
// interaction operator H of every cluster is built once as dense blocks
// (FP64 build from the same rotate/axial/phase tables as the direct path,
// stored complex FP64 and complex FP32), so that applying it to all
// right-hand sides of a cluster is a complex GEMM:
//   * H is block-diagonal in the helicity mode p (the axial translation of
//     axialtrancoefrecurrence, mstm-modules-33.f90:1566-1676, acts per mode),
//     so each cluster holds two M x M matrices, M = N L(L+2);
//   * the BiCG of cbicg (mstm-solver-8.f90:654-949) runs on all columns
//     together, each column with its own scalars; both of its products
//     (A p and A^H w) are one grouped FP32 GEMM launch on MFMA
//     (v_mfma_f32_32x32x2f32, FP32 accumulation) over a work list of
//     (cluster, mode, operation, row tile) jobs taken by a persistent grid;
//   * the FP32 inner solve sits in FP64 iterative refinement whose residual
//     uses the FP64 operator; the inner iterations are replayed as a HIP
//     graph;
//   * the persistent grid size and job order are picked by timing the
//     candidates on the actual operator (autotune).
// Vector layout ("batched"): per cluster c, mode p, row = sphere*NB + packed
// (m,n) index (MSTM's a(0:L+1,1:L) packing), R columns contiguous per row;
// column 2*d + q is direction d, polarization q.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mstm_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int BR = 64;  // largest number of right-hand sides per cluster

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef cplx<float> Z32;
typedef cplx<double> Z64;

struct GemmJob {
    int c, mode, op, rt;
};
struct RowInfo {
    int v0, v1;  // batched rows of modes 0 and 1
    int an;      // complex offset of the sphere's T block of this degree
    int c;       // cluster
    int mi, nb;  // MSTM-layout index of the mode-0 element, block size L(L+2)
};
struct Chunk {
    int g0, g1;  // global mode-0 rows [g0, g1) of one cluster
};

// packed index e of a mode block of order L -> (m, n)
__device__ __forceinline__ void unpack_mn(int e, int L, int &m, int &n)
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

// ---------------------------------------------------------------------------
// Dense blocks of one unordered pair (i < j) from its table (launch_list_tables
// layout, target i, source j, orders L): forward block H_ij (rows i, cols j)
// and reverse block H_ji, both modes, FP64, written as FP64 and FP32.
//   H_ij[(k,t),(k',q)] = e^{-ik phi} e^{ik' phi} sum_m d^t(m,k) A_m(t,q) d^q(m,k')
//   H_ji[(k,t),(k',q)] = (-1)^{k+k'} e^{-ik phi} e^{ik' phi}
//                        sum_m d^t(m,-k) A_m(q,t) d^q(m,-k')
// (the rotate / axial / rotate-back sequence of coefficient_translation,
// mstm-modules-33.f90:6697-6749, written out per matrix element; the reverse
// block is the shift_op form of the same table).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void expand_kernel(const double *__restrict__ tab, int slot_elems,
                                                     const int2 *__restrict__ pairs, int L, int M,
                                                     Z64 *__restrict__ H64, Z32 *__restrict__ H32)
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
    const int64_t MM = (int64_t)M * M;
    const int tot = 4 * NB * NB;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int sc = e % NB, tr = (e / NB) % NB, mode = (e / (NB * NB)) & 1, dir = e / (2 * NB * NB);
        int kt, t, kq, qd;
        unpack_mn(tr, L, kt, t);
        unpack_mn(sc, L, kq, qd);
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
        double vr = sr * cs - si * sn, vi = sr * sn + si * cs;
        if (dir && ((kt + kq) & 1)) {
            vr = -vr;
            vi = -vi;
        }
        const int64_t row = (int64_t)(dir ? pr.y : pr.x) * NB + tr, col = (int64_t)(dir ? pr.x : pr.y) * NB + sc;
        const int64_t idx = mode * MM + row * M + col;
        H64[idx] = cmk(vr, vi);
        H32[idx] = cmk((float)vr, (float)vi);
    }
}

// ---------------------------------------------------------------------------
// Grouped complex GEMM, FP32 on MFMA: for every job, a 64-row x R tile of
// Y = H X (op 0, inputs in0 -> out0) or Y = H^H X (op 1, in1 -> out1) of
// cluster c, mode p.  Persistent: each workgroup takes jobs from a counter.
// cact (optional, inside the inner solve): cact[2c + h] = 0 when every
// column of half h (columns 32h .. 32h + 31) of cluster c has stopped; a job
// with both halves stopped is skipped whole, and in the double-buffered form
// the waves of a stopped half skip their products.
// 4 waves per workgroup, each a 32x32 quadrant; complex products as four
// real MFMAs (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gemm32_kernel(const GemmJob *__restrict__ jobs, int njobs,
                                                     int *__restrict__ counter, const int *__restrict__ cM,
                                                     const int64_t *__restrict__ choff,
                                                     const int64_t *__restrict__ cvoff, const Z32 *__restrict__ H,
                                                     const Z32 *__restrict__ in0, Z32 *__restrict__ out0,
                                                     const Z32 *__restrict__ in1, Z32 *__restrict__ out1, int R,
                                                     const int *__restrict__ cact)
{
    __shared__ float As[2][64][33];
    __shared__ float Bs[2][32][65];
    __shared__ int sjob;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    for (;;) {
        if (tid == 0) sjob = atomicAdd(counter, 1);
        __syncthreads();
        const int jb = sjob;
        __syncthreads();
        if (jb >= njobs) break;
        const GemmJob J = jobs[jb];
        if (cact && !(cact[2 * J.c] | cact[2 * J.c + 1])) continue;  // every column has stopped
        const int M = cM[J.c];
        const Z32 *Hc = H + choff[J.c] + (int64_t)J.mode * M * M;
        const int64_t vb = (cvoff[J.c] + (int64_t)J.mode * M) * R;
        const Z32 *X = (J.op ? in1 : in0) + vb;
        Z32 *Y = (J.op ? out1 : out0) + vb;
        const int row0 = J.rt * 64;
        floatx16 cr, ci;
#pragma unroll
        for (int r = 0; r < 16; r++) cr[r] = ci[r] = 0.0f;
        for (int k0 = 0; k0 < M; k0 += 32) {
            if (!J.op) {
#pragma unroll
                for (int it = 0; it < 8; it++) {
                    const int kk = tid & 31, rr = (tid >> 5) + 8 * it;
                    const int gr = row0 + rr, gk = k0 + kk;
                    Z32 v = cmk(0.0f, 0.0f);
                    if (gr < M && gk < M) v = Hc[(int64_t)gr * M + gk];
                    As[0][rr][kk] = v.re;
                    As[1][rr][kk] = v.im;
                }
            } else {
#pragma unroll
                for (int it = 0; it < 8; it++) {
                    const int rr = tid & 63, kk = (tid >> 6) + 4 * it;
                    const int gr = row0 + rr, gk = k0 + kk;
                    Z32 v = cmk(0.0f, 0.0f);
                    if (gr < M && gk < M) v = Hc[(int64_t)gk * M + gr];
                    As[0][rr][kk] = v.re;
                    As[1][rr][kk] = -v.im;
                }
            }
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int cc = tid & 63, kk = (tid >> 6) + 4 * it;
                const int gk = k0 + kk;
                Z32 v = cmk(0.0f, 0.0f);
                if (gk < M && cc < R) v = X[(int64_t)gk * R + cc];
                Bs[0][kk][cc] = v.re;
                Bs[1][kk][cc] = v.im;
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int ka = 2 * s + (lane >> 5);
                const float ar = As[0][wr * 32 + (lane & 31)][ka], ai = As[1][wr * 32 + (lane & 31)][ka];
                const float br = Bs[0][ka][wc * 32 + (lane & 31)], bi = Bs[1][ka][wc * 32 + (lane & 31)];
                cr = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, cr, 0, 0, 0);
                cr = __builtin_amdgcn_mfma_f32_32x32x2f32(-ai, bi, cr, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bi, ci, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, ci, 0, 0, 0);
            }
            __syncthreads();
        }
        const int col = wc * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = row0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (row < M && col < R) Y[(int64_t)row * R + col] = cmk(cr[r], ci[r]);
        }
    }
}

// Second variant of the grouped FP32 GEMM: the next 64 x 32 k-chunk is
// fetched into registers while the current one is multiplied (two LDS
// buffers, one barrier per chunk), and the LDS keeps each lane's four
// successive k-pairs contiguous so one 16-byte read feeds four MFMA steps
// (k = 2s + h stored at h*16 + s; rows padded to 36 floats).
__global__ __launch_bounds__(256) void gemm32b_kernel(const GemmJob *__restrict__ jobs, int njobs,
                                                      int *__restrict__ counter, const int *__restrict__ cM,
                                                      const int64_t *__restrict__ choff,
                                                      const int64_t *__restrict__ cvoff, const Z32 *__restrict__ H,
                                                      const Z32 *__restrict__ in0, Z32 *__restrict__ out0,
                                                      const Z32 *__restrict__ in1, Z32 *__restrict__ out1, int R,
                                                      const int *__restrict__ cact)
{
    typedef float float4v __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) float As[2][2][64][36];
    __shared__ __align__(16) float Bs[2][2][64][36];
    __shared__ int sjob;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    for (;;) {
        if (tid == 0) sjob = atomicAdd(counter, 1);
        __syncthreads();
        const int jb = sjob;
        __syncthreads();
        if (jb >= njobs) break;
        const GemmJob J = jobs[jb];
        if (cact && !(cact[2 * J.c] | cact[2 * J.c + 1])) continue;  // every column has stopped
        const bool on = !cact || cact[2 * J.c + wc];                   // this wave's 32 columns
        const int M = cM[J.c];
        const Z32 *Hc = H + choff[J.c] + (int64_t)J.mode * M * M;
        const int64_t vb = (cvoff[J.c] + (int64_t)J.mode * M) * R;
        const Z32 *X = (J.op ? in1 : in0) + vb;
        Z32 *Y = (J.op ? out1 : out0) + vb;
        const int row0 = J.rt * 64;
        Z32 ra[8], rb[8];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                int rr, kk;
                if (!J.op) {
                    kk = tid & 31;
                    rr = (tid >> 5) + 8 * it;
                } else {
                    rr = tid & 63;
                    kk = (tid >> 6) + 4 * it;
                }
                const int gr = row0 + rr, gk = k0 + kk;
                Z32 v = cmk(0.0f, 0.0f);
                if (gr < M && gk < M) v = J.op ? Hc[(int64_t)gk * M + gr] : Hc[(int64_t)gr * M + gk];
                ra[it] = J.op ? cmk(v.re, -v.im) : v;
            }
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int cc = tid & 63, kk = (tid >> 6) + 4 * it;
                const int gk = k0 + kk;
                rb[it] = (gk < M && cc < R) ? X[(int64_t)gk * R + cc] : cmk(0.0f, 0.0f);
            }
        };
        auto stash = [&](int buf) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                int rr, kk;
                if (!J.op) {
                    kk = tid & 31;
                    rr = (tid >> 5) + 8 * it;
                } else {
                    rr = tid & 63;
                    kk = (tid >> 6) + 4 * it;
                }
                const int ks = (kk & 1) * 16 + (kk >> 1);
                As[buf][0][rr][ks] = ra[it].re;
                As[buf][1][rr][ks] = ra[it].im;
            }
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int cc = tid & 63, kk = (tid >> 6) + 4 * it;
                const int ks = (kk & 1) * 16 + (kk >> 1);
                Bs[buf][0][cc][ks] = rb[it].re;
                Bs[buf][1][cc][ks] = rb[it].im;
            }
        };
        floatx16 cr, ci;
#pragma unroll
        for (int r = 0; r < 16; r++) cr[r] = ci[r] = 0.0f;
        const int nch = (M + 31) / 32;
        fetch(0);
        stash(0);
        __syncthreads();
        const int arow = wr * 32 + (lane & 31), bcol = wc * 32 + (lane & 31), h = lane >> 5;
        for (int ch = 0; ch < nch; ch++) {
            const int buf = ch & 1;
            if (ch + 1 < nch) fetch((ch + 1) * 32);
            if (on) {
#pragma unroll
                for (int sb = 0; sb < 4; sb++) {
                    const float4v a_r = *reinterpret_cast<const float4v *>(&As[buf][0][arow][h * 16 + 4 * sb]);
                    const float4v a_i = *reinterpret_cast<const float4v *>(&As[buf][1][arow][h * 16 + 4 * sb]);
                    const float4v b_r = *reinterpret_cast<const float4v *>(&Bs[buf][0][bcol][h * 16 + 4 * sb]);
                    const float4v b_i = *reinterpret_cast<const float4v *>(&Bs[buf][1][bcol][h * 16 + 4 * sb]);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        cr = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r[u], b_r[u], cr, 0, 0, 0);
                        cr = __builtin_amdgcn_mfma_f32_32x32x2f32(-a_i[u], b_i[u], cr, 0, 0, 0);
                        ci = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r[u], b_i[u], ci, 0, 0, 0);
                        ci = __builtin_amdgcn_mfma_f32_32x32x2f32(a_i[u], b_r[u], ci, 0, 0, 0);
                    }
                }
            }
            if (ch + 1 < nch) stash(buf ^ 1);
            __syncthreads();
        }
        const int col = wc * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = row0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (on && row < M && col < R) Y[(int64_t)row * R + col] = cmk(cr[r], ci[r]);
        }
    }
}

// FP64 grouped GEMM Y = H X (op 0 jobs only), VALU, 64x64 tile, 4x4 per thread
__global__ __launch_bounds__(256) void gemm64_kernel(const GemmJob *__restrict__ jobs, int njobs,
                                                     int *__restrict__ counter, const int *__restrict__ cM,
                                                     const int64_t *__restrict__ choff,
                                                     const int64_t *__restrict__ cvoff, const Z64 *__restrict__ H,
                                                     const Z64 *__restrict__ X0, Z64 *__restrict__ Y0, int R)
{
    __shared__ Z64 As[64][17];
    __shared__ Z64 Bs[16][65];
    __shared__ int sjob;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (;;) {
        if (tid == 0) sjob = atomicAdd(counter, 1);
        __syncthreads();
        const int jb = sjob;
        __syncthreads();
        if (jb >= njobs) break;
        const GemmJob J = jobs[jb];
        const int M = cM[J.c];
        const Z64 *Hc = H + choff[J.c] + (int64_t)J.mode * M * M;
        const int64_t vb = (cvoff[J.c] + (int64_t)J.mode * M) * R;
        const Z64 *X = X0 + vb;
        Z64 *Y = Y0 + vb;
        const int row0 = J.rt * 64;
        Z64 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] = cmk(0.0, 0.0);
        for (int k0 = 0; k0 < M; k0 += 16) {
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int kk = tid & 15, rr = (tid >> 4) + 16 * it;
                const int gr = row0 + rr, gk = k0 + kk;
                As[rr][kk] = (gr < M && gk < M) ? Hc[(int64_t)gr * M + gk] : cmk(0.0, 0.0);
            }
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int cc = tid & 63, kk = (tid >> 6) + 4 * it;
                const int gk = k0 + kk;
                Bs[kk][cc] = (gk < M && cc < R) ? X[(int64_t)gk * R + cc] : cmk(0.0, 0.0);
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < 16; kk++) {
                Z64 av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; a++) av[a] = As[ty * 4 + a][kk];
#pragma unroll
                for (int b = 0; b < 4; b++) bv[b] = Bs[kk][tx * 4 + b];
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++) cfma(acc[a][b], av[a], bv[b]);
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int row = row0 + ty * 4 + a;
            if (row >= M) continue;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int col = tx * 4 + b;
                if (col < R) Y[(int64_t)row * R + col] = acc[a][b];
            }
        }
    }
}

__global__ void reset_counter_kernel(int *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0;
}

// T (or T^H) applied per row pair: out_p = sum_q T(p,q) in_q with
// T = [[t0, t2], [t1, t3]] (multmiecoeffmult, mstm-modules-33.f90:5633-5751)
template <typename Z, bool HERM>
__device__ __forceinline__ void tmul2(const Z64 *__restrict__ an, int ai, Z a0, Z a1, Z &o0, Z &o1)
{
    typedef decltype(a0.re) R_;
    Z64 t0 = an[ai], t1 = an[ai + 1], t2 = an[ai + 2], t3 = an[ai + 3];
    if (HERM) {  // T^H = [[t0*, t1*], [t2*, t3*]]
        Z64 u1 = cconj(t2), u2 = cconj(t1);
        t0 = cconj(t0);
        t3 = cconj(t3);
        t1 = u1;
        t2 = u2;
    }
    const Z c0 = cmk((R_)t0.re, (R_)t0.im), c1 = cmk((R_)t1.re, (R_)t1.im), c2 = cmk((R_)t2.re, (R_)t2.im),
            c3 = cmk((R_)t3.re, (R_)t3.im);
    o0 = cadd(cmul(c0, a0), cmul(c2, a1));
    o1 = cadd(cmul(c1, a0), cmul(c3, a1));
}

// out = T^H in (FP32) over every (row, column)
__global__ void tmulh32_kernel(int ng, const RowInfo *__restrict__ ri, const Z64 *__restrict__ an, int R,
                               const Z32 *__restrict__ in, Z32 *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ng * R) return;
    const int g = (int)(t / R), col = (int)(t % R);
    const RowInfo q = ri[g];
    const int64_t e0 = (int64_t)q.v0 * R + col, e1 = (int64_t)q.v1 * R + col;
    Z32 o0, o1;
    tmul2<Z32, true>(an, q.an, in[e0], in[e1], o0, o1);
    out[e0] = o0;
    out[e1] = o1;
}

// workgroup-level reduction of NV complex values per column: 4 row-lanes x 64 columns
template <int NV>
__device__ __forceinline__ void col_reduce(Z64 (&s)[NV], Z64 *__restrict__ partial, int R)
{
    __shared__ Z64 red[NV][4][BR];
    const int col = threadIdx.x & 63, rs = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < NV; v++) red[v][rs][col] = s[v];
    __syncthreads();
    if (rs == 0 && col < R)
#pragma unroll
        for (int v = 0; v < NV; v++) {
            Z64 a = cadd(cadd(red[v][0][col], red[v][1][col]), cadd(red[v][2][col], red[v][3][col]));
            partial[(int64_t)col * NV + v] = a;
        }
}

// inner-solve start (d = 0): q = w = conj(r), p = r; partials sum r r, sum |r|^2
__global__ __launch_bounds__(256) void bt_init_kernel(const Chunk *__restrict__ chunks, const RowInfo *__restrict__ ri,
                                                      int R, const int *__restrict__ active, Z32 *__restrict__ d,
                                                      const Z32 *__restrict__ r, Z32 *__restrict__ q,
                                                      Z32 *__restrict__ w, Z32 *__restrict__ p,
                                                      Z64 *__restrict__ partials)
{
    const Chunk ch = chunks[blockIdx.x];
    const int col = threadIdx.x & 63, rs = threadIdx.x >> 6;
    Z64 s[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
    if (col < R) {
        for (int g = ch.g0 + rs; g < ch.g1; g += 4) {
            const RowInfo qi = ri[g];
            const bool on = active[qi.c * BR + col] != 0;
            for (int md = 0; md < 2; md++) {
                const int64_t e = (int64_t)(md ? qi.v1 : qi.v0) * R + col;
                const Z32 rv = on ? r[e] : cmk(0.0f, 0.0f);
                d[e] = cmk(0.0f, 0.0f);
                q[e] = cconj(rv);
                w[e] = cconj(rv);
                p[e] = rv;
                const Z64 r64 = cmk((double)rv.re, (double)rv.im);
                cfma(s[0], r64, r64);
                s[1].re += r64.re * r64.re + r64.im * r64.im;
            }
        }
    }
    col_reduce<2>(s, partials + (int64_t)blockIdx.x * R * 2, R);
}

// cap = p - T y1 ; caw = w - y2 ; partial sum conj(w) cap
__global__ __launch_bounds__(256) void bt_post_kernel(const Chunk *__restrict__ chunks, const RowInfo *__restrict__ ri,
                                                      const Z64 *__restrict__ an, int R,
                                                      const int *__restrict__ active, const Z32 *__restrict__ p,
                                                      const Z32 *__restrict__ w, Z32 *__restrict__ y1,
                                                      Z32 *__restrict__ y2, Z64 *__restrict__ partials)
{
    const Chunk ch = chunks[blockIdx.x];
    const int col = threadIdx.x & 63, rs = threadIdx.x >> 6;
    Z64 s[1] = {cmk(0.0, 0.0)};
    if (col < R) {
        for (int g = ch.g0 + rs; g < ch.g1; g += 4) {
            const RowInfo qi = ri[g];
            if (!active[qi.c * BR + col]) continue;
            const int64_t e0 = (int64_t)qi.v0 * R + col, e1 = (int64_t)qi.v1 * R + col;
            Z32 t0, t1;
            tmul2<Z32, false>(an, qi.an, y1[e0], y1[e1], t0, t1);
            const Z32 c0 = csub(p[e0], t0), c1 = csub(p[e1], t1);
            y1[e0] = c0;
            y1[e1] = c1;
            const Z32 w0 = w[e0], w1 = w[e1];
            y2[e0] = csub(w0, y2[e0]);
            y2[e1] = csub(w1, y2[e1]);
            cfma(s[0], cmk((double)w0.re, -(double)w0.im), cmk((double)c0.re, (double)c0.im));
            cfma(s[0], cmk((double)w1.re, -(double)w1.im), cmk((double)c1.re, (double)c1.im));
        }
    }
    col_reduce<1>(s, partials + (int64_t)blockIdx.x * R, R);
}

// d += cak p ; r -= cak cap ; q -= conj(cak) caw ; partials <q, r>, <r, r>
__global__ __launch_bounds__(256) void bt_update_b_kernel(const Chunk *__restrict__ chunks,
                                                          const RowInfo *__restrict__ ri, int R,
                                                          const int *__restrict__ active,
                                                          const Z64 *__restrict__ cak, Z32 *__restrict__ d,
                                                          Z32 *__restrict__ r, Z32 *__restrict__ q,
                                                          const Z32 *__restrict__ p, const Z32 *__restrict__ cap,
                                                          const Z32 *__restrict__ caw, Z64 *__restrict__ partials)
{
    const Chunk ch = chunks[blockIdx.x];
    const int col = threadIdx.x & 63, rs = threadIdx.x >> 6;
    Z64 s[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
    if (col < R) {
        for (int g = ch.g0 + rs; g < ch.g1; g += 4) {
            const RowInfo qi = ri[g];
            const int sc = qi.c * BR + col;
            if (!active[sc]) continue;
            const Z64 k64 = cak[sc];
            const Z32 k = cmk((float)k64.re, (float)k64.im), kc = cconj(k);
            for (int md = 0; md < 2; md++) {
                const int64_t e = (int64_t)(md ? qi.v1 : qi.v0) * R + col;
                Z32 dv = d[e];
                cfma(dv, k, p[e]);
                d[e] = dv;
                Z32 rv = r[e];
                cfma(rv, cmk(-k.re, -k.im), cap[e]);
                r[e] = rv;
                Z32 qv = q[e];
                cfma(qv, cmk(-kc.re, -kc.im), caw[e]);
                q[e] = qv;
                const Z64 r64 = cmk((double)rv.re, (double)rv.im), q64 = cmk((double)qv.re, -(double)qv.im);
                cfma(s[0], q64, r64);
                s[1].re += r64.re * r64.re + r64.im * r64.im;
            }
        }
    }
    col_reduce<2>(s, partials + (int64_t)blockIdx.x * R * 2, R);
}

// p = r + cbk p ; w = q + conj(cbk) w
__global__ void bt_update_c_kernel(int ng, const RowInfo *__restrict__ ri, int R, const int *__restrict__ active,
                                   const Z64 *__restrict__ cbk, Z32 *__restrict__ p, Z32 *__restrict__ w,
                                   const Z32 *__restrict__ r, const Z32 *__restrict__ q)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ng * R) return;
    const int g = (int)(t / R), col = (int)(t % R);
    const RowInfo qi = ri[g];
    const int sc = qi.c * BR + col;
    if (!active[sc]) return;
    const Z32 k = cmk((float)cbk[sc].re, (float)cbk[sc].im), kc = cconj(k);
    for (int md = 0; md < 2; md++) {
        const int64_t e = (int64_t)(md ? qi.v1 : qi.v0) * R + col;
        Z32 pv = r[e];
        cfma(pv, k, p[e]);
        p[e] = pv;
        Z32 wv = q[e];
        cfma(wv, kc, w[e]);
        w[e] = wv;
    }
}

// per-column scalars (one thread per cluster x column)
// stage 0 (init): csk = sum r r, enorm = sum |r|^2, active = want && enorm > 0
// stage 1 (a):    cak = csk / den
// stage 2 (b):    err = <r,r>/enorm, convergence, cbk, csk
__global__ void bt_scalar_kernel(int stage, int ncl, int R, const int *__restrict__ ch0,
                                 const int *__restrict__ ch1, const Z64 *__restrict__ partials,
                                 int *__restrict__ active, const int *__restrict__ want, Z64 *__restrict__ csk,
                                 Z64 *__restrict__ cak, Z64 *__restrict__ cbk, double *__restrict__ enorm,
                                 const double *__restrict__ eps_in, double *__restrict__ err, int *__restrict__ iters,
                                 int *__restrict__ itot)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncl * BR) return;
    const int c = t / BR, col = t % BR;
    if (col >= R) return;
    const int NV = (stage == 1) ? 1 : 2;
    Z64 s0 = cmk(0.0, 0.0), s1 = cmk(0.0, 0.0);
    for (int b = ch0[c]; b < ch1[c]; b++) {
        s0 = cadd(s0, partials[((int64_t)b * R + col) * NV]);
        if (NV == 2) s1 = cadd(s1, partials[((int64_t)b * R + col) * NV + 1]);
    }
    if (stage == 0) {
        csk[t] = s0;
        enorm[t] = s1.re;
        active[t] = (want[t] && s1.re > 0.0) ? 1 : 0;
        iters[t] = 0;
        err[t] = s1.re > 0.0 ? 1.0 : 0.0;
        return;
    }
    if (!active[t]) return;
    if (stage == 1) {
        if (s0.re == 0.0 && s0.im == 0.0) {
            active[t] = 0;
            cak[t] = cmk(0.0, 0.0);
        } else {
            cak[t] = cdiv(csk[t], s0);
        }
        return;
    }
    const double e = s1.re / enorm[t];
    err[t] = e;
    iters[t] += 1;
    itot[t] += 1;
    const Z64 ck = csk[t];
    if (e < eps_in[t] || (ck.re == 0.0 && ck.im == 0.0)) {
        active[t] = 0;
        return;
    }
    cbk[t] = cdiv(s0, ck);
    csk[t] = s0;
}

// FP64 residual r = b - x + T y (y = H x), partial |r|^2 (and |b|^2); r32 = r
__global__ __launch_bounds__(256) void bt_resid_kernel(const Chunk *__restrict__ chunks,
                                                       const RowInfo *__restrict__ ri, const Z64 *__restrict__ an,
                                                       int R, int xzero, const Z64 *__restrict__ b,
                                                       const Z64 *__restrict__ x, const Z64 *__restrict__ y,
                                                       Z32 *__restrict__ r32, Z64 *__restrict__ partials)
{
    const Chunk ch = chunks[blockIdx.x];
    const int col = threadIdx.x & 63, rs = threadIdx.x >> 6;
    Z64 s[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
    if (col < R) {
        for (int g = ch.g0 + rs; g < ch.g1; g += 4) {
            const RowInfo qi = ri[g];
            const int64_t e0 = (int64_t)qi.v0 * R + col, e1 = (int64_t)qi.v1 * R + col;
            Z64 t0 = cmk(0.0, 0.0), t1 = cmk(0.0, 0.0);
            if (!xzero) tmul2<Z64, false>(an, qi.an, y[e0], y[e1], t0, t1);
            Z64 r0 = b[e0], r1 = b[e1];
            if (!xzero) {
                r0 = cadd(csub(r0, x[e0]), t0);
                r1 = cadd(csub(r1, x[e1]), t1);
            }
            r32[e0] = cmk((float)r0.re, (float)r0.im);
            r32[e1] = cmk((float)r1.re, (float)r1.im);
            s[0].re += r0.re * r0.re + r0.im * r0.im + r1.re * r1.re + r1.im * r1.im;
            const Z64 b0 = b[e0], b1 = b[e1];
            s[1].re += b0.re * b0.re + b0.im * b0.im + b1.re * b1.re + b1.im * b1.im;
        }
    }
    col_reduce<2>(s, partials + (int64_t)blockIdx.x * R * 2, R);
}

// outer scalars: err = |r|^2/|b|^2; want = err >= eps; eps_in for the inner solve
__global__ void bt_outer_kernel(int ncl, int R, const int *__restrict__ ch0, const int *__restrict__ ch1,
                                const Z64 *__restrict__ partials, double eps, double *__restrict__ oerr,
                                int *__restrict__ want, double *__restrict__ eps_in, int *__restrict__ nwant)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncl * BR) return;
    const int c = t / BR, col = t % BR;
    if (col >= R) {
        want[t] = 0;
        return;
    }
    double rr = 0.0, bb = 0.0;
    for (int b = ch0[c]; b < ch1[c]; b++) {
        rr += partials[((int64_t)b * R + col) * 2].re;
        bb += partials[((int64_t)b * R + col) * 2 + 1].re;
    }
    const double e = bb > 0.0 ? rr / bb : 0.0;
    oerr[t] = e;
    const int w = e >= eps ? 1 : 0;
    want[t] = w;
    eps_in[t] = rr > 0.0 ? fmax(0.5 * eps * bb / rr, 1e-12) : 1.0;
    if (w) atomicAdd(nwant, 1);
}

// cact[2c + h] = 1 while any column of half h of cluster c (columns
// 32h .. 32h + 31) is active in the inner solve
__global__ void cluster_active_kernel(int ncl, const int *__restrict__ active, int *__restrict__ cact)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ncl) return;
    const int c = t >> 1, h = t & 1;
    int a = 0;
    for (int j = 32 * h; j < 32 * h + 32; j++) a |= active[c * BR + j];
    cact[t] = a != 0;
}

__global__ void count_active_kernel(int n, const int *__restrict__ active, int *__restrict__ out)
{
    __shared__ int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    int v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += active[i];
    atomicAdd(&s, v);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = s;
}

// x += d (FP64 += FP32)
__global__ void add_d_kernel(int64_t n, Z64 *__restrict__ x, const Z32 *__restrict__ d)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    x[t] = cadd(x[t], cmk((double)d[t].re, (double)d[t].im));
}

// MSTM layout <-> batched layout for one direction's column pair:
// mstm[q*ldm + mi + mode*nb] <-> bat[row(mode) * R + dcol + q], q = 0, 1
__global__ void layout_kernel(int ng, const RowInfo *__restrict__ ri, int R, int dcol, int64_t ldm, int to_batched,
                              Z64 *__restrict__ mstm, Z64 *__restrict__ bat)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ng * 2) return;
    const int g = t >> 1, q = t & 1;
    const RowInfo qi = ri[g];
    for (int md = 0; md < 2; md++) {
        const int64_t eb = (int64_t)(md ? qi.v1 : qi.v0) * R + dcol + q;
        const int64_t em = (int64_t)q * ldm + qi.mi + md * qi.nb;
        if (to_batched) bat[eb] = mstm[em];
        else mstm[em] = bat[eb];
    }
}

}  // namespace

namespace {

// in-place T in the batched layout, FP64 (b = T f)
__global__ void tmul64_kernel(int ng, const RowInfo *__restrict__ ri, const Z64 *__restrict__ an, int R,
                              Z64 *__restrict__ v)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ng * R) return;
    const int g = (int)(t / R), col = (int)(t % R);
    const RowInfo q = ri[g];
    const int64_t e0 = (int64_t)q.v0 * R + col, e1 = (int64_t)q.v1 * R + col;
    Z64 o0, o1;
    tmul2<Z64, false>(an, q.an, v[e0], v[e1], o0, o1);
    v[e0] = o0;
    v[e1] = o1;
}

// host copy between the batched device layout and per-cluster MSTM-layout
// host arrays: host[(colbase_c + col) * neqns_c + mstm index] (clusters
// concatenated, each (R, neqns_c) complex)
__global__ void gather_host_kernel(int ng, const RowInfo *__restrict__ ri, const int64_t *__restrict__ hbase,
                                   const int *__restrict__ cne, int R, int to_batched, Z64 *__restrict__ host,
                                   Z64 *__restrict__ bat, const int *__restrict__ cmi0)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ng * R) return;
    const int g = (int)(t / R), col = (int)(t % R);
    const RowInfo q = ri[g];
    const int ne = cne[q.c];
    for (int md = 0; md < 2; md++) {
        const int64_t eb = (int64_t)(md ? q.v1 : q.v0) * R + col;
        const int64_t eh = hbase[q.c] + (int64_t)col * ne + (q.mi - cmi0[q.c]) + md * q.nb;
        if (to_batched) bat[eb] = host[eh];
        else host[eh] = bat[eb];
    }
}

}  // namespace

hipError_t launch_mie(const double *x, const double *mre, const double *mim, double mmed, int nsph, double qeps,
                      int *order, double *an, int ldan, hipStream_t st);

struct mstm_hip_batch {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    int ncl = 0, nsph = 0, ndir = 0, R = 0, Lmax = 0;
    int64_t neqns = 0;  // MSTM layout, all spheres
    double rimed = 1.0;
    std::vector<int> counts, sph0, cL, cM, cne, cmi0;
    std::vector<int64_t> choff, cvoff, hbase;
    std::vector<double> pos, radius, ri, alpha, sinc;
    std::vector<int> order, offset, sdir;
    // device sphere data
    double *d_pos = nullptr, *d_radius = nullptr, *d_mre = nullptr, *d_mim = nullptr;
    Z64 *d_an = nullptr;  // 2x2 T blocks per degree, stride 4 Lmax per sphere
    int *d_order = nullptr, *d_offset = nullptr, *d_mieoff = nullptr;
    // operators
    Z64 *d_H64 = nullptr;
    Z32 *d_H32 = nullptr;
    int64_t hsize = 0;
    bool built = false;
    // rows, chunks, jobs
    int ng = 0, nchunks = 0;
    int64_t nrows = 0;
    RowInfo *d_ri = nullptr;
    Chunk *d_chunks = nullptr;
    int *d_ch0 = nullptr, *d_ch1 = nullptr, *d_cM = nullptr, *d_cne = nullptr, *d_cmi0 = nullptr;
    int64_t *d_choff = nullptr, *d_cvoff = nullptr, *d_hbase = nullptr;
    GemmJob *d_jobs32[2] = {nullptr, nullptr}, *d_jobs64 = nullptr;
    int njobs32 = 0, njobs64 = 0;
    int *d_counter = nullptr;
    // vectors
    Z64 *f64 = nullptr, *b64 = nullptr, *x64 = nullptr, *y64 = nullptr;
    Z32 *r32 = nullptr, *d32 = nullptr, *q32 = nullptr, *w32 = nullptr, *p32 = nullptr, *y1 = nullptr, *y2 = nullptr,
        *tw = nullptr;
    Z64 *partials = nullptr;
    // column state [ncl * BR]
    int *d_active = nullptr, *d_want = nullptr, *d_iters = nullptr, *d_itot = nullptr, *d_cnt = nullptr;
    int *d_cact = nullptr;  // per cluster and half of its columns: any column active in the inner solve
    Z64 *d_csk = nullptr, *d_cak = nullptr, *d_cbk = nullptr;
    double *d_enorm = nullptr, *d_epsin = nullptr, *d_err = nullptr, *d_oerr = nullptr;
    // graph of GRAPH_K inner iterations
    hipGraphExec_t gexec = nullptr;
    // tuning and statistics
    int grid32 = 1024, lpt = 1, kvar = 0;
    double build_ms = 0, setup_ms = 0, tune_ms = 0, gemm32_ms = 0, gemm64_ms = 0, solve_ms = 0;
    double inner_ms = 0;
    long gemm32_launches = 0, outer_steps = 0, inner_iters = 0;
    double flops32 = 0, bytes32 = 0;
    std::vector<double> tune_table;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static const int GRAPH_K = 8;

static int bfail(mstm_hip_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}
#define BCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "%s: %s", #x, hipGetErrorString(e_));  \
    } while (0)

template <typename P> static hipError_t balloc(P **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    return hipMalloc((void **)p, bytes);
}
template <typename P> static void bfree(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static void batch_release(mstm_hip_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    bfree(b->d_pos), bfree(b->d_radius), bfree(b->d_mre), bfree(b->d_mim), bfree(b->d_an);
    bfree(b->d_order), bfree(b->d_offset), bfree(b->d_mieoff);
    bfree(b->d_H64), bfree(b->d_H32);
    bfree(b->d_ri), bfree(b->d_chunks), bfree(b->d_ch0), bfree(b->d_ch1), bfree(b->d_cM), bfree(b->d_cne),
        bfree(b->d_cmi0);
    bfree(b->d_choff), bfree(b->d_cvoff), bfree(b->d_hbase);
    bfree(b->d_jobs32[0]), bfree(b->d_jobs32[1]), bfree(b->d_jobs64), bfree(b->d_counter);
    bfree(b->f64), bfree(b->b64), bfree(b->x64), bfree(b->y64);
    bfree(b->r32), bfree(b->d32), bfree(b->q32), bfree(b->w32), bfree(b->p32), bfree(b->y1), bfree(b->y2),
        bfree(b->tw);
    bfree(b->partials);
    bfree(b->d_active), bfree(b->d_want), bfree(b->d_iters), bfree(b->d_itot), bfree(b->d_cnt), bfree(b->d_cact);
    bfree(b->d_csk), bfree(b->d_cak), bfree(b->d_cbk), bfree(b->d_enorm), bfree(b->d_epsin), bfree(b->d_err),
        bfree(b->d_oerr);
    b->built = false;
}

static inline int grid_of(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// one grouped FP32 GEMM: y1 = H p (op 0) and y2 = H^H tw (op 1)
static hipError_t gemm32(mstm_hip_batch *b, const Z32 *in0, Z32 *out0, const Z32 *in1, Z32 *out1, int grid,
                         const GemmJob *jobs, const int *cact = nullptr)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, b->st, b->d_counter);
    if (b->kvar == 0)
        hipLaunchKernelGGL(gemm32_kernel, dim3(std::min(grid, b->njobs32)), dim3(256), 0, b->st, jobs, b->njobs32,
                           b->d_counter, b->d_cM, b->d_choff, b->d_cvoff, b->d_H32, in0, out0, in1, out1, b->R, cact);
    else
        hipLaunchKernelGGL(gemm32b_kernel, dim3(std::min(grid, b->njobs32)), dim3(256), 0, b->st, jobs, b->njobs32,
                           b->d_counter, b->d_cM, b->d_choff, b->d_cvoff, b->d_H32, in0, out0, in1, out1, b->R, cact);
    return hipGetLastError();
}

static hipError_t gemm64(mstm_hip_batch *b, const Z64 *x, Z64 *y)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, b->st, b->d_counter + 1);
    hipLaunchKernelGGL(gemm64_kernel, dim3(std::min(1024, b->njobs64)), dim3(256), 0, b->st, b->d_jobs64, b->njobs64,
                       b->d_counter + 1, b->d_cM, b->d_choff, b->d_cvoff, b->d_H64, x, y, b->R);
    return hipGetLastError();
}

// the launches of one inner BiCG iteration (cbicg :860-933 per column)
static hipError_t inner_iteration(mstm_hip_batch *b)
{
    const int R = b->R, ncl = b->ncl;
    const int64_t ne = (int64_t)b->ng * R;
    hipLaunchKernelGGL(tmulh32_kernel, dim3(grid_of(ne, 256)), dim3(256), 0, b->st, b->ng, b->d_ri, b->d_an, R,
                       b->w32, b->tw);
    // the products of clusters whose columns have all stopped are skipped
    hipLaunchKernelGGL(cluster_active_kernel, dim3(grid_of(2 * ncl, 256)), dim3(256), 0, b->st, ncl, b->d_active,
                       b->d_cact);
    hipError_t e = gemm32(b, b->p32, b->y1, b->tw, b->y2, b->grid32, b->d_jobs32[b->lpt], b->d_cact);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bt_post_kernel, dim3(b->nchunks), dim3(256), 0, b->st, b->d_chunks, b->d_ri, b->d_an, R,
                       b->d_active, b->p32, b->w32, b->y1, b->y2, b->partials);
    hipLaunchKernelGGL(bt_scalar_kernel, dim3(grid_of(ncl * BR, 256)), dim3(256), 0, b->st, 1, ncl, R, b->d_ch0,
                       b->d_ch1, b->partials, b->d_active, b->d_want, b->d_csk, b->d_cak, b->d_cbk, b->d_enorm,
                       b->d_epsin, b->d_err, b->d_iters, b->d_itot);
    hipLaunchKernelGGL(bt_update_b_kernel, dim3(b->nchunks), dim3(256), 0, b->st, b->d_chunks, b->d_ri, R,
                       b->d_active, b->d_cak, b->d32, b->r32, b->q32, b->p32, b->y1, b->y2, b->partials);
    hipLaunchKernelGGL(bt_scalar_kernel, dim3(grid_of(ncl * BR, 256)), dim3(256), 0, b->st, 2, ncl, R, b->d_ch0,
                       b->d_ch1, b->partials, b->d_active, b->d_want, b->d_csk, b->d_cak, b->d_cbk, b->d_enorm,
                       b->d_epsin, b->d_err, b->d_iters, b->d_itot);
    hipLaunchKernelGGL(bt_update_c_kernel, dim3(grid_of(ne, 256)), dim3(256), 0, b->st, b->ng, b->d_ri, R,
                       b->d_active, b->d_cbk, b->p32, b->w32, b->r32, b->q32);
    return hipGetLastError();
}

static int count_active(mstm_hip_batch *b, int *out)
{
    hipLaunchKernelGGL(count_active_kernel, dim3(1), dim3(256), 0, b->st, b->ncl * BR, b->d_active, b->d_cnt);
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(out, b->d_cnt, sizeof(int), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    return 0;
}

static int build_operators(mstm_hip_batch *b)
{
    const int LM = b->Lmax;
    std::vector<double> vh;
    vcc_table(LM, vh);
    double *d_vcc = nullptr;
    BCHK(balloc(&d_vcc, sizeof(double) * vh.size()));
    BCHK(hipMemcpy(d_vcc, vh.data(), sizeof(double) * vh.size(), hipMemcpyHostToDevice));
    const int CH = 8192;  // pairs per table chunk (multiple of the 16-slot groups)
    const int se_max = slot_len(LM, LM, 2);
    double *tab = nullptr, *d_rt = nullptr;
    int2 *d_pairs = nullptr;
    int rc = 0;
    hipError_t e = balloc(&tab, sizeof(double) * (size_t)CH * se_max);
    if (e == hipSuccess) e = balloc(&d_rt, sizeof(double) * 3 * CH);
    if (e == hipSuccess) e = balloc(&d_pairs, sizeof(int2) * CH);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_H64, 0, sizeof(Z64) * b->hsize, b->st);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_H32, 0, sizeof(Z32) * b->hsize, b->st);
    std::vector<double> rt(3 * CH);
    std::vector<int2> pr(CH);
    for (int c = 0; c < b->ncl && e == hipSuccess; c++) {
        const int n = b->counts[c], L = b->cL[c], s0 = b->sph0[c];
        const int se = slot_len(L, L, 2);
        std::vector<int2> all;
        all.reserve((size_t)n * (n - 1) / 2);
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) all.push_back(make_int2(i, j));
        for (size_t p0 = 0; p0 < all.size() && e == hipSuccess; p0 += CH) {
            const int np = (int)std::min<size_t>(CH, all.size() - p0);
            for (int q = 0; q < np; q++) {
                const int2 ij = all[p0 + q];
                pr[q] = ij;
                for (int a = 0; a < 3; a++)
                    rt[3 * q + a] = b->pos[3 * (s0 + ij.x) + a] - b->pos[3 * (s0 + ij.y) + a];
            }
            e = hipMemcpyAsync(d_rt, rt.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, b->st);
            if (e == hipSuccess) e = hipMemcpyAsync(d_pairs, pr.data(), sizeof(int2) * np, hipMemcpyHostToDevice, b->st);
            if (e == hipSuccess) e = launch_list_tables<double>(tab, d_rt, np, L, L, 3, b->rimed, d_vcc, LM, b->st);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(expand_kernel, dim3(np), dim3(256), sizeof(double) * se, b->st, tab, se, d_pairs, L,
                                   b->cM[c], b->d_H64 + b->choff[c], b->d_H32 + b->choff[c]);
                e = hipGetLastError();
            }
            // host buffers rt/pr are reused by the next chunk: wait for the copies
            if (e == hipSuccess) e = hipStreamSynchronize(b->st);
        }
    }
    if (e != hipSuccess) rc = bfail(b, MSTM_HIP_EHIP, "build_operators: %s", hipGetErrorString(e));
    bfree(tab);
    bfree(d_rt);
    bfree(d_pairs);
    bfree(d_vcc);
    return rc;
}

static int build_rhs(mstm_hip_batch *b)
{
    // f (MSTM layout, both polarizations) per direction -> batched column pair
    Z64 *pm = nullptr;
    BCHK(balloc(&pm, sizeof(Z64) * 2 * b->neqns));
    hipError_t e = hipSuccess;
    for (int d = 0; d < b->ndir && e == hipSuccess; d++) {
        e = launch_planewave((double *)pm, (int)(2 * b->neqns), b->nsph, b->d_pos, b->d_order, b->d_offset,
                             b->alpha[d], b->sinc[d], b->sdir[d], b->rimed, b->st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(layout_kernel, dim3(grid_of((int64_t)b->ng * 2, 256)), dim3(256), 0, b->st, b->ng,
                               b->d_ri, b->R, 2 * d, b->neqns, 1, pm, b->f64);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(b->b64, b->f64, sizeof(Z64) * b->nrows * b->R, hipMemcpyDeviceToDevice, b->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(tmul64_kernel, dim3(grid_of((int64_t)b->ng * b->R, 256)), dim3(256), 0, b->st, b->ng,
                           b->d_ri, b->d_an, b->R, b->b64);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    bfree(pm);
    if (e != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "build_rhs: %s", hipGetErrorString(e));
    return 0;
}

static int autotune(mstm_hip_batch *b)
{
    const int grids[] = {512, 1024, 2048};
    double best = 1e30;
    int bestv = 0;
    b->tune_table.clear();
    for (int kv = 0; kv < 2; kv++)
    for (int o = 0; o < 2; o++)
        for (int g : grids) {
            b->kvar = kv;
            float ms = 0.0f, tot = 0.0f;
            BCHK(gemm32(b, b->p32, b->y1, b->tw, b->y2, g, b->d_jobs32[o]));  // warm-up
            for (int rep = 0; rep < 2; rep++) {
                BCHK(hipEventRecord(b->ev0, b->st));
                BCHK(gemm32(b, b->p32, b->y1, b->tw, b->y2, g, b->d_jobs32[o]));
                BCHK(hipEventRecord(b->ev1, b->st));
                BCHK(hipEventSynchronize(b->ev1));
                BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                tot += ms;
            }
            const double avg = tot / 2.0;
            b->tune_table.push_back(kv * 2 + o);
            b->tune_table.push_back(g);
            b->tune_table.push_back(avg);
            if (avg < best) {
                best = avg;
                b->grid32 = g;
                b->lpt = o;
                bestv = kv;
            }
        }
    b->kvar = bestv;
    b->gemm32_ms = best;
    return 0;
}

extern "C" {

int mstm_hip_batch_create(mstm_hip_batch **out, int device)
{
    if (!out) return MSTM_HIP_EARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MSTM_HIP_EHIP;
    auto *b = new mstm_hip_batch();
    b->device = device;
    *out = b;
    BCHK(hipSetDevice(device));
    BCHK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    BCHK(hipEventCreate(&b->ev0));
    BCHK(hipEventCreate(&b->ev1));
    return 0;
}

int mstm_hip_batch_destroy(mstm_hip_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    if (b->st) (void)hipStreamSynchronize(b->st);
    batch_release(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return 0;
}

const char *mstm_hip_batch_last_error(const mstm_hip_batch *b) { return b ? b->err.c_str() : "null batch"; }

int mstm_hip_batch_set(mstm_hip_batch *b, int ncluster, const int *counts, const double *pos, const double *radius,
                       const double *ri, double ri_medium, double mie_eps, int ndir, const double *alpha,
                       const double *beta)
{
    if (!b || ncluster <= 0 || !counts || !pos || !radius || !ri || ndir <= 0 || !alpha || !beta)
        return bfail(b, MSTM_HIP_EARG, "batch_set: bad arguments");
    if (2 * ndir > BR) return bfail(b, MSTM_HIP_EARG, "batch_set: %d directions > %d", ndir, BR / 2);
    BCHK(hipSetDevice(b->device));
    batch_release(b);
    auto t0 = std::chrono::steady_clock::now();
    b->ncl = ncluster;
    b->counts.assign(counts, counts + ncluster);
    b->nsph = 0;
    b->sph0.resize(ncluster);
    for (int c = 0; c < ncluster; c++) {
        if (counts[c] <= 0) return bfail(b, MSTM_HIP_EARG, "batch_set: cluster %d is empty", c);
        b->sph0[c] = b->nsph;
        b->nsph += counts[c];
    }
    const int n = b->nsph;
    b->pos.assign(pos, pos + 3 * n);
    b->radius.assign(radius, radius + n);
    b->ri.assign(ri, ri + 2 * n);
    b->rimed = ri_medium;
    b->ndir = ndir;
    b->R = 2 * ndir;
    b->alpha.assign(alpha, alpha + ndir);
    b->sinc.resize(ndir);
    b->sdir.resize(ndir);
    for (int d = 0; d < ndir; d++) {  // fixedorsoln's (alpha, sinc, dir) form of (alpha, beta)
        b->sinc[d] = ri_medium * std::sin(beta[d]);
        b->sdir[d] = std::cos(beta[d]) >= 0.0 ? 1 : 2;
    }
    std::vector<double> mre(n), mim(n);
    for (int i = 0; i < n; i++) mre[i] = ri[2 * i], mim[i] = ri[2 * i + 1];
    BCHK(balloc(&b->d_pos, sizeof(double) * 3 * n));
    BCHK(balloc(&b->d_radius, sizeof(double) * n));
    BCHK(balloc(&b->d_mre, sizeof(double) * n));
    BCHK(balloc(&b->d_mim, sizeof(double) * n));
    BCHK(balloc(&b->d_order, sizeof(int) * n));
    BCHK(hipMemcpy(b->d_pos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_radius, radius, sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_mre, mre.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_mim, mim.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(launch_mie(b->d_radius, b->d_mre, b->d_mim, ri_medium, n, mie_eps, b->d_order, nullptr, 0, b->st));
    b->order.resize(n);
    BCHK(hipMemcpyAsync(b->order.data(), b->d_order, sizeof(int) * n, hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    // per-cluster order (one order per cluster: the dense blocks are square)
    b->cL.resize(ncluster);
    b->Lmax = 0;
    for (int c = 0; c < ncluster; c++) {
        const int L = b->order[b->sph0[c]];
        for (int i = 1; i < counts[c]; i++)
            if (b->order[b->sph0[c] + i] != L)
                return bfail(b, MSTM_HIP_EUNSUP, "batch_set: cluster %d mixes sphere orders (%d, %d)", c, L,
                             b->order[b->sph0[c] + i]);
        if (L > 8) return bfail(b, MSTM_HIP_EUNSUP, "batch_set: order %d > 8", L);
        b->cL[c] = L;
        b->Lmax = std::max(b->Lmax, L);
    }
    b->offset.assign(n + 1, 0);
    for (int i = 0; i < n; i++) b->offset[i + 1] = b->offset[i] + 2 * b->order[i] * (b->order[i] + 2);
    b->neqns = b->offset[n];
    std::vector<int> mieoff(n);
    for (int i = 0; i < n; i++) mieoff[i] = 4 * b->Lmax * i;
    BCHK(balloc(&b->d_offset, sizeof(int) * (n + 1)));
    BCHK(balloc(&b->d_mieoff, sizeof(int) * n));
    BCHK(hipMemcpy(b->d_offset, b->offset.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_mieoff, mieoff.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    BCHK(balloc(&b->d_an, sizeof(Z64) * 4 * (size_t)b->Lmax * n));
    BCHK(hipMemsetAsync(b->d_an, 0, sizeof(Z64) * 4 * (size_t)b->Lmax * n, b->st));
    BCHK(launch_mie(b->d_radius, b->d_mre, b->d_mim, ri_medium, n, mie_eps, b->d_order, (double *)b->d_an,
                    4 * b->Lmax, b->st));
    // cluster geometry of the batched layout
    b->cM.resize(ncluster);
    b->cne.resize(ncluster);
    b->cmi0.resize(ncluster);
    b->choff.resize(ncluster);
    b->cvoff.resize(ncluster);
    b->hbase.resize(ncluster);
    int64_t ho = 0, vo = 0, hb = 0;
    for (int c = 0; c < ncluster; c++) {
        const int L = b->cL[c], NB = L * (L + 2);
        b->cM[c] = counts[c] * NB;
        b->cne[c] = 2 * b->cM[c];
        b->cmi0[c] = b->offset[b->sph0[c]];
        b->choff[c] = ho;
        b->cvoff[c] = vo;
        b->hbase[c] = hb;
        ho += 2 * (int64_t)b->cM[c] * b->cM[c];
        vo += 2 * (int64_t)b->cM[c];
        hb += (int64_t)b->R * b->cne[c];
    }
    b->hsize = ho;
    b->nrows = vo;
    // rows, chunks, jobs
    std::vector<RowInfo> rows;
    std::vector<Chunk> chunks;
    std::vector<int> ch0(ncluster), ch1(ncluster);
    for (int c = 0; c < ncluster; c++) {
        const int L = b->cL[c], NB = L * (L + 2);
        const int g0 = (int)rows.size();
        for (int i = 0; i < counts[c]; i++) {
            const int ig = b->sph0[c] + i;
            for (int e = 0; e < NB; e++) {
                const int row = e % (L + 2), colm = e / (L + 2) + 1, nd = (row <= colm) ? colm : row - 1;
                RowInfo q;
                q.v0 = (int)(b->cvoff[c] + (int64_t)i * NB + e);
                q.v1 = q.v0 + b->cM[c];
                q.an = mieoff[ig] + 4 * (nd - 1);
                q.c = c;
                q.mi = b->offset[ig] + e;
                q.nb = NB;
                rows.push_back(q);
            }
        }
        const int g1 = (int)rows.size();
        ch0[c] = (int)chunks.size();
        for (int g = g0; g < g1; g += 256) chunks.push_back(Chunk{g, std::min(g + 256, g1)});
        ch1[c] = (int)chunks.size();
    }
    if (b->nrows * b->R > (int64_t)1 << 40) return bfail(b, MSTM_HIP_EARG, "batch_set: too large");
    b->ng = (int)rows.size();
    b->nchunks = (int)chunks.size();
    std::vector<GemmJob> j32, j64;
    for (int c = 0; c < ncluster; c++) {
        const int nt = (b->cM[c] + 63) / 64;
        for (int mode = 0; mode < 2; mode++)
            for (int rt = 0; rt < nt; rt++) {
                j32.push_back(GemmJob{c, mode, 0, rt});
                j32.push_back(GemmJob{c, mode, 1, rt});
                j64.push_back(GemmJob{c, mode, 0, rt});
            }
    }
    std::vector<GemmJob> j32l = j32;  // longest (largest M) first
    std::stable_sort(j32l.begin(), j32l.end(), [&](const GemmJob &x, const GemmJob &y) { return b->cM[x.c] > b->cM[y.c]; });
    std::stable_sort(j64.begin(), j64.end(), [&](const GemmJob &x, const GemmJob &y) { return b->cM[x.c] > b->cM[y.c]; });
    b->njobs32 = (int)j32.size();
    b->njobs64 = (int)j64.size();
    BCHK(balloc(&b->d_ri, sizeof(RowInfo) * rows.size()));
    BCHK(balloc(&b->d_chunks, sizeof(Chunk) * chunks.size()));
    BCHK(balloc(&b->d_ch0, sizeof(int) * ncluster));
    BCHK(balloc(&b->d_ch1, sizeof(int) * ncluster));
    BCHK(balloc(&b->d_cM, sizeof(int) * ncluster));
    BCHK(balloc(&b->d_cne, sizeof(int) * ncluster));
    BCHK(balloc(&b->d_cmi0, sizeof(int) * ncluster));
    BCHK(balloc(&b->d_choff, sizeof(int64_t) * ncluster));
    BCHK(balloc(&b->d_cvoff, sizeof(int64_t) * ncluster));
    BCHK(balloc(&b->d_hbase, sizeof(int64_t) * ncluster));
    BCHK(balloc(&b->d_jobs32[0], sizeof(GemmJob) * j32.size()));
    BCHK(balloc(&b->d_jobs32[1], sizeof(GemmJob) * j32.size()));
    BCHK(balloc(&b->d_jobs64, sizeof(GemmJob) * j64.size()));
    BCHK(balloc(&b->d_counter, sizeof(int) * 2));
    BCHK(hipMemcpy(b->d_ri, rows.data(), sizeof(RowInfo) * rows.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_chunks, chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_ch0, ch0.data(), sizeof(int) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_ch1, ch1.data(), sizeof(int) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_cM, b->cM.data(), sizeof(int) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_cne, b->cne.data(), sizeof(int) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_cmi0, b->cmi0.data(), sizeof(int) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_choff, b->choff.data(), sizeof(int64_t) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_cvoff, b->cvoff.data(), sizeof(int64_t) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_hbase, b->hbase.data(), sizeof(int64_t) * ncluster, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_jobs32[0], j32.data(), sizeof(GemmJob) * j32.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_jobs32[1], j32l.data(), sizeof(GemmJob) * j32.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_jobs64, j64.data(), sizeof(GemmJob) * j64.size(), hipMemcpyHostToDevice));
    // vectors and column state
    const size_t nv = (size_t)b->nrows * b->R;
    BCHK(balloc(&b->f64, sizeof(Z64) * nv));
    BCHK(balloc(&b->b64, sizeof(Z64) * nv));
    BCHK(balloc(&b->x64, sizeof(Z64) * nv));
    BCHK(balloc(&b->y64, sizeof(Z64) * nv));
    Z32 **v32[] = {&b->r32, &b->d32, &b->q32, &b->w32, &b->p32, &b->y1, &b->y2, &b->tw};
    for (auto pp : v32) {
        BCHK(balloc(pp, sizeof(Z32) * nv));
        BCHK(hipMemsetAsync(*pp, 0, sizeof(Z32) * nv, b->st));
    }
    BCHK(balloc(&b->partials, sizeof(Z64) * 2 * (size_t)b->nchunks * b->R));
    const size_t ns = (size_t)ncluster * BR;
    BCHK(balloc(&b->d_active, sizeof(int) * ns));
    BCHK(balloc(&b->d_cact, sizeof(int) * 2 * (size_t)ncluster));
    BCHK(balloc(&b->d_want, sizeof(int) * ns));
    BCHK(balloc(&b->d_iters, sizeof(int) * ns));
    BCHK(balloc(&b->d_itot, sizeof(int) * ns));
    BCHK(balloc(&b->d_cnt, sizeof(int) * 2));
    BCHK(balloc(&b->d_csk, sizeof(Z64) * ns));
    BCHK(balloc(&b->d_cak, sizeof(Z64) * ns));
    BCHK(balloc(&b->d_cbk, sizeof(Z64) * ns));
    BCHK(balloc(&b->d_enorm, sizeof(double) * ns));
    BCHK(balloc(&b->d_epsin, sizeof(double) * ns));
    BCHK(balloc(&b->d_err, sizeof(double) * ns));
    BCHK(balloc(&b->d_oerr, sizeof(double) * ns));
    BCHK(hipMemsetAsync(b->d_active, 0, sizeof(int) * ns, b->st));
    BCHK(hipMemsetAsync(b->d_want, 0, sizeof(int) * ns, b->st));
    BCHK(hipMemsetAsync(b->d_itot, 0, sizeof(int) * ns, b->st));
    BCHK(hipMemsetAsync(b->d_oerr, 0, sizeof(double) * ns, b->st));
    BCHK(hipStreamSynchronize(b->st));
    b->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int mstm_hip_batch_orders(const mstm_hip_batch *b, int *order)
{
    if (!b || !order) return MSTM_HIP_EARG;
    for (int c = 0; c < b->ncl; c++) order[c] = b->cL[c];
    return 0;
}

// dense operators (FP64 build; FP64 + FP32 store), right-hand sides, autotune
int mstm_hip_batch_setup(mstm_hip_batch *b, int tune)
{
    if (!b || !b->ncl) return bfail(b, MSTM_HIP_ESTATE, "batch_setup: no clusters");
    BCHK(hipSetDevice(b->device));
    if (!b->d_H64) {
        BCHK(balloc(&b->d_H64, sizeof(Z64) * b->hsize));
        BCHK(balloc(&b->d_H32, sizeof(Z32) * b->hsize));
    }
    auto t0 = std::chrono::steady_clock::now();
    int rc = build_operators(b);
    if (rc) return rc;
    b->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    rc = build_rhs(b);
    if (rc) return rc;
    b->flops32 = 0;
    b->bytes32 = 0;
    for (int c = 0; c < b->ncl; c++) {
        const double M = b->cM[c];
        b->flops32 += 2.0 * 2.0 * M * M * b->R * 8.0;               // 2 modes x (H, H^H), complex MAC = 8 flops
        b->bytes32 += 2.0 * 2.0 * M * M * 8.0 + 2.0 * 2.0 * 2.0 * M * b->R * 8.0;  // H twice + in/out vectors
    }
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (tune) {
        rc = autotune(b);
        if (rc) return rc;
    }
    b->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    b->built = true;
    return 0;
}

// Solve every (cluster, direction, polarization) to |r|^2/|b|^2 < eps.
// iters[c*2*ndir + col]: inner FP32 iterations; err: final FP64 |r|^2/|b|^2.
int mstm_hip_batch_solve(mstm_hip_batch *b, double eps, int niter, int use_graph, int *iters, double *err)
{
    if (!b || !b->built) return bfail(b, MSTM_HIP_ESTATE, "batch_solve: call mstm_hip_batch_setup first");
    if (niter <= 0 || !(eps > 0.0)) return bfail(b, MSTM_HIP_EARG, "batch_solve: niter/eps");
    BCHK(hipSetDevice(b->device));
    const int R = b->R, ncl = b->ncl;
    const size_t nv = (size_t)b->nrows * R, ns = (size_t)ncl * BR;
    auto t0 = std::chrono::steady_clock::now();
    BCHK(hipMemsetAsync(b->x64, 0, sizeof(Z64) * nv, b->st));
    BCHK(hipMemsetAsync(b->d_itot, 0, sizeof(int) * ns, b->st));
    b->outer_steps = 0;
    b->inner_iters = 0;
    b->inner_ms = 0;
    const int max_ref = 6;
    for (int ref = 0; ref <= max_ref; ref++) {
        if (ref > 0) BCHK(gemm64(b, b->x64, b->y64));
        BCHK(hipMemsetAsync(b->d_cnt + 1, 0, sizeof(int), b->st));
        hipLaunchKernelGGL(bt_resid_kernel, dim3(b->nchunks), dim3(256), 0, b->st, b->d_chunks, b->d_ri, b->d_an, R,
                           ref == 0 ? 1 : 0, b->b64, b->x64, b->y64, b->r32, b->partials);
        hipLaunchKernelGGL(bt_outer_kernel, dim3(grid_of(ns, 256)), dim3(256), 0, b->st, ncl, R, b->d_ch0, b->d_ch1,
                           b->partials, eps, b->d_oerr, b->d_want, b->d_epsin, b->d_cnt + 1);
        BCHK(hipGetLastError());
        int nwant = 0;
        BCHK(hipMemcpyAsync(&nwant, b->d_cnt + 1, sizeof(int), hipMemcpyDeviceToHost, b->st));
        BCHK(hipStreamSynchronize(b->st));
        if (nwant == 0 || ref == max_ref) break;
        b->outer_steps++;
        // inner FP32 BiCG on every column still wanted, from d = 0
        auto ti = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(bt_init_kernel, dim3(b->nchunks), dim3(256), 0, b->st, b->d_chunks, b->d_ri, R, b->d_want,
                           b->d32, b->r32, b->q32, b->w32, b->p32, b->partials);
        hipLaunchKernelGGL(bt_scalar_kernel, dim3(grid_of(ns, 256)), dim3(256), 0, b->st, 0, ncl, R, b->d_ch0,
                           b->d_ch1, b->partials, b->d_active, b->d_want, b->d_csk, b->d_cak, b->d_cbk, b->d_enorm,
                           b->d_epsin, b->d_err, b->d_iters, b->d_itot);
        BCHK(hipGetLastError());
        if (use_graph && !b->gexec) {
            hipGraph_t g = nullptr;
            BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (int k = 0; k < GRAPH_K && ce == hipSuccess; k++) ce = inner_iteration(b);
            hipError_t ee = hipStreamEndCapture(b->st, &g);
            if (ce != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "graph capture: %s", hipGetErrorString(ce));
            if (ee != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "graph capture end: %s", hipGetErrorString(ee));
            hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "graph instantiate: %s", hipGetErrorString(ie));
        }
        int done = 0;
        while (done < niter) {
            if (use_graph) {
                BCHK(hipGraphLaunch(b->gexec, b->st));
            } else {
                for (int k = 0; k < GRAPH_K; k++) BCHK(inner_iteration(b));
            }
            done += GRAPH_K;
            b->gemm32_launches += GRAPH_K;
            int na = 0;
            int rc = count_active(b, &na);
            if (rc) return rc;
            if (na == 0) break;
        }
        b->inner_iters += done;
        b->inner_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ti).count();
        hipLaunchKernelGGL(add_d_kernel, dim3(grid_of((int64_t)nv, 256)), dim3(256), 0, b->st, (int64_t)nv, b->x64,
                           b->d32);
        BCHK(hipGetLastError());
    }
    BCHK(hipStreamSynchronize(b->st));
    b->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<int> it(ns);
    std::vector<double> er(ns);
    BCHK(hipMemcpy(it.data(), b->d_itot, sizeof(int) * ns, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(er.data(), b->d_oerr, sizeof(double) * ns, hipMemcpyDeviceToHost));
    for (int c = 0; c < ncl; c++)
        for (int col = 0; col < R; col++) {
            if (iters) iters[c * R + col] = it[c * BR + col];
            if (err) err[c * R + col] = er[c * BR + col];
        }
    return 0;
}

// solution (which = 0), right-hand side b = T f (1) or incident field f (2),
// per cluster (R, neqns_c) complex in MSTM layout, clusters concatenated
int mstm_hip_batch_vectors(mstm_hip_batch *b, int which, double *out)
{
    if (!b || !b->ncl || !out) return bfail(b, MSTM_HIP_EARG, "batch_vectors: args");
    BCHK(hipSetDevice(b->device));
    const Z64 *src = which == 0 ? b->x64 : (which == 1 ? b->b64 : b->f64);
    const int64_t tot = b->hbase.back() + (int64_t)b->R * b->cne.back();
    Z64 *h = nullptr;
    BCHK(balloc(&h, sizeof(Z64) * tot));
    hipLaunchKernelGGL(gather_host_kernel, dim3(grid_of((int64_t)b->ng * b->R, 256)), dim3(256), 0, b->st, b->ng,
                       b->d_ri, b->d_hbase, b->d_cne, b->R, 0, h, const_cast<Z64 *>(src), b->d_cmi0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, h, sizeof(Z64) * tot, hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    bfree(h);
    if (e != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "batch_vectors: %s", hipGetErrorString(e));
    return 0;
}

// y = H x (op 0) or H^H x (op 1) with the FP32 MFMA (precision 1) or FP64
// (precision 0, op 0 only) operator; x, y in the layout of batch_vectors
int mstm_hip_batch_product(mstm_hip_batch *b, int precision, int op, const double *x, double *y)
{
    if (!b || !b->built || !x || !y) return bfail(b, MSTM_HIP_ESTATE, "batch_product: no operator");
    if (precision == MSTM_HIP_FP64 && op != 0) return bfail(b, MSTM_HIP_EUNSUP, "batch_product: FP64 H^H");
    BCHK(hipSetDevice(b->device));
    const int64_t tot = b->hbase.back() + (int64_t)b->R * b->cne.back();
    const size_t nv = (size_t)b->nrows * b->R;
    Z64 *h = nullptr;
    BCHK(balloc(&h, sizeof(Z64) * tot));
    hipError_t e = hipMemcpyAsync(h, x, sizeof(Z64) * tot, hipMemcpyHostToDevice, b->st);
    const int gb = grid_of((int64_t)b->ng * b->R, 256);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather_host_kernel, dim3(gb), dim3(256), 0, b->st, b->ng, b->d_ri, b->d_hbase, b->d_cne,
                           b->R, 1, h, b->y64, b->d_cmi0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        if (precision == MSTM_HIP_FP64) {
            Z64 *tmp = nullptr;
            e = balloc(&tmp, sizeof(Z64) * nv);
            if (e == hipSuccess) e = gemm64(b, b->y64, tmp);
            if (e == hipSuccess) e = hipMemcpyAsync(b->y64, tmp, sizeof(Z64) * nv, hipMemcpyDeviceToDevice, b->st);
            if (e == hipSuccess) e = hipStreamSynchronize(b->st);
            bfree(tmp);
        } else {
            // FP64 -> FP32 on the host (test entry, not a hot path)
            Z64 *hx = (Z64 *)malloc(sizeof(Z64) * nv);
            Z32 *hx32 = (Z32 *)malloc(sizeof(Z32) * nv);
            e = hipMemcpyAsync(hx, b->y64, sizeof(Z64) * nv, hipMemcpyDeviceToHost, b->st);
            if (e == hipSuccess) e = hipStreamSynchronize(b->st);
            for (size_t q = 0; q < nv && e == hipSuccess; q++) hx32[q] = cmk((float)hx[q].re, (float)hx[q].im);
            Z32 *in = op ? b->tw : b->p32, *out = op ? b->y2 : b->y1;
            if (e == hipSuccess) e = hipMemcpyAsync(in, hx32, sizeof(Z32) * nv, hipMemcpyHostToDevice, b->st);
            if (e == hipSuccess) e = gemm32(b, b->p32, b->y1, b->tw, b->y2, b->grid32, b->d_jobs32[b->lpt]);
            if (e == hipSuccess) e = hipMemcpyAsync(hx32, out, sizeof(Z32) * nv, hipMemcpyDeviceToHost, b->st);
            if (e == hipSuccess) e = hipStreamSynchronize(b->st);
            for (size_t q = 0; q < nv && e == hipSuccess; q++) hx[q] = cmk((double)hx32[q].re, (double)hx32[q].im);
            if (e == hipSuccess) e = hipMemcpyAsync(b->y64, hx, sizeof(Z64) * nv, hipMemcpyHostToDevice, b->st);
            if (e == hipSuccess) e = hipStreamSynchronize(b->st);
            free(hx);
            free(hx32);
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather_host_kernel, dim3(gb), dim3(256), 0, b->st, b->ng, b->d_ri, b->d_hbase, b->d_cne,
                           b->R, 0, h, b->y64, b->d_cmi0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, h, sizeof(Z64) * tot, hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    bfree(h);
    if (e != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "batch_product: %s", hipGetErrorString(e));
    return 0;
}

// per (cluster, direction): qtot (mstm-scatprops-26.f90:872-912) of the
// solution, q[(c*ndir + d)*9 + 3*k + j], k = unpolarized/pol 1/pol 2,
// j = ext/abs/sca, with the cross-section radius (sum a^3)^0.333333
int mstm_hip_batch_qeff(mstm_hip_batch *b, double *q)
{
    if (!b || !b->built || !q) return bfail(b, MSTM_HIP_ESTATE, "batch_qeff: args");
    BCHK(hipSetDevice(b->device));
    Z64 *am = nullptr, *pm = nullptr;
    double *dq = nullptr;
    BCHK(balloc(&am, sizeof(Z64) * 2 * b->neqns));
    BCHK(balloc(&pm, sizeof(Z64) * 2 * b->neqns));
    BCHK(balloc(&dq, sizeof(double) * 9 * b->nsph));
    std::vector<double> hq(9 * (size_t)b->nsph);
    std::vector<double> w(b->nsph);
    for (int c = 0; c < b->ncl; c++) {
        double s3 = 0.0;
        for (int i = 0; i < b->counts[c]; i++) s3 += std::pow(b->radius[b->sph0[c] + i], 3);
        const double xv = std::pow(s3, (double)0.333333f);
        for (int i = 0; i < b->counts[c]; i++) {
            const double a = b->radius[b->sph0[c] + i];
            w[b->sph0[c] + i] = a * a / (xv * xv);
        }
    }
    hipError_t e = hipSuccess;
    const int gb = grid_of((int64_t)b->ng * 2, 256);
    for (int d = 0; d < b->ndir && e == hipSuccess; d++) {
        hipLaunchKernelGGL(layout_kernel, dim3(gb), dim3(256), 0, b->st, b->ng, b->d_ri, b->R, 2 * d, b->neqns, 0, am,
                           b->x64);
        hipLaunchKernelGGL(layout_kernel, dim3(gb), dim3(256), 0, b->st, b->ng, b->d_ri, b->R, 2 * d, b->neqns, 0, pm,
                           b->f64);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = launch_qeff((const double *)am, (const double *)pm, (int)(2 * b->neqns), b->nsph, b->d_radius,
                            b->d_order, b->d_offset, b->d_mieoff, (const double *)b->d_an, b->rimed, dq, b->st);
        if (e == hipSuccess) e = hipMemcpyAsync(hq.data(), dq, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, b->st);
        if (e == hipSuccess) e = hipStreamSynchronize(b->st);
        for (int c = 0; c < b->ncl && e == hipSuccess; c++) {
            double t[9] = {0};
            for (int i = 0; i < b->counts[c]; i++) {
                const int ig = b->sph0[c] + i;
                for (int k = 0; k < 9; k++) t[k] += w[ig] * hq[9 * (size_t)ig + k];
            }
            for (int k = 0; k < 3; k++) t[3 * k + 2] = t[3 * k] + t[3 * k + 2] - t[3 * k + 1];
            for (int k = 0; k < 9; k++) q[((size_t)c * b->ndir + d) * 9 + k] = t[k];
        }
    }
    bfree(am);
    bfree(pm);
    bfree(dq);
    if (e != hipSuccess) return bfail(b, MSTM_HIP_EHIP, "batch_qeff: %s", hipGetErrorString(e));
    return 0;
}

const char *mstm_hip_batch_stats_names(void)
{
    return "setup_ms,build_ms,tune_ms,solve_ms,inner_ms,gemm32_ms,gemm32_flops,gemm32_bytes,grid32,lpt,"
           "outer_steps,inner_iters,hsize,nrows,njobs32,gemm32_variant";
}

int mstm_hip_batch_stats(const mstm_hip_batch *b, double *out, int nout)
{
    if (!b || !out) return 0;
    double v[] = {b->setup_ms, b->build_ms, b->tune_ms, b->solve_ms, b->inner_ms, b->gemm32_ms, b->flops32,
                  b->bytes32, (double)b->grid32, (double)b->lpt, (double)b->outer_steps, (double)b->inner_iters,
                  (double)b->hsize, (double)b->nrows, (double)b->njobs32, (double)b->kvar};
    int k = std::min(nout, (int)(sizeof v / sizeof v[0]));
    for (int q = 0; q < k; q++) out[q] = v[q];
    return k;
}

// autotune table: (2*variant + longest-first, grid, ms) triples; returns the number of values
int mstm_hip_batch_tune_table(const mstm_hip_batch *b, double *out, int nout)
{
    if (!b || !out) return 0;
    int k = std::min(nout, (int)b->tune_table.size());
    for (int q = 0; q < k; q++) out[q] = b->tune_table[q];
    return k;
}

}  // extern "C"
