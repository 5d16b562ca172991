// learned.hip -- kernels of the learned initial guess: neighbour features,
// the FP32 MLP (dense layers, SiLU, loss, Adam) and the maps between the
// network's per-sphere outputs and MSTM coefficient vectors.  The model is
// described in learned.hpp.  Nothing here has a counterpart in MSTM; the
// coefficient packing is the reference's (mstm-modules-33.f90:2365-2377) and
// the incident phase is that of layerplanewavecoef (mstm-modules-33.f90:3901-3965)
// with beta = 0.
#include <algorithm>

#include "learned.hpp"

namespace {

constexpr int GT = 64;   // GEMM tile (rows and columns)
constexpr int GK = 16;   // GEMM k-step
constexpr int LOSS_PARTS = 512;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z)
{
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// features: one thread per (sphere, shell).  Global features (shell-0 thread):
//   0 depth below the entry surface / cutoff (clipped to 1)
//   1 height below the exit surface / cutoff (clipped to 1)
//   2 size parameter in the medium / 4
//   3 Re(m/m_medium) - 1, 4 Im(m/m_medium)
//   5 number of neighbours inside the cutoff / 150
// Shell s (edges 0, 1.25D, 2D, 3D, 4D with D the sphere diameter, cutoff 4D),
// sums over neighbours j with unit vector u = (r_j - r_i)/|r_j - r_i|:
//   count, u (3), traceless second moments (5), cos/sin(k dz),
//   (D/r) cos/sin(k (dz + r)), (D/r)^3;  every sum divided by 20.
// The entry/exit surfaces are the extreme z - a_j / z + a_j over the spheres
// whose projections on the xy plane overlap sphere i.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void lig_features_kernel(int nsph, const double *__restrict__ pos,
                                                           const double *__restrict__ radius,
                                                           const double *__restrict__ mre,
                                                           const double *__restrict__ mim, double rimed,
                                                           float *__restrict__ feat, float *__restrict__ depth)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsph * LIG_NSHELL) return;
    const int i = t / LIG_NSHELL, s = t % LIG_NSHELL;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2], ai = radius[i];
    const double D = 2.0 * ai, rc = 4.0 * D;
    const double r_lo = s == 0 ? 0.0 : (s == 1 ? 1.25 * D : (s == 2 ? 2.0 * D : 3.0 * D));
    const double r_hi = s == 0 ? 1.25 * D : (s == 1 ? 2.0 * D : (s == 2 ? 3.0 * D : rc));
    const double k = rimed;
    float acc[LIG_SHELLF];
#pragma unroll
    for (int q = 0; q < LIG_SHELLF; q++) acc[q] = 0.0f;
    double zmin = zi - ai, zmax = zi + ai;
    int ncut = 0;
    for (int j = 0; j < nsph; j++) {
        const double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi, aj = radius[j];
        const double rho2 = dx * dx + dy * dy;
        const double w = ai + aj;
        if (rho2 < w * w) {
            zmin = fmin(zmin, zi + dz - aj);
            zmax = fmax(zmax, zi + dz + aj);
        }
        if (j == i) continue;
        const double r2 = rho2 + dz * dz;
        if (r2 >= rc * rc) continue;
        ncut++;
        if (r2 < r_lo * r_lo || r2 >= r_hi * r_hi) continue;
        const double r = sqrt(r2);
        const float ux = (float)(dx / r), uy = (float)(dy / r), uz = (float)(dz / r);
        acc[0] += 1.0f;
        acc[1] += ux;
        acc[2] += uy;
        acc[3] += uz;
        acc[4] += ux * ux - uy * uy;
        acc[5] += 2.0f * uz * uz - ux * ux - uy * uy;
        acc[6] += ux * uy;
        acc[7] += ux * uz;
        acc[8] += uy * uz;
        const double ph = k * dz, ph2 = k * (dz + r), wr = D / r;
        acc[9] += (float)cos(ph);
        acc[10] += (float)sin(ph);
        acc[11] += (float)(wr * cos(ph2));
        acc[12] += (float)(wr * sin(ph2));
        acc[13] += (float)(wr * wr * wr);
    }
    float *f = feat + (size_t)i * LIG_NFEAT;
#pragma unroll
    for (int q = 0; q < LIG_SHELLF; q++) f[LIG_GLOBALF + s * LIG_SHELLF + q] = acc[q] * (1.0f / 20.0f);
    if (s == 0) {
        f[0] = (float)fmin((zi - zmin) / rc, 1.0);
        f[1] = (float)fmin((zmax - zi) / rc, 1.0);
        f[2] = (float)(k * ai / 4.0);
        f[3] = (float)(mre[i] / k - 1.0);
        f[4] = (float)(mim[i] / k);
        f[5] = (float)ncut / 150.0f;
        for (int q = LIG_GLOBALF + LIG_NSHELL * LIG_SHELLF; q < LIG_NFEAT; q++) f[q] = 0.0f;
        depth[i] = (float)(zi - zmin);
    }
}

// Training sample of one polarization: target = a_i exp(-i k z_i) and
// base = b_i exp(-i k z_i) in the order-Lnet layout (zeros where n > L_i).
// Output rows hold [pol 0 | pol 1], each 2*Lnet(Lnet+2) complex.
__global__ void lig_samples_kernel(int nsph, int pol, const int *__restrict__ order, const int *__restrict__ offset,
                                   const double *__restrict__ pos, double rimed,
                                   const cplx<double> *__restrict__ amnp, const cplx<double> *__restrict__ b, int Lnet,
                                   float *__restrict__ target, float *__restrict__ base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsph) return;
    const int L = order[i], NB = L * (L + 2), NBn = Lnet * (Lnet + 2);
    const double ph = rimed * pos[3 * i + 2];
    const cplx<double> e = cmk(cos(ph), -sin(ph));
    const size_t row = (size_t)i * 8 * NBn + (size_t)pol * 4 * NBn;
    float *ty = target + row, *tb = base + row;
    for (int q = 0; q < 4 * NBn; q++) ty[q] = tb[q] = 0.0f;
    const int nmax = imn(L, Lnet);
    for (int mode = 0; mode < 2; mode++)
        for (int n = 1; n <= nmax; n++)
            for (int m = -n; m <= n; m++) {
                const int src = offset[i] + mode * NB + amn(m, n, L);
                const int dst = 2 * (mode * NBn + amn(m, n, Lnet));
                const cplx<double> y = cmul(amnp[src], e), bb = cmul(b[src], e);
                ty[dst] = (float)y.re;
                ty[dst + 1] = (float)y.im;
                tb[dst] = (float)bb.re;
                tb[dst + 1] = (float)bb.im;
            }
}

// x_i = exp(i kappa d_i) (b_i + exp(i k z_i) out_i) for one polarization.
__global__ void lig_combine_kernel(int nsph, int pol, const int *__restrict__ order, const int *__restrict__ offset,
                                   const double *__restrict__ pos, double rimed, const float *__restrict__ out,
                                   int Lnet, const float *__restrict__ depth, const float *__restrict__ kappa,
                                   const cplx<double> *__restrict__ b, cplx<double> *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsph) return;
    const int L = order[i], NB = L * (L + 2), NBn = Lnet * (Lnet + 2);
    const double ph = rimed * pos[3 * i + 2];
    const cplx<double> eph = cmk(cos(ph), sin(ph));
    const double d = depth[i], amp = exp(-(double)kappa[1] * d), th = (double)kappa[0] * d;
    const cplx<double> E = cmk(amp * cos(th), amp * sin(th));
    const float *o = out + (size_t)i * 8 * NBn + (size_t)pol * 4 * NBn;
    for (int mode = 0; mode < 2; mode++)
        for (int n = 1; n <= L; n++)
            for (int m = -n; m <= n; m++) {
                const int idx = offset[i] + mode * NB + amn(m, n, L);
                cplx<double> v = b[idx];
                if (n <= Lnet) {
                    const int src = 2 * (mode * NBn + amn(m, n, Lnet));
                    cfma(v, eph, cmk((double)o[src], (double)o[src + 1]));
                }
                x[idx] = cmul(E, v);
            }
}

// ---------------------------------------------------------------------------
// C[M x N] = op(A) op(B) + bias; op(A)(m,k) = A[m*lda+k] (or A[k*lda+m] if TA),
// op(B)(k,n) = B[k*ldb+n] (or B[n*ldb+k] if TB).  64x64 tile per 256-thread
// workgroup, 4x4 outputs per thread, 16-deep k steps staged through LDS.
// act = 1 applies SiLU; Z (optional) receives the pre-activation.
// ---------------------------------------------------------------------------
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float *__restrict__ A, int lda,
                                                   const float *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                   float *__restrict__ C, int ldc, float *__restrict__ Z, int act)
{
    __shared__ float As[GK][GT + 4];
    __shared__ float Bs[GK][GT + 4];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += GK) {
        for (int e = tid; e < GT * GK; e += 256) {
            int mm, kk;
            if (TA) {
                mm = e % GT;
                kk = e / GT;
            } else {
                kk = e % GK;
                mm = e / GK;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.0f;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[kk][mm] = v;
        }
        for (int e = tid; e < GT * GK; e += 256) {
            int nn, kk;
            if (TB) {
                kk = e % GK;
                nn = e / GK;
            } else {
                nn = e % GT;
                kk = e / GT;
            }
            const int gn = n0 + nn, gk = k0 + kk;
            float v = 0.0f;
            if (gn < N && gk < K) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk++) {
            float av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = As[kk][ty * 4 + a];
#pragma unroll
            for (int b = 0; b < 4; b++) bv[b] = Bs[kk][tx * 4 + b];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = fmaf(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int gm = m0 + ty * 4 + a;
        if (gm >= M) continue;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int gn = n0 + tx * 4 + b;
            if (gn >= N) continue;
            float v = acc[a][b] + (bias ? bias[gn] : 0.0f);
            if (Z) Z[(size_t)gm * ldc + gn] = v;
            C[(size_t)gm * ldc + gn] = act ? silu(v) : v;
        }
    }
}

__global__ void silu_bwd_kernel(int n, const float *__restrict__ z, float *__restrict__ dh)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dh[i] *= silu_grad(z[i]);
}

// out[c] = sum_r a[r*cols + c]; one workgroup per 64 columns, rows split over 4 waves
__global__ __launch_bounds__(256) void colsum_kernel(int rows, int cols, const float *__restrict__ a,
                                                     float *__restrict__ out)
{
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + threadIdx.x % 64, w = threadIdx.x / 64;
    float s = 0.0f;
    if (c < cols)
        for (int r = w; r < rows; r += 4) s += a[(size_t)r * cols + c];
    red[w][threadIdx.x % 64] = s;
    __syncthreads();
    if (w == 0 && c < cols) out[c] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Loss over B samples of NC complex outputs: sum |exp(i kappa d)(base + O) - Y|^2 * inv_norm.
// Writes dO and per-workgroup partial (loss, dL/dkappa_re, dL/dkappa_im) sums.
__global__ __launch_bounds__(256) void lig_loss_kernel(int B, int NC, const float *__restrict__ O,
                                                       const float *__restrict__ Y, const float *__restrict__ base,
                                                       const float *__restrict__ depth,
                                                       const float *__restrict__ kappa, float *__restrict__ dO,
                                                       float inv_norm, double *__restrict__ partials)
{
    __shared__ double red[3][256];
    const float kr = kappa[0], ki = kappa[1];
    double sl = 0.0, skr = 0.0, ski = 0.0;
    const size_t tot = (size_t)B * NC;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const size_t s = t / NC;
        const float d = depth[s];
        const float amp = __expf(-ki * d);
        float sn, cs;
        __sincosf(kr * d, &sn, &cs);
        const float Er = amp * cs, Ei = amp * sn;
        const float ur = base[2 * t] + O[2 * t], ui = base[2 * t + 1] + O[2 * t + 1];
        const float yr = Er * ur - Ei * ui, yi = Er * ui + Ei * ur;   // yhat
        const float er = yr - Y[2 * t], ei = yi - Y[2 * t + 1];       // error
        // g = conj(e) E
        const float gr = er * Er + ei * Ei, gi = er * Ei - ei * Er;
        dO[2 * t] = 2.0f * gr * inv_norm;
        dO[2 * t + 1] = -2.0f * gi * inv_norm;
        // conj(e) yhat
        const float hr = er * yr + ei * yi, hi = er * yi - ei * yr;
        sl += (double)(er * er + ei * ei);
        skr += -2.0 * (double)d * (double)hi;
        ski += -2.0 * (double)d * (double)hr;
    }
    red[0][threadIdx.x] = sl;
    red[1][threadIdx.x] = skr;
    red[2][threadIdx.x] = ski;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) partials[3 * blockIdx.x + q] = red[q][0];
}

// sums the partials: loss_acc += loss * lscale, dkappa = dL/dkappa * gscale
__global__ __launch_bounds__(256) void lig_loss_finish_kernel(int nparts, const double *__restrict__ partials,
                                                              float gscale, double lscale, float *__restrict__ dkappa,
                                                              double *__restrict__ loss_acc)
{
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += 256)
        for (int q = 0; q < 3; q++) s[q] += partials[3 * b + q];
    for (int q = 0; q < 3; q++) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (loss_acc) loss_acc[0] += red[0][0] * lscale;
        if (dkappa) {
            dkappa[0] = (float)(red[1][0] * gscale);
            dkappa[1] = (float)(red[2][0] * gscale);
        }
    }
}

__global__ void adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

__global__ void gather_rows_kernel(int rows, int cols, const int *__restrict__ idx, const float *__restrict__ src,
                                   float *__restrict__ dst)
{
    const size_t tot = (size_t)rows * cols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t / cols, c = t % cols;
        dst[t] = src[(size_t)idx[r] * cols + c];
    }
}

int grid_for(size_t n, int bs, int cap) { return (int)std::min<size_t>((n + bs - 1) / bs, (size_t)cap); }

}  // namespace

int lig_loss_parts() { return LOSS_PARTS; }

hipError_t launch_lig_features(int nsph, const double *pos, const double *radius, const double *mre,
                               const double *mim, double rimed, float *feat, float *depth, hipStream_t st)
{
    if (nsph <= 0) return hipSuccess;
    const int nt = nsph * LIG_NSHELL;
    hipLaunchKernelGGL(lig_features_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, nsph, pos, radius, mre, mim,
                       rimed, feat, depth);
    return hipGetLastError();
}

hipError_t launch_lig_samples(int nsph, int pol, const int *order, const int *offset, const double *pos, double rimed,
                              const double *amnp, const double *b, int Lnet, float *target, float *base,
                              hipStream_t st)
{
    if (nsph <= 0) return hipSuccess;
    hipLaunchKernelGGL(lig_samples_kernel, dim3((nsph + 63) / 64), dim3(64), 0, st, nsph, pol, order, offset, pos,
                       rimed, reinterpret_cast<const cplx<double> *>(amnp), reinterpret_cast<const cplx<double> *>(b),
                       Lnet, target, base);
    return hipGetLastError();
}

hipError_t launch_lig_combine(int nsph, int pol, const int *order, const int *offset, const double *pos, double rimed,
                              const float *out, int Lnet, const float *depth, const float *kappa, const double *b,
                              double *x, hipStream_t st)
{
    if (nsph <= 0) return hipSuccess;
    hipLaunchKernelGGL(lig_combine_kernel, dim3((nsph + 63) / 64), dim3(64), 0, st, nsph, pol, order, offset, pos,
                       rimed, out, Lnet, depth, kappa, reinterpret_cast<const cplx<double> *>(b),
                       reinterpret_cast<cplx<double> *>(x));
    return hipGetLastError();
}

hipError_t launch_gemm(bool ta, bool tb, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                       const float *bias, float *C, int ldc, float *Z, int act, hipStream_t st)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    if (ta && tb) hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, bias, C, ldc, Z, act);
    else if (ta) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, bias, C, ldc, Z, act);
    else if (tb) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, bias, C, ldc, Z, act);
    else hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, bias, C, ldc, Z, act);
    return hipGetLastError();
}

hipError_t launch_silu_bwd(int n, const float *z, float *dh, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(silu_bwd_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, z, dh);
    return hipGetLastError();
}

hipError_t launch_colsum(int rows, int cols, const float *a, float *out, hipStream_t st)
{
    if (cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(256), 0, st, rows, cols, a, out);
    return hipGetLastError();
}

hipError_t launch_lig_loss(int B, int NC, const float *O, const float *Y, const float *base, const float *depth,
                           const float *kappa, float *dO, float inv_norm, double *partials, int nparts,
                           hipStream_t st)
{
    hipLaunchKernelGGL(lig_loss_kernel, dim3(nparts), dim3(256), 0, st, B, NC, O, Y, base, depth, kappa, dO, inv_norm,
                       partials);
    return hipGetLastError();
}

hipError_t launch_lig_loss_finish(int nparts, const double *partials, float gscale, double lscale, float *dkappa,
                                  double *loss_acc, hipStream_t st)
{
    hipLaunchKernelGGL(lig_loss_finish_kernel, dim3(1), dim3(256), 0, st, nparts, partials, gscale, lscale, dkappa,
                       loss_acc);
    return hipGetLastError();
}

hipError_t launch_adam(int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, n, p, g, m, v, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError();
}

hipError_t launch_gather_rows(int rows, int cols, const int *idx, const float *src, float *dst, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for((size_t)rows * cols, 256, 8192)), dim3(256), 0, st, rows,
                       cols, idx, src, dst);
    return hipGetLastError();
}

// ===========================================================================
// Host side: MLP parameter layout, forward/backward drivers, trainer C-ABI.
// Parameters are one flat FP32 array: for each layer l = 0..nlayers
// (nlayers hidden SiLU layers of width `hidden`, then a linear output layer)
// the weight W_l [out x in] row-major followed by the bias b_l [out]; then
// kappa (re, im).  Input width LIG_NFEAT, output width 8*Lnet(Lnet+2).
// ===========================================================================
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mstm_hip.h"

size_t mlp_nparams(const MlpShape &s)
{
    size_t n = 0;
    int in = s.nin;
    for (int l = 0; l <= s.nlayers; l++) {
        const int out = (l == s.nlayers) ? s.nout : s.hidden;
        n += (size_t)out * in + out;
        in = out;
    }
    return n + 2;
}

static size_t mlp_woff(const MlpShape &s, int l)
{
    size_t n = 0;
    int in = s.nin;
    for (int q = 0; q < l; q++) {
        const int out = (q == s.nlayers) ? s.nout : s.hidden;
        n += (size_t)out * in + out;
        in = out;
    }
    return n;
}

hipError_t mlp_forward(const MlpShape &s, const float *params, const float *X, int B, float *const *H, float *const *Z,
                       float *out, hipStream_t st)
{
    const float *in = X;
    int nin = s.nin;
    for (int l = 0; l <= s.nlayers; l++) {
        const bool last = (l == s.nlayers);
        const int nout = last ? s.nout : s.hidden;
        const float *W = params + mlp_woff(s, l), *b = W + (size_t)nout * nin;
        float *dst = last ? out : H[l];
        hipError_t e = launch_gemm(false, true, B, nout, nin, in, nin, W, nin, b, dst, nout,
                                   (!last && Z) ? Z[l] : nullptr, last ? 0 : 1, st);
        if (e != hipSuccess) return e;
        in = dst;
        nin = nout;
    }
    return hipSuccess;
}

// gradient of the loss w.r.t. every parameter except kappa, given dO
hipError_t mlp_backward(const MlpShape &s, const float *params, const float *X, int B, float *const *H,
                        float *const *Z, const float *dO, float *dA, float *dB, float *grad, hipStream_t st)
{
    hipError_t e;
    const int h = s.hidden;
    {   // output layer
        const size_t o = mlp_woff(s, s.nlayers);
        const float *W = params + o;
        e = launch_gemm(true, false, s.nout, h, B, dO, s.nout, H[s.nlayers - 1], h, nullptr, grad + o, h, nullptr, 0,
                        st);
        if (e != hipSuccess) return e;
        e = launch_colsum(B, s.nout, dO, grad + o + (size_t)s.nout * h, st);
        if (e != hipSuccess) return e;
        e = launch_gemm(false, false, B, h, s.nout, dO, s.nout, W, h, nullptr, dA, h, nullptr, 0, st);
        if (e != hipSuccess) return e;
    }
    for (int l = s.nlayers - 1; l >= 0; l--) {
        const int nin = (l == 0) ? s.nin : h;
        const float *in = (l == 0) ? X : H[l - 1];
        const size_t o = mlp_woff(s, l);
        e = launch_silu_bwd(B * h, Z[l], dA, st);  // dA: dL/dZ_l
        if (e != hipSuccess) return e;
        e = launch_gemm(true, false, h, nin, B, dA, h, in, nin, nullptr, grad + o, nin, nullptr, 0, st);
        if (e != hipSuccess) return e;
        e = launch_colsum(B, h, dA, grad + o + (size_t)h * nin, st);
        if (e != hipSuccess) return e;
        if (l > 0) {
            e = launch_gemm(false, false, B, nin, h, dA, h, params + o, nin, nullptr, dB, nin, nullptr, 0, st);
            if (e != hipSuccess) return e;
            std::swap(dA, dB);
        }
    }
    return hipSuccess;
}

void mlp_init(const MlpShape &s, unsigned seed, std::vector<float> &p)
{
    p.assign(mlp_nparams(s), 0.0f);
    std::mt19937 rng(seed);
    int in = s.nin;
    for (int l = 0; l <= s.nlayers; l++) {
        const bool last = (l == s.nlayers);
        const int out = last ? s.nout : s.hidden;
        std::normal_distribution<float> nd(0.0f, (last ? 0.01f : 1.0f) / std::sqrt((float)in));
        float *W = p.data() + mlp_woff(s, l);
        for (size_t q = 0; q < (size_t)out * in; q++) W[q] = nd(rng);
        in = out;
    }
}

struct mstm_hip_trainer {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    MlpShape s{};
    size_t np = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    long step = 0;
    // data
    int nsamp = 0;
    double yscale = 1.0;
    float *X = nullptr, *Y = nullptr, *Bs = nullptr, *Dp = nullptr;
    int *idx = nullptr;
    // batch buffers
    int cap = 0;
    float *bx = nullptr, *by = nullptr, *bb = nullptr, *bd = nullptr, *out = nullptr, *dO = nullptr, *dA = nullptr,
          *dB = nullptr;
    std::vector<float *> H, Z;
    double *partials = nullptr, *lacc = nullptr;
};

static int tfail(mstm_hip_trainer *t, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (t) t->err = buf;
    return code;
}

#define TCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return tfail(t, MSTM_HIP_EHIP, "%s: %s", #x, hipGetErrorString(e_));  \
    } while (0)

template <typename T> static hipError_t talloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return hipSuccess;
    return hipMalloc((void **)p, n * sizeof(T));
}
template <typename T> static void tfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

static void free_batch(mstm_hip_trainer *t)
{
    tfree(t->bx), tfree(t->by), tfree(t->bb), tfree(t->bd), tfree(t->out), tfree(t->dO), tfree(t->dA), tfree(t->dB);
    for (auto &q : t->H) tfree(q);
    for (auto &q : t->Z) tfree(q);
    t->cap = 0;
}

static int ensure_batch(mstm_hip_trainer *t, int B)
{
    if (B <= t->cap) return 0;
    free_batch(t);
    const MlpShape &s = t->s;
    TCHK(talloc(&t->bx, (size_t)B * s.nin));
    TCHK(talloc(&t->by, (size_t)B * s.nout));
    TCHK(talloc(&t->bb, (size_t)B * s.nout));
    TCHK(talloc(&t->bd, (size_t)B));
    TCHK(talloc(&t->out, (size_t)B * s.nout));
    TCHK(talloc(&t->dO, (size_t)B * s.nout));
    TCHK(talloc(&t->dA, (size_t)B * s.hidden));
    TCHK(talloc(&t->dB, (size_t)B * s.hidden));
    t->H.assign(s.nlayers, nullptr);
    t->Z.assign(s.nlayers, nullptr);
    for (int l = 0; l < s.nlayers; l++) {
        TCHK(talloc(&t->H[l], (size_t)B * s.hidden));
        TCHK(talloc(&t->Z[l], (size_t)B * s.hidden));
    }
    t->cap = B;
    return 0;
}

// forward + loss (+ backward into t->g when grad) on B rows already in the batch buffers
static int batch_pass(mstm_hip_trainer *t, int B, bool grad, double lscale)
{
    const MlpShape &s = t->s;
    const int NC = s.nout / 2;
    const float inv_norm = (float)(1.0 / ((double)B * t->yscale));
    TCHK(mlp_forward(s, t->p, t->bx, B, t->H.data(), t->Z.data(), t->out, t->st));
    const int nparts = lig_loss_parts();
    TCHK(launch_lig_loss(B, NC, t->out, t->by, t->bb, t->bd, t->p + t->np - 2, t->dO, inv_norm, t->partials, nparts,
                         t->st));
    TCHK(launch_lig_loss_finish(nparts, t->partials, inv_norm, lscale / t->yscale, grad ? t->g + t->np - 2 : nullptr,
                                t->lacc, t->st));
    if (grad)
        TCHK(mlp_backward(s, t->p, t->bx, B, t->H.data(), t->Z.data(), t->dO, t->dA, t->dB, t->g, t->st));
    return 0;
}

static int load_rows(mstm_hip_trainer *t, const int *idx, int B)
{
    const MlpShape &s = t->s;
    TCHK(launch_gather_rows(B, s.nin, idx, t->X, t->bx, t->st));
    TCHK(launch_gather_rows(B, s.nout, idx, t->Y, t->by, t->st));
    TCHK(launch_gather_rows(B, s.nout, idx, t->Bs, t->bb, t->st));
    TCHK(launch_gather_rows(B, 1, idx, t->Dp, t->bd, t->st));
    return 0;
}

extern "C" {

int mstm_hip_lig_nparams(int hidden, int nlayers, int Lnet)
{
    if (hidden <= 0 || nlayers <= 0 || Lnet <= 0) return 0;
    MlpShape s{LIG_NFEAT, hidden, nlayers, 8 * Lnet * (Lnet + 2)};
    return (int)mlp_nparams(s);
}

int mstm_hip_trainer_create(mstm_hip_trainer **out, int device, int hidden, int nlayers, int Lnet, unsigned seed)
{
    if (!out) return MSTM_HIP_EARG;
    *out = nullptr;
    if (hidden <= 0 || hidden > 1024 || nlayers <= 0 || nlayers > 8 || Lnet <= 0 || Lnet > MSTM_MAX_L)
        return MSTM_HIP_EARG;
    auto *t = new mstm_hip_trainer();
    t->device = device;
    t->s = MlpShape{LIG_NFEAT, hidden, nlayers, 8 * Lnet * (Lnet + 2)};
    t->np = mlp_nparams(t->s);
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = talloc(&t->p, t->np);
    if (e == hipSuccess) e = talloc(&t->g, t->np);
    if (e == hipSuccess) e = talloc(&t->m, t->np);
    if (e == hipSuccess) e = talloc(&t->v, t->np);
    if (e == hipSuccess) e = talloc(&t->partials, (size_t)3 * lig_loss_parts());
    if (e == hipSuccess) e = talloc(&t->lacc, 1);
    std::vector<float> p;
    mlp_init(t->s, seed, p);
    if (e == hipSuccess) e = hipMemcpy(t->p, p.data(), sizeof(float) * t->np, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(t->m, 0, sizeof(float) * t->np);
    if (e == hipSuccess) e = hipMemset(t->v, 0, sizeof(float) * t->np);
    if (e == hipSuccess) e = hipMemset(t->g, 0, sizeof(float) * t->np);
    *out = t;
    return e == hipSuccess ? MSTM_HIP_OK : tfail(t, MSTM_HIP_EHIP, "trainer_create: %s", hipGetErrorString(e));
}

int mstm_hip_trainer_destroy(mstm_hip_trainer *t)
{
    if (!t) return MSTM_HIP_OK;
    (void)hipSetDevice(t->device);
    if (t->st) (void)hipStreamSynchronize(t->st);
    free_batch(t);
    tfree(t->p), tfree(t->g), tfree(t->m), tfree(t->v), tfree(t->partials), tfree(t->lacc);
    tfree(t->X), tfree(t->Y), tfree(t->Bs), tfree(t->Dp), tfree(t->idx);
    if (t->st) (void)hipStreamDestroy(t->st);
    delete t;
    return MSTM_HIP_OK;
}

const char *mstm_hip_trainer_last_error(const mstm_hip_trainer *t) { return t ? t->err.c_str() : "null trainer"; }

int mstm_hip_trainer_params(mstm_hip_trainer *t, float *params, int nparams, int set)
{
    if (!t || !params) return MSTM_HIP_EARG;
    if ((size_t)nparams != t->np) return tfail(t, MSTM_HIP_EARG, "params: expected %zu values, got %d", t->np, nparams);
    TCHK(hipSetDevice(t->device));
    if (set) {
        TCHK(hipMemcpyAsync(t->p, params, sizeof(float) * t->np, hipMemcpyHostToDevice, t->st));
        TCHK(hipMemsetAsync(t->m, 0, sizeof(float) * t->np, t->st));
        TCHK(hipMemsetAsync(t->v, 0, sizeof(float) * t->np, t->st));
        t->step = 0;
    } else {
        TCHK(hipMemcpyAsync(params, t->p, sizeof(float) * t->np, hipMemcpyDeviceToHost, t->st));
    }
    TCHK(hipStreamSynchronize(t->st));
    return 0;
}

int mstm_hip_trainer_set_data(mstm_hip_trainer *t, int nsamp, const float *feat, const float *target,
                              const float *base, const float *depth)
{
    if (!t || nsamp <= 0 || !feat || !target || !base || !depth) return MSTM_HIP_EARG;
    TCHK(hipSetDevice(t->device));
    tfree(t->X), tfree(t->Y), tfree(t->Bs), tfree(t->Dp), tfree(t->idx);
    const MlpShape &s = t->s;
    TCHK(talloc(&t->X, (size_t)nsamp * s.nin));
    TCHK(talloc(&t->Y, (size_t)nsamp * s.nout));
    TCHK(talloc(&t->Bs, (size_t)nsamp * s.nout));
    TCHK(talloc(&t->Dp, (size_t)nsamp));
    TCHK(talloc(&t->idx, (size_t)nsamp));
    TCHK(hipMemcpy(t->X, feat, sizeof(float) * nsamp * s.nin, hipMemcpyHostToDevice));
    TCHK(hipMemcpy(t->Y, target, sizeof(float) * nsamp * s.nout, hipMemcpyHostToDevice));
    TCHK(hipMemcpy(t->Bs, base, sizeof(float) * nsamp * s.nout, hipMemcpyHostToDevice));
    TCHK(hipMemcpy(t->Dp, depth, sizeof(float) * nsamp, hipMemcpyHostToDevice));
    double ys = 0.0;
    for (size_t q = 0; q < (size_t)nsamp * s.nout; q++) ys += (double)target[q] * target[q];
    t->yscale = ys > 0.0 ? ys / nsamp : 1.0;
    t->nsamp = nsamp;
    return 0;
}

// One pass over the data in shuffled mini-batches with Adam.  mean_loss is
// the sample mean of |yhat - y|^2 / <|y|^2> seen during the pass.
int mstm_hip_trainer_epoch(mstm_hip_trainer *t, int batch, float lr, unsigned seed, double *mean_loss)
{
    if (!t || batch <= 0) return MSTM_HIP_EARG;
    if (!t->nsamp) return tfail(t, MSTM_HIP_ESTATE, "epoch: no data");
    TCHK(hipSetDevice(t->device));
    int rc = ensure_batch(t, batch);
    if (rc) return rc;
    std::vector<int> perm(t->nsamp);
    for (int q = 0; q < t->nsamp; q++) perm[q] = q;
    std::mt19937 rng(seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    TCHK(hipMemcpyAsync(t->idx, perm.data(), sizeof(int) * t->nsamp, hipMemcpyHostToDevice, t->st));
    TCHK(hipMemsetAsync(t->lacc, 0, sizeof(double), t->st));
    const float b1 = 0.9f, b2 = 0.999f;
    for (int r0 = 0; r0 < t->nsamp; r0 += batch) {
        const int B = std::min(batch, t->nsamp - r0);
        rc = load_rows(t, t->idx + r0, B);
        if (rc) return rc;
        rc = batch_pass(t, B, true, 1.0);
        if (rc) return rc;
        t->step++;
        const float c1 = 1.0f / (1.0f - std::pow(b1, (float)t->step));
        const float c2 = 1.0f / (1.0f - std::pow(b2, (float)t->step));
        TCHK(launch_adam((int)t->np, t->p, t->g, t->m, t->v, lr, b1, b2, 1e-8f, c1, c2, t->st));
    }
    double l = 0.0;
    TCHK(hipMemcpyAsync(&l, t->lacc, sizeof(double), hipMemcpyDeviceToHost, t->st));
    TCHK(hipStreamSynchronize(t->st));
    if (mean_loss) *mean_loss = l / t->nsamp;
    return 0;
}

// loss on the whole data set, no update
int mstm_hip_trainer_loss(mstm_hip_trainer *t, int batch, double *loss)
{
    if (!t || batch <= 0 || !loss) return MSTM_HIP_EARG;
    if (!t->nsamp) return tfail(t, MSTM_HIP_ESTATE, "loss: no data");
    TCHK(hipSetDevice(t->device));
    int rc = ensure_batch(t, batch);
    if (rc) return rc;
    std::vector<int> ident(t->nsamp);
    for (int q = 0; q < t->nsamp; q++) ident[q] = q;
    TCHK(hipMemcpyAsync(t->idx, ident.data(), sizeof(int) * t->nsamp, hipMemcpyHostToDevice, t->st));
    TCHK(hipMemsetAsync(t->lacc, 0, sizeof(double), t->st));
    for (int r0 = 0; r0 < t->nsamp; r0 += batch) {
        const int B = std::min(batch, t->nsamp - r0);
        rc = load_rows(t, t->idx + r0, B);
        if (rc) return rc;
        rc = batch_pass(t, B, false, 1.0);
        if (rc) return rc;
    }
    double l = 0.0;
    TCHK(hipMemcpyAsync(&l, t->lacc, sizeof(double), hipMemcpyDeviceToHost, t->st));
    TCHK(hipStreamSynchronize(t->st));
    *loss = l / t->nsamp;
    return 0;
}

// Gradient (all parameters, kappa last) and loss of the mean loss over the
// given samples, at the current parameters.  For checking the kernels.
int mstm_hip_trainer_grad(mstm_hip_trainer *t, int nsamp, const float *feat, const float *target, const float *base,
                          const float *depth, float *grad, double *loss)
{
    if (!t || nsamp <= 0 || !feat || !target || !base || !depth || !grad) return MSTM_HIP_EARG;
    TCHK(hipSetDevice(t->device));
    int rc = ensure_batch(t, nsamp);
    if (rc) return rc;
    const MlpShape &s = t->s;
    TCHK(hipMemcpyAsync(t->bx, feat, sizeof(float) * nsamp * s.nin, hipMemcpyHostToDevice, t->st));
    TCHK(hipMemcpyAsync(t->by, target, sizeof(float) * nsamp * s.nout, hipMemcpyHostToDevice, t->st));
    TCHK(hipMemcpyAsync(t->bb, base, sizeof(float) * nsamp * s.nout, hipMemcpyHostToDevice, t->st));
    TCHK(hipMemcpyAsync(t->bd, depth, sizeof(float) * nsamp, hipMemcpyHostToDevice, t->st));
    double ys = 0.0;
    for (size_t q = 0; q < (size_t)nsamp * s.nout; q++) ys += (double)target[q] * target[q];
    const double keep = t->yscale;
    t->yscale = ys > 0.0 ? ys / nsamp : 1.0;
    TCHK(hipMemsetAsync(t->lacc, 0, sizeof(double), t->st));
    rc = batch_pass(t, nsamp, true, 1.0 / nsamp);
    t->yscale = keep;
    if (rc) return rc;
    TCHK(hipMemcpyAsync(grad, t->g, sizeof(float) * t->np, hipMemcpyDeviceToHost, t->st));
    double l = 0.0;
    TCHK(hipMemcpyAsync(&l, t->lacc, sizeof(double), hipMemcpyDeviceToHost, t->st));
    TCHK(hipStreamSynchronize(t->st));
    if (loss) *loss = l;
    return 0;
}

// network outputs for nsamp feature rows
int mstm_hip_trainer_forward(mstm_hip_trainer *t, int nsamp, const float *feat, float *out)
{
    if (!t || nsamp <= 0 || !feat || !out) return MSTM_HIP_EARG;
    TCHK(hipSetDevice(t->device));
    int rc = ensure_batch(t, nsamp);
    if (rc) return rc;
    const MlpShape &s = t->s;
    TCHK(hipMemcpyAsync(t->bx, feat, sizeof(float) * nsamp * s.nin, hipMemcpyHostToDevice, t->st));
    TCHK(mlp_forward(s, t->p, t->bx, nsamp, t->H.data(), nullptr, t->out, t->st));
    TCHK(hipMemcpyAsync(out, t->out, sizeof(float) * nsamp * s.nout, hipMemcpyDeviceToHost, t->st));
    TCHK(hipStreamSynchronize(t->st));
    return 0;
}

}  // extern "C"
