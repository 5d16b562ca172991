// krylov.hip -- FP64 vector kernels of the complex biconjugate-gradient
// iteration (cbicg, mstm-solver-8.f90:654-949) and of the mixed-precision
// refinement.  Inner products are written as per-block partial sums that the
// host adds in a fixed order (deterministic results).
#include "common.hpp"
#include "kernels.hpp"

#define KB 256  // threads per block
#define KGRID 512

__device__ __forceinline__ void block_sum3(cplx<double> *v, int nv)
{
    __shared__ cplx<double> sh[3][KB];
    for (int q = 0; q < nv; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = KB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < nv; q++) sh[q][threadIdx.x] = cadd(sh[q][threadIdx.x], sh[q][threadIdx.x + s]);
        __syncthreads();
    }
    for (int q = 0; q < nv; q++) v[q] = sh[q][0];
}

// partial[b] = sum (conj_a ? conj(a) : a) * b over the block's grid-stride share
__global__ void zdot_kernel(const cplx<double> *__restrict__ a, const cplx<double> *__restrict__ b, int n, int conj_a,
                            cplx<double> *__restrict__ partials)
{
    cplx<double> s = cmk(0.0, 0.0);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> x = a[i];
        if (conj_a) x = cconj(x);
        cfma(s, x, b[i]);
    }
    block_sum3(&s, 1);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

hipError_t launch_zdot(const double *a, const double *b, int n, int conj_a, double *partials, double *, hipStream_t st)
{
    hipLaunchKernelGGL(zdot_kernel, dim3(KGRID), dim3(KB), 0, st, reinterpret_cast<const cplx<double> *>(a),
                       reinterpret_cast<const cplx<double> *>(b), n, conj_a, reinterpret_cast<cplx<double> *>(partials));
    return hipGetLastError();
}

// cap = cp - cap ; caw = cw - caw ; partial <cw, cap>   (:897-903)
__global__ void bicg_a_kernel(cplx<double> *__restrict__ cap, cplx<double> *__restrict__ caw,
                              const cplx<double> *__restrict__ cp, const cplx<double> *__restrict__ cw, int n,
                              cplx<double> *__restrict__ partials)
{
    cplx<double> s = cmk(0.0, 0.0);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> a = csub(cp[i], cap[i]);
        cap[i] = a;
        caw[i] = csub(cw[i], caw[i]);
        cfma(s, cconj(cw[i]), a);
    }
    block_sum3(&s, 1);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// anp += cak cp ; cr -= cak cap ; cq -= conj(cak) caw ;
// partials <cq, cr>, <cr, cr>   (:912-920)
__global__ void bicg_b_kernel(cplx<double> *__restrict__ anp, cplx<double> *__restrict__ cr,
                              cplx<double> *__restrict__ cq, const cplx<double> *__restrict__ cp,
                              const cplx<double> *__restrict__ cap, const cplx<double> *__restrict__ caw,
                              cplx<double> cak, int n, cplx<double> *__restrict__ partials)
{
    cplx<double> s[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
    cplx<double> cakc = cconj(cak);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> a = anp[i];
        cfma(a, cak, cp[i]);
        anp[i] = a;
        cplx<double> r = cr[i];
        cfma(r, cscale(cak, -1.0), cap[i]);
        cr[i] = r;
        cplx<double> q = cq[i];
        cfma(q, cscale(cakc, -1.0), caw[i]);
        cq[i] = q;
        cfma(s[0], cconj(q), r);
        cfma(s[1], cconj(r), r);
    }
    block_sum3(s, 2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s[0];
        partials[2 * blockIdx.x + 1] = s[1];
    }
}

// cp = cr + cbk cp ; cw = cq + conj(cbk) cw   (:928-929)
__global__ void bicg_c_kernel(cplx<double> *__restrict__ cp, cplx<double> *__restrict__ cw,
                              const cplx<double> *__restrict__ cr, const cplx<double> *__restrict__ cq,
                              cplx<double> cbk, int n)
{
    cplx<double> cbkc = cconj(cbk);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> p = cr[i];
        cfma(p, cbk, cp[i]);
        cp[i] = p;
        cplx<double> w = cq[i];
        cfma(w, cbkc, cw[i]);
        cw[i] = w;
    }
}

// cr = pnp - anp + cr ; cq = cw = conj(cr) ; cp = cr ; partial sum cr*cr
// (bilinear, :828-834)
__global__ void bicg_init_kernel(cplx<double> *__restrict__ cr, cplx<double> *__restrict__ cq,
                                 cplx<double> *__restrict__ cw, cplx<double> *__restrict__ cp,
                                 const cplx<double> *__restrict__ pnp, const cplx<double> *__restrict__ anp, int n,
                                 cplx<double> *__restrict__ partials)
{
    cplx<double> s = cmk(0.0, 0.0);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> r = cadd(csub(pnp[i], anp[i]), cr[i]);
        cr[i] = r;
        cq[i] = cconj(r);
        cw[i] = cconj(r);
        cp[i] = r;
        cfma(s, r, r);
    }
    block_sum3(&s, 1);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// y = a*x + b*y (complex scalars a, b)
__global__ void zaxpby_kernel(cplx<double> a, const cplx<double> *__restrict__ x, cplx<double> b,
                              cplx<double> *__restrict__ y, int n)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> v = cmul(b, y[i]);
        cfma(v, a, x[i]);
        y[i] = v;
    }
}

int krylov_grid() { return KGRID; }

// ---------------------------------------------------------------------------
// Device-resident BiCG scalars (see BicgScal in kernels.hpp).  The vector
// kernels of a system whose state is inactive return at once, so the host
// may launch one iteration for both systems and learn only afterwards which
// of them stopped.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(KB) void bicg_scal_kernel(int step, const cplx<double> *__restrict__ partials,
                                                       int64_t region, BicgScal *__restrict__ sc, int ns,
                                                       int *__restrict__ gate)
{
    __shared__ cplx<double> red[2][KB];
    const int nv = step == 3 ? 2 : 1;
    for (int s = 0; s < ns; s++) {
        BicgScal &S = sc[s];
        const bool skip = step >= 1 && !S.active;  // uniform: read by all threads
        if (!skip) {
            const cplx<double> *p = partials + region * s;
            for (int q = 0; q < nv; q++) {
                cplx<double> v = cmk(0.0, 0.0);
                for (int b = threadIdx.x; b < KGRID; b += KB) v = cadd(v, p[(int64_t)nv * b + q]);
                red[q][threadIdx.x] = v;
            }
            __syncthreads();
            for (int h = KB / 2; h > 0; h >>= 1) {
                if (threadIdx.x < h)
                    for (int q = 0; q < nv; q++) red[q][threadIdx.x] = cadd(red[q][threadIdx.x], red[q][threadIdx.x + h]);
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            const cplx<double> v0 = red[0][0], v1 = red[1][0];
            const cplx<double> zero = cmk(0.0, 0.0);
            if (step == 0) {
                S.enorm = v0.re;
                S.active = v0.re != 0.0;
                S.it = 0;
                S.brk = 0;
                S.cur = 0;
                S.err = 0.0;
                S.cak = S.cbk = zero;
            } else if (skip) {
                S.cak = S.cbk = zero;
            } else if (step == 1) {
                S.csk = v0;
                if (v0.re == 0.0 && v0.im == 0.0) S.active = 0;  // it stays 0
            } else if (step == 2) {
                S.cur++;
                if (v0.re == 0.0 && v0.im == 0.0) {  // breakdown: the Fortran loop exits here
                    S.active = 0;
                    S.it = S.cur;
                    S.brk = 1;
                    S.cak = zero;
                } else {
                    S.cak = cdiv(S.csk, v0);
                }
            } else {
                S.err = v1.re / S.enorm;
                if (S.err < S.eps || (S.csk.re == 0.0 && S.csk.im == 0.0)) {
                    S.active = 0;
                    S.it = S.cur;
                    S.cbk = zero;
                } else {
                    S.cbk = cdiv(v0, S.csk);
                    S.csk = v0;
                }
            }
        }
        __syncthreads();
    }
    if (gate && threadIdx.x == 0) {
        int any = 0;
        for (int s = 0; s < ns; s++) any |= sc[s].active != 0;
        *gate = any;
    }
}

// The vector steps of ns systems in one launch: system s = blockIdx.y has
// its vectors at the base pointers plus s times a stride (complex elements:
// sv for x, r, q; sp for p, w, Ap, A^H w), its scalars at sc[s] and its
// partial sums at partials + s * region; a stopped system returns at once.
__global__ void bicg_a_dev_kernel(cplx<double> *__restrict__ cap, cplx<double> *__restrict__ caw,
                                  const cplx<double> *__restrict__ cp, const cplx<double> *__restrict__ cw, int64_t sp,
                                  int n, cplx<double> *__restrict__ partials, int64_t region,
                                  const BicgScal *__restrict__ sc)
{
    const int64_t y = blockIdx.y;
    if (!sc[y].active) return;
    cap += y * sp;
    caw += y * sp;
    cp += y * sp;
    cw += y * sp;
    partials += y * region;
    cplx<double> s = cmk(0.0, 0.0);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> a = csub(cp[i], cap[i]);
        cap[i] = a;
        caw[i] = csub(cw[i], caw[i]);
        cfma(s, cconj(cw[i]), a);
    }
    block_sum3(&s, 1);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void bicg_b_dev_kernel(cplx<double> *__restrict__ anp, cplx<double> *__restrict__ cr,
                                  cplx<double> *__restrict__ cq, const cplx<double> *__restrict__ cp,
                                  const cplx<double> *__restrict__ cap, const cplx<double> *__restrict__ caw,
                                  int64_t sv, int64_t sp, const BicgScal *__restrict__ sc, int n,
                                  cplx<double> *__restrict__ partials, int64_t region)
{
    const int64_t y = blockIdx.y;
    if (!sc[y].active) return;
    anp += y * sv;
    cr += y * sv;
    cq += y * sv;
    cp += y * sp;
    cap += y * sp;
    caw += y * sp;
    partials += y * region;
    const cplx<double> cak = sc[y].cak, cakc = cconj(cak);
    cplx<double> s[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> a = anp[i];
        cfma(a, cak, cp[i]);
        anp[i] = a;
        cplx<double> r = cr[i];
        cfma(r, cscale(cak, -1.0), cap[i]);
        cr[i] = r;
        cplx<double> q = cq[i];
        cfma(q, cscale(cakc, -1.0), caw[i]);
        cq[i] = q;
        cfma(s[0], cconj(q), r);
        cfma(s[1], cconj(r), r);
    }
    block_sum3(s, 2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s[0];
        partials[2 * blockIdx.x + 1] = s[1];
    }
}

__global__ void bicg_c_dev_kernel(cplx<double> *__restrict__ cp, cplx<double> *__restrict__ cw,
                                  const cplx<double> *__restrict__ cr, const cplx<double> *__restrict__ cq, int64_t sv,
                                  int64_t sp, const BicgScal *__restrict__ sc, int n)
{
    const int64_t y = blockIdx.y;
    if (!sc[y].active) return;
    cp += y * sp;
    cw += y * sp;
    cr += y * sv;
    cq += y * sv;
    const cplx<double> cbk = sc[y].cbk, cbkc = cconj(cbk);
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> p = cr[i];
        cfma(p, cbk, cp[i]);
        cp[i] = p;
        cplx<double> w = cq[i];
        cfma(w, cbkc, cw[i]);
        cw[i] = w;
    }
}

hipError_t launch_bicg_scal(int step, const double *partials, int64_t region, BicgScal *sc, int ns, hipStream_t st,
                            int *gate)
{
    if (ns < 1 || ns > 2 || step < 0 || step > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bicg_scal_kernel, dim3(1), dim3(KB), 0, st, step,
                       reinterpret_cast<const cplx<double> *>(partials), region, sc, ns, gate);
    return hipGetLastError();
}
hipError_t launch_bicg_a_dev(double *cap, double *caw, const double *cp, const double *cw, int64_t sp, int n,
                             double *partials, int64_t region, const BicgScal *sc, int ns, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_a_dev_kernel, dim3(KGRID, ns), dim3(KB), 0, st, (Z *)cap, (Z *)caw, (const Z *)cp,
                       (const Z *)cw, sp / 2, n, (Z *)partials, region, sc);
    return hipGetLastError();
}
hipError_t launch_bicg_b_dev(double *anp, double *cr, double *cq, const double *cp, const double *cap,
                             const double *caw, int64_t sv, int64_t sp, const BicgScal *sc, int n, double *partials,
                             int64_t region, int ns, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_b_dev_kernel, dim3(KGRID, ns), dim3(KB), 0, st, (Z *)anp, (Z *)cr, (Z *)cq, (const Z *)cp,
                       (const Z *)cap, (const Z *)caw, sv / 2, sp / 2, sc, n, (Z *)partials, region);
    return hipGetLastError();
}
hipError_t launch_bicg_c_dev(double *cp, double *cw, const double *cr, const double *cq, int64_t sv, int64_t sp,
                             const BicgScal *sc, int n, int ns, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_c_dev_kernel, dim3(KGRID, ns), dim3(KB), 0, st, (Z *)cp, (Z *)cw, (const Z *)cr,
                       (const Z *)cq, sv / 2, sp / 2, sc, n);
    return hipGetLastError();
}
hipError_t launch_bicg_a(double *cap, double *caw, const double *cp, const double *cw, int n, double *partials,
                         hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_a_kernel, dim3(KGRID), dim3(KB), 0, st, (Z *)cap, (Z *)caw, (const Z *)cp, (const Z *)cw, n,
                       (Z *)partials);
    return hipGetLastError();
}
hipError_t launch_bicg_b(double *anp, double *cr, double *cq, const double *cp, const double *cap, const double *caw,
                         double cakr, double caki, int n, double *partials, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_b_kernel, dim3(KGRID), dim3(KB), 0, st, (Z *)anp, (Z *)cr, (Z *)cq, (const Z *)cp,
                       (const Z *)cap, (const Z *)caw, cmk(cakr, caki), n, (Z *)partials);
    return hipGetLastError();
}
hipError_t launch_bicg_c(double *cp, double *cw, const double *cr, const double *cq, double cbkr, double cbki, int n,
                         hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_c_kernel, dim3(KGRID), dim3(KB), 0, st, (Z *)cp, (Z *)cw, (const Z *)cr, (const Z *)cq,
                       cmk(cbkr, cbki), n);
    return hipGetLastError();
}
hipError_t launch_bicg_init(double *cr, double *cq, double *cw, double *cp, const double *pnp, const double *anp, int n,
                            double *partials, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(bicg_init_kernel, dim3(KGRID), dim3(KB), 0, st, (Z *)cr, (Z *)cq, (Z *)cw, (Z *)cp,
                       (const Z *)pnp, (const Z *)anp, n, (Z *)partials);
    return hipGetLastError();
}
hipError_t launch_zaxpby(double ar, double ai, const double *x, double br, double bi, double *y, int n, hipStream_t st)
{
    typedef cplx<double> Z;
    hipLaunchKernelGGL(zaxpby_kernel, dim3(KGRID), dim3(KB), 0, st, cmk(ar, ai), (const Z *)x, cmk(br, bi), (Z *)y, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Block Gram-Schmidt pieces of the GMRES inner solver: nvec vectors V_j at
// V + j*stride (complex), one vector w.
// multi_zdot: partial[b*nvec + j] = sum over the block's share of conj(V_j) w
// reduce_cols: out[j] = sum_b partial[b*nvec + j] (fixed order)
// multi_axpy: w -= sum_j h_j V_j (sign = -1) or w += ... (sign = +1)
// ---------------------------------------------------------------------------
#define MV_MAX 128
__global__ __launch_bounds__(KB) void multi_zdot_kernel(const cplx<double> *__restrict__ V, int64_t stride, int nvec,
                                                        const cplx<double> *__restrict__ w, int n,
                                                        cplx<double> *__restrict__ partials)
{
    __shared__ cplx<double> red[KB];
    for (int j = 0; j < nvec; j++) {
        cplx<double> s = cmk(0.0, 0.0);
        const cplx<double> *v = V + (int64_t)j * stride;
        for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) cfma(s, cconj(v[i]), w[i]);
        red[threadIdx.x] = s;
        __syncthreads();
        for (int h = KB / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + h]);
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(int64_t)j * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// out[j] = sum over the nblk block partials of vector j (stored j-major);
// one workgroup per vector, fixed reduction order
__global__ __launch_bounds__(KB) void reduce_cols_kernel(const cplx<double> *__restrict__ partials, int nblk,
                                                         cplx<double> *__restrict__ out)
{
    __shared__ cplx<double> red[KB];
    const int j = blockIdx.x;
    cplx<double> s = cmk(0.0, 0.0);
    for (int b = threadIdx.x; b < nblk; b += KB) s = cadd(s, partials[(int64_t)j * nblk + b]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = KB / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

__global__ void multi_axpy_kernel(const cplx<double> *__restrict__ V, int64_t stride, int nvec,
                                  const cplx<double> *__restrict__ h, double sign, cplx<double> *__restrict__ w, int n)
{
    __shared__ cplx<double> hs[MV_MAX];
    for (int j = threadIdx.x; j < nvec; j += blockDim.x) hs[j] = cscale(h[j], sign);
    __syncthreads();
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        cplx<double> a = w[i];
        for (int j = 0; j < nvec; j++) cfma(a, hs[j], V[(int64_t)j * stride + i]);
        w[i] = a;
    }
}

int multi_vec_max() { return MV_MAX; }

// out (device, nvec complex) = V^H w; partials needs KGRID*nvec complex
hipError_t launch_multi_zdot(const double *V, int64_t stride, int nvec, const double *w, int n, double *partials,
                             double *out, hipStream_t st)
{
    if (nvec <= 0) return hipSuccess;
    hipLaunchKernelGGL(multi_zdot_kernel, dim3(KGRID), dim3(KB), 0, st, reinterpret_cast<const cplx<double> *>(V),
                       stride / 2, nvec, reinterpret_cast<const cplx<double> *>(w), n,
                       reinterpret_cast<cplx<double> *>(partials));
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(nvec), dim3(KB), 0, st, reinterpret_cast<const cplx<double> *>(partials),
                       KGRID, reinterpret_cast<cplx<double> *>(out));
    return hipGetLastError();
}

hipError_t launch_multi_axpy(const double *V, int64_t stride, int nvec, const double *h, double sign, double *w, int n,
                             hipStream_t st)
{
    if (nvec <= 0) return hipSuccess;
    if (nvec > MV_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(KGRID), dim3(KB), 0, st, reinterpret_cast<const cplx<double> *>(V),
                       stride / 2, nvec, reinterpret_cast<const cplx<double> *>(h), sign,
                       reinterpret_cast<cplx<double> *>(w), n);
    return hipGetLastError();
}
