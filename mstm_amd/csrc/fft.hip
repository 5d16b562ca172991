// fft.hip -- FFT node-grid sphere interaction on MI355X.
//
// Replaces fft_external_to_external_expansion (fft_translation-5.f90:72-336)
// and the routines it calls:
//   node_selection                 :654-825   (host, same rules)
//   local_sphere_to_node_translation :482-652 (list translation kernels)
//   fft_translation_initialization :901-974   (node kernel, FP64 -> FP32)
//   fft_node_to_node_translation   :827-899   (Stockham FFT passes +
//                                               per-frequency contraction)
//   local_sphere_to_sphere_expansion :338-480 (near-field pair list)
// The con_tran m-flips of :202-217/306-321 are applied by the shared
// prep/post kernels of direct.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include <type_traits>

#include "common.hpp"
#include "fft.hpp"
#include "kernels.hpp"
#include "specfun.hpp"

#define LMAX_LIST MSTM_TAB_L  // largest order in a list translation (node order): the tables' limit

// ===========================================================================
// Generic directed translations over a slot list (runtime orders).
// Slot q uses the interleaved table of launch_list_tables (target order LT
// <- source order LS).  DIR = 0 applies it forward (source order LS ->
// target LT), DIR = 1 in reverse (sop = .true.: m-flip, axial transposed,
// opposite phase).  Source element (rhs r, mode p, coefficient e) is read at
// src + src_off[q] + r*rs + p*ms + e*es; the result for slot q is written
// contiguously to dst[q][r][p][NB(target)].
// One workgroup per slot (128 or 256 threads) serves its four columns (r, p); each
// stage is spread over the threads and kept in LDS (dynamic size, from the
// orders):
//   xp(n,k)  = e^{i k phi} a(k,n)                          n <= Ls
//   y_m(n)   = sum_k d^n_{m,k} xp(n,k)                      (rotation)
//   z_m(t)   = sum_s A_m(t,s) y_m(s)                        (axial)
//   W(t,k)   = e^{-i k phi} sum_m d^t_{m,k} z_m(t)          (back rotation)
// ===========================================================================
// NT threads per workgroup (128 or 256, by the size of the stages)
template <typename T, int DIR, int NT>
__global__ __launch_bounds__(NT) void list_translate_kernel(const T *__restrict__ tab, int nslots, int LT,
                                                                    int LS, const cplx<T> *__restrict__ src,
                                                                    const int64_t *__restrict__ src_off, int64_t rs,
                                                                    int64_t ms, int64_t es,
                                                                    const unsigned char *__restrict__ ident,
                                                                    cplx<T> *__restrict__ dst)
{
    constexpr int VW = 16 / sizeof(T);
    const int q = blockIdx.x, tid = threadIdx.x;
    if (q >= nslots) return;
    const int Ls = DIR ? LT : LS, Lt = DIR ? LS : LT;
    const int NBS = Ls * (Ls + 2), NBT = Lt * (Lt + 2);
    const int nmin = imn(LT, LS), nmax = imx(LT, LS), NM = 2 * nmin + 1;
    const int SL = slot_len(LT, LS, VW);
    const T *p0 = tab + (int64_t)(q / MSTM_SLOTS) * SL * MSTM_SLOTS + (q % MSTM_SLOTS) * VW;
    const cplx<T> *xs = src + src_off[q];
    cplx<T> *out = dst + (int64_t)q * 4 * NBT;
    auto srcv = [&](int c, int n, int k) -> cplx<T> {
        return xs[(c >> 1) * rs + (c & 1) * ms + (int64_t)amn(k, n, Ls) * es];
    };
    if (ident[q]) {  // zero translation of regular waves: truncated copy (mtransfer)
        for (int e = tid; e < 4 * NBT; e += NT) {
            const int c = e / NBT, w = e - c * NBT;
            int t = (int)sqrtf((float)(w + 1));
            while (t * (t + 2) < w + 1) t++;
            while ((t - 1) * (t + 1) >= w + 1) t--;
            const int k = w + 1 - t * (t + 1);
            out[(int64_t)c * NBT + amn(k, t, Lt)] = (t <= Ls) ? srcv(c, t, k) : cmk((T)0, (T)0);
        }
        return;
    }
    // section offsets of the slot's table (block b = 0..NM-1 in the order
    // 0,-1,1,...; NM: the phase), worked out once instead of by the
    // layout functions' loops in every thread and term
    __shared__ int boff[2 * LMAX_LIST + 2], baxo[2 * LMAX_LIST + 1], bapl[2 * LMAX_LIST + 1];
    if (tid == 0) {
        int o = 0;
        for (int b = 0; b < NM; b++) {
            const int m1 = imx(1, iab(mblock_m(b)));
            const int rl = pad_to((nmax + 1) * (nmax + 1) - m1 * m1, VW);  // rot_len
            const int al = pad_to(2 * (LT - m1 + 1) * (LS - m1 + 1), VW);  // ax_len
            boff[b] = o;
            baxo[b] = o + rl;
            bapl[b] = al;
            o += rl + 2 * al;
        }
        boff[NM] = o;
    }
    extern __shared__ __attribute__((aligned(16))) unsigned char lt_smem[];
    cplx<T> *xp = reinterpret_cast<cplx<T> *>(lt_smem);  // [c][nk(n,k)], NBS each
    cplx<T> *ys = xp + 4 * NBS;                          // [c][m+nmin][n], n = 0..Ls
    cplx<T> *zs = ys + 4 * NM * (Ls + 1);                // [c][m+nmin][t], t = 0..Lt
    cplx<T> *ph = zs + 4 * NM * (Lt + 1);                // e^{i k phi}, k = -nmax..nmax
    // the slot's table, copied from its interleaved chunks once by all
    // threads (the stages below read each value many times)
    const size_t stage_bytes = sizeof(cplx<T>) * ((size_t)4 * nmax * (nmax + 2) + (size_t)8 * NM * (nmax + 1) + (2 * nmax + 1));
    T *tl = reinterpret_cast<T *>(lt_smem + pad_to((int)stage_bytes, 16));  // as list_translate_lds lays it out
    {
        typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V;
        const V *gsrc = reinterpret_cast<const V *>(p0);
        V *d = reinterpret_cast<V *>(tl);
        for (int c = tid; c < SL / VW; c += NT) d[c] = gsrc[(int64_t)c * MSTM_SLOTS];
    }
    __syncthreads();
    auto E = [&](int e) -> T { return tl[e]; };
    const cplx<T> e1 = cmk(E(boff[NM]), DIR ? -E(boff[NM] + 1) : E(boff[NM] + 1));
    for (int k = tid - nmax; k <= nmax; k += NT) {
        cplx<T> r = cmk((T)1, (T)0), b = k < 0 ? cconj(e1) : e1;
        for (int u = 0; u < iab(k); u++) r = cmul(r, b);
        ph[k + nmax] = r;
    }
    __syncthreads();
    // phase-shifted source, m-flipped for the reverse direction
    for (int e = tid; e < 4 * NBS; e += NT) {
        const int c = e / NBS, w = e - c * NBS;
        int n = (int)sqrtf((float)(w + 1));
        while (n * (n + 2) < w + 1) n++;
        while ((n - 1) * (n + 1) >= w + 1) n--;
        const int k = w + 1 - n * (n + 1);
        cplx<T> v;
        if (!DIR) v = srcv(c, n, k);
        else {
            v = srcv(c, n, -k);
            if (k & 1) v = cmk(-v.re, -v.im);
        }
        xp[c * NBS + w] = cmul(ph[k + nmax], v);
    }
    __syncthreads();
    // rotation rows of block m start at sec_off(b); degree n at + sum_{q=m1}^{n-1} (2q+1) = n^2 - m1^2
    auto bofm = [](int m) { return m == 0 ? 0 : (m < 0 ? 2 * (-m) - 1 : 2 * m); };
    // forward rotation y_m(n), n = m1..Ls
    for (int e = tid; e < 4 * NM * Ls; e += NT) {
        const int c = e / (NM * Ls), rem = e - c * NM * Ls, mi = rem / Ls, n = rem - mi * Ls + 1;
        const int m = mi - nmin, m1 = imx(1, iab(m));
        if (n < m1) continue;
        const int e0 = boff[bofm(m)] + n * n - m1 * m1;
        cplx<T> acc = cmk((T)0, (T)0);
        for (int k = -n; k <= n; k++) crfma(acc, E(e0 + k + n), xp[c * NBS + nk(n, k)]);
        ys[(c * NM + mi) * (Ls + 1) + n] = acc;
    }
    __syncthreads();
    // axial z_m(t), t = m1..Lt (mode-specific matrix, transposed for DIR = 1)
    for (int e = tid; e < 4 * NM * Lt; e += NT) {
        const int c = e / (NM * Lt), rem = e - c * NM * Lt, mi = rem / Lt, t = rem - mi * Lt + 1;
        const int m = mi - nmin, m1 = imx(1, iab(m)), p = c & 1;
        if (t < m1) continue;
        const int ea = baxo[bofm(m)] + p * bapl[bofm(m)];
        const int NR = LT - m1 + 1;
        cplx<T> acc = cmk((T)0, (T)0);
        for (int s = m1; s <= Ls; s++) {
            const int ex = DIR ? ea + 2 * ((s - m1) + NR * (t - m1)) : ea + 2 * ((t - m1) + NR * (s - m1));
            cfma(acc, cmk(E(ex), E(ex + 1)), ys[(c * NM + mi) * (Ls + 1) + s]);
        }
        zs[(c * NM + mi) * (Lt + 1) + t] = acc;
    }
    __syncthreads();
    // back rotation with the output phase; m-flip back for the reverse direction
    for (int e = tid; e < 4 * NBT; e += NT) {
        const int c = e / NBT, w = e - c * NBT;
        int t = (int)sqrtf((float)(w + 1));
        while (t * (t + 2) < w + 1) t++;
        while ((t - 1) * (t + 1) >= w + 1) t--;
        const int k = w + 1 - t * (t + 1);
        const int kk = DIR ? -k : k;  // W(t, kk) lands at (k, t)
        cplx<T> acc = cmk((T)0, (T)0);
        const int mm = imn(t, nmin);
        for (int m = -mm; m <= mm; m++) {
            const int m1 = imx(1, iab(m));
            const int e0 = boff[bofm(m)] + t * t - m1 * m1;
            crfma(acc, E(e0 + kk + t), zs[(c * NM + m + nmin) * (Lt + 1) + t]);
        }
        cplx<T> v = cmul(ph[-kk + nmax], acc);
        if (DIR && (k & 1)) v = cmk(-v.re, -v.im);
        out[(int64_t)c * NBT + amn(k, t, Lt)] = v;
    }
}

// threads of a translation's workgroup: 256 when its largest stage (the
// rotations, 4 columns x (2 min(LT, LS) + 1) blocks x max(LT, LS) degrees)
// has more than 256 items, else 128 (small orders leave the rest idle)
static int list_translate_threads(int LT, int LS)
{
    return 4 * (2 * imn(LT, LS) + 1) * imx(LT, LS) > 256 ? 256 : 128;
}
template <typename T, int DIR>
static hipError_t launch_list_translate(dim3 g, size_t lds, hipStream_t st, const T *tab, int nslots, int LT, int LS,
                                        const cplx<T> *src, const int64_t *off, int64_t rs, int64_t ms, int64_t es,
                                        const unsigned char *ident, cplx<T> *dst)
{
    if (list_translate_threads(LT, LS) == 256)
        hipLaunchKernelGGL((list_translate_kernel<T, DIR, 256>), g, dim3(256), lds, st, tab, nslots, LT, LS, src, off,
                           rs, ms, es, ident, dst);
    else
        hipLaunchKernelGGL((list_translate_kernel<T, DIR, 128>), g, dim3(128), lds, st, tab, nslots, LT, LS, src, off,
                           rs, ms, es, ident, dst);
    return hipGetLastError();
}

template <typename T> static size_t list_translate_lds(int LT, int LS)
{
    const int nmin = imn(LT, LS), nmax = imx(LT, LS), NM = 2 * nmin + 1, L1 = imx(LT, LS) + 1;
    // stages, then the slot's table (slot_len reals, at a 16-byte boundary)
    const size_t stages = sizeof(cplx<T>) * (4 * (size_t)nmax * (nmax + 2) + 2 * 4 * (size_t)NM * L1 + (2 * nmax + 1));
    return pad_to((int)stages, 16) + sizeof(T) * (size_t)slot_len(LT, LS, 16 / (int)sizeof(T));
}

// anode[r][p][n][cell] = sum over spheres of the cell of asph[i][r][p][n]
template <typename T>
__global__ void node_reduce_kernel(const cplx<T> *__restrict__ asph, const int *__restrict__ cstart,
                                   const int *__restrict__ cmem, int ncells, int nblk, cplx<T> *__restrict__ anode)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)4 * nblk * ncells) return;
    int cell = (int)(q % ncells);
    int64_t t = q / ncells;  // (r*2 + p)*nblk + n
    int n = (int)(t % nblk), rp = (int)(t / nblk);
    cplx<T> s = cmk((T)0, (T)0);
    for (int j = cstart[cell]; j < cstart[cell + 1]; j++) s = cadd(s, asph[((int64_t)cmem[j] * 4 + rp) * nblk + n]);
    anode[q] = s;
}

// out (partial layout, one part) = far field + near-field pair terms
template <typename T>
__global__ void sphere_reduce_kernel(const cplx<T> *__restrict__ far, const cplx<T> *__restrict__ nf,
                                     const cplx<T> *__restrict__ nr, const int *__restrict__ sstart,
                                     const int *__restrict__ sent, int nsph, int npad, int NBP,
                                     cplx<T> *__restrict__ out)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nsph * 4 * NBP) return;
    int e = (int)(q % NBP);
    int64_t t = q / NBP;
    int rp = (int)(t % 4);
    int i = (int)(t / 4);
    cplx<T> s = far[((int64_t)i * 4 + rp) * NBP + e];
    for (int j = sstart[i]; j < sstart[i + 1]; j++) {
        int ent = sent[j], pq = ent >> 1;
        const cplx<T> *src = (ent & 1) ? nr : nf;
        s = cadd(s, src[((int64_t)pq * 4 + rp) * NBP + e]);
    }
    int r = rp >> 1, p = rp & 1;
    out[(((int64_t)r * npad + i) * 2 + p) * NBP + e] = s;
}

// ===========================================================================
// Stockham FFT passes along one axis of a batch of 3-D arrays.  Input and
// output arrays differ only in the length along `axis` (nin -> nout):
// zero padding on input (nin < n) and truncation on output (nout < n).
// Radix 2/3/4/5 stages; P pencils per workgroup staged in LDS.
// sign = -1 forward, +1 inverse (no scaling).
// ===========================================================================
struct FftPlan {
    int n, nstage, radix[16];
};

// one radix-R butterfly of a Stockham stage (R = 2..5), twiddled inputs;
// R is a compile-time constant so the work arrays stay in registers.
// tw[j] = exp(sign 2 pi i j / n): the stage twiddles exp(sign 2 pi i kk r /
// (Ns R)) are tw[kk r n / (Ns R)] (Ns R divides n) and the butterfly's own
// roots exp(sign 2 pi i kr / R) are tw[kr n / R].
template <typename T, int R>
__device__ __forceinline__ void stage_radix(const cplx<T> *__restrict__ A, cplx<T> *__restrict__ B,
                                            const cplx<T> *__restrict__ tw, int p, int jj, int n, int nr, int Ns)
{
    const int kk = jj % Ns;
    const int step = n / (Ns * R);
    cplx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        cplx<T> x = A[p * n + jj + r * nr];
        if (kk != 0 && r != 0) x = cmul(x, tw[kk * r * step]);
        v[r] = x;
    }
    cplx<T> root[R];
#pragma unroll
    for (int q = 0; q < R; q++) root[q] = tw[q * nr];
    const int idxD = (jj / Ns) * Ns * R + kk;
#pragma unroll
    for (int k = 0; k < R; k++) {
        cplx<T> s = cmk((T)0, (T)0);
#pragma unroll
        for (int r = 0; r < R; r++) cfma(s, root[(k * r) % R], v[r]);
        B[p * n + idxD + k * Ns] = s;
    }
}

template <typename T>
__global__ void fft_pass_kernel(const cplx<T> *__restrict__ in, cplx<T> *__restrict__ out, int3 din, int3 dout,
                                int axis, FftPlan plan, int nin, int nout, int sign, int P)
{
    extern __shared__ unsigned char smem[];
    cplx<T> *A = reinterpret_cast<cplx<T> *>(smem);
    const int n = plan.n;
    cplx<T> *B = A + P * n;
    cplx<T> *tw = B + P * n;  // n roots of unity, computed once per workgroup
    const int d[3] = {din.x, din.y, din.z};
    const int o[3] = {dout.x, dout.y, dout.z};
    // non-axis dimensions (fast, slow)
    const int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
    const int npen = d[a0] * d[a1];
    const int64_t vin = (int64_t)d[0] * d[1] * d[2], vout = (int64_t)o[0] * o[1] * o[2];
    const int b = blockIdx.y;
    const int pen0 = blockIdx.x * P;
    const int np = imn(P, npen - pen0);
    const int64_t sin_[3] = {1, d[0], (int64_t)d[0] * d[1]};
    const int64_t sout[3] = {1, o[0], (int64_t)o[0] * o[1]};
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int j = tid; j < n; j += bs) {
        T sn, cs;
        if (sizeof(T) == 8) {
            double sd, cd;
            sincospi(2.0 * sign * (double)j / (double)n, &sd, &cd);
            sn = (T)sd;
            cs = (T)cd;
        } else {
            float sf, cf;
            sincospif(2.0f * sign * (float)j / (float)n, &sf, &cf);
            sn = (T)sf;
            cs = (T)cf;
        }
        tw[j] = cmk(cs, sn);
    }
    // load
    for (int idx = tid; idx < P * n; idx += bs) {
        int p, j;
        if (axis == 0) {
            p = idx / n;
            j = idx % n;
        } else {
            j = idx / P;
            p = idx % P;
        }
        cplx<T> v = cmk((T)0, (T)0);
        if (p < np && j < nin) {
            int pen = pen0 + p;
            int c0 = pen % d[a0], c1 = pen / d[a0];
            int64_t off = (int64_t)c0 * sin_[a0] + (int64_t)c1 * sin_[a1] + (int64_t)j * sin_[axis];
            v = in[(int64_t)b * vin + off];
        }
        A[p * n + j] = v;
    }
    __syncthreads();
    int Ns = 1;
    for (int s = 0; s < plan.nstage; s++) {
        const int R = plan.radix[s];
        const int nr = n / R;
        for (int w = tid; w < P * nr; w += bs) {
            const int p = w / nr, jj = w % nr;
            switch (R) {
            case 2: stage_radix<T, 2>(A, B, tw, p, jj, n, nr, Ns); break;
            case 3: stage_radix<T, 3>(A, B, tw, p, jj, n, nr, Ns); break;
            case 4: stage_radix<T, 4>(A, B, tw, p, jj, n, nr, Ns); break;
            default: stage_radix<T, 5>(A, B, tw, p, jj, n, nr, Ns); break;
            }
        }
        __syncthreads();
        cplx<T> *t = A;
        A = B;
        B = t;
        Ns *= R;
    }
    // store
    for (int idx = tid; idx < P * nout; idx += bs) {
        int p, j;
        if (axis == 0) {
            p = idx / nout;
            j = idx % nout;
        } else {
            j = idx / P;
            p = idx % P;
        }
        if (p >= np) continue;
        int pen = pen0 + p;
        int c0 = pen % d[a0], c1 = pen / d[a0];
        int64_t off = (int64_t)c0 * sout[a0] + (int64_t)c1 * sout[a1] + (int64_t)j * sout[axis];
        out[(int64_t)b * vout + off] = A[p * n + j];
    }
}

static FftPlan make_plan(int n)
{
    FftPlan pl{};
    pl.n = n;
    int m = n;
    while (m % 4 == 0) {
        pl.radix[pl.nstage++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        pl.radix[pl.nstage++] = 2;
        m /= 2;
    }
    while (m % 3 == 0) {
        pl.radix[pl.nstage++] = 3;
        m /= 3;
    }
    while (m % 5 == 0) {
        pl.radix[pl.nstage++] = 5;
        m /= 5;
    }
    if (m != 1) pl.n = -1;
    return pl;
}

template <typename T>
static hipError_t fft_axis(const cplx<T> *in, cplx<T> *out, int batch, int3 din, int3 dout, int axis, int n, int nin,
                           int nout, int sign, hipStream_t st)
{
    FftPlan pl = make_plan(n);
    if (pl.n < 0) return hipErrorInvalidValue;
    // P pencils per workgroup: about two elements per thread (small
    // workgroups, many of them: a pass is short and latency-bound), two
    // buffers of P pencils and the n roots of unity within 64 KB of LDS
    const int P = std::max(1, std::min(512 / n, (65536 / (int)sizeof(cplx<T>) / n - 1) / 2));
    const size_t sh = (2 * (size_t)P + 1) * n * sizeof(cplx<T>);
    if (sh > 65536) return hipErrorInvalidValue;  // an axis of more than 1365 (FP64) or 2730 (FP32) points
    const int d[3] = {din.x, din.y, din.z};
    int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
    int npen = d[a0] * d[a1];
    dim3 g((npen + P - 1) / P, batch);
    hipLaunchKernelGGL(fft_pass_kernel<T>, g, dim3(256), sh, st, in, out, din, dout, axis, pl, nin, nout, sign, P);
    return hipGetLastError();
}

// ===========================================================================
// Per-frequency contraction g(k,l) = sum_n K(p,l,n,k) a(k,n) for both RHS
// (the pointwise products of fft_node_to_node_translation, :864-872).
// The node-kernel spectrum is FP32 (the reference stores complex(4)); the
// data precision is T.  One thread per (frequency, mode, block of CONTRACT_LC l).
// ===========================================================================
#define CONTRACT_LC 4  // output degrees l per thread: each source value read serves LC of them
// NL (<= CONTRACT_LC) output degrees of one thread: a compile-time count, so
// that every load of a step has no guard and all of them (and those of the
// next step) are in flight together
template <typename T, int NL>
__device__ __forceinline__ void contract_degrees(const cplx<float> *__restrict__ K, const cplx<T> *__restrict__ a,
                                                 cplx<T> *__restrict__ g, int M, int nblk, int k, int p, int l0)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    cplx<T> acc[2][NL];
#pragma unroll
    for (int l = 0; l < NL; l++) acc[0][l] = acc[1][l] = cmk((T)0, (T)0);
    const int64_t ls = (int64_t)nblk * M;  // K stride between degrees l
    const cplx<float> *kp = K + ((int64_t)p * nblk + l0) * ls + k;
    const cplx<T> *ap0 = a + (int64_t)p * nblk * M + k, *ap1 = a + (int64_t)(2 + p) * nblk * M + k;
#pragma unroll 2
    for (int n = 0; n < nblk; n++) {
        const cplx<T> a0 = ap0[(int64_t)n * M], a1 = ap1[(int64_t)n * M];
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const f2v kv = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(kp + l * ls + (int64_t)n * M));
            const cplx<T> kk = cmk((T)kv[0], (T)kv[1]);
            cfma(acc[0][l], kk, a0);
            cfma(acc[1][l], kk, a1);
        }
    }
#pragma unroll
    for (int l = 0; l < NL; l++) {
        g[((int64_t)(0 * 2 + p) * nblk + l0 + l) * M + k] = acc[0][l];
        g[((int64_t)(1 * 2 + p) * nblk + l0 + l) * M + k] = acc[1][l];
    }
}

template <typename T>
__global__ void contract_kernel(const cplx<float> *__restrict__ K, const cplx<T> *__restrict__ a,
                                cplx<T> *__restrict__ g, int M, int nblk)
{
    constexpr int LC = CONTRACT_LC;
    static_assert(LC == 4, "the degree-count switch below covers 1..4");
    // workgroup -> (frequency block x, mode p, degree block z): the Z degree
    // blocks of one (x, p) read the same source fields, so they are dealt to
    // one XCD (workgroups go to MI355X's 8 XCDs round robin) one after another
    // and share its L2; the grid holds 8 * ceil(2 X / 8) * Z workgroups
    const int X = (M + 127) / 128, Z = (nblk + LC - 1) / LC;
    const int wg = blockIdx.x, slot = wg >> 3;
    const int c = (wg & 7) + 8 * (slot / Z);  // (x, p) combination
    if (c >= 2 * X) return;
    const int k = (c >> 1) * blockDim.x + threadIdx.x;
    const int p = c & 1;
    const int l0 = (slot % Z) * LC;
    if (k >= M) return;
    switch (imn(LC, nblk - l0)) {  // uniform over the workgroup
    case 4: contract_degrees<T, 4>(K, a, g, M, nblk, k, p, l0); break;
    case 3: contract_degrees<T, 3>(K, a, g, M, nblk, k, p, l0); break;
    case 2: contract_degrees<T, 2>(K, a, g, M, nblk, k, p, l0); break;
    default: contract_degrees<T, 1>(K, a, g, M, nblk, k, p, l0); break;
    }
}

// ===========================================================================
// Node kernel H(d*offset) for every point of the 2N-periodic grid
// (fft_translation_initialization :923-959), dense matrices by the
// vector-coupling formula of gentranmatrix (mstm-modules-33.f90:1762-1888),
// FP64.  One workgroup per grid point; rows l in [l0, l1) only (chunks).
// Ks[((p*nl + (l-l0))*nblk + n)*M + g], complex FP64.
// ===========================================================================
__global__ void node_kernel_build(double *__restrict__ Ks, int3 N, double dcell, int LG, int nblk,
                                  const int *__restrict__ hole, int nhole, const double *__restrict__ tc,
                                  const int *__restrict__ kl_of, const int *__restrict__ mn_of, double rimed, int l0,
                                  int l1)
{
    const int M = 8 * N.x * N.y * N.z;
    const int g = blockIdx.x;
    const int i0 = g % (2 * N.x), i1 = (g / (2 * N.x)) % (2 * N.y), i2 = g / (4 * N.x * N.y);
    const int W = 2 * LG;
    __shared__ double ymn[(4 * LMAX_LIST + 1) * (2 * LMAX_LIST + 1)];
    __shared__ cplx<double> hn[2 * LMAX_LIST + 1], eph[4 * LMAX_LIST + 1];
    const int nl = l1 - l0;
    // source offsets per axis: (n, sign) pairs mapping to index i
    int cx[2][2], cy[2][2], cz[2][2], ncx = 0, ncy = 0, ncz = 0;
    auto srcs = [](int i, int Nn, int c[2][2], int &nc) {
        if (i < Nn) {
            c[0][0] = i;
            c[0][1] = 1;
            nc = 1;
        } else if (i == Nn) {
            c[0][0] = Nn;
            c[0][1] = 1;
            c[1][0] = Nn;
            c[1][1] = -1;
            nc = 2;
        } else {
            c[0][0] = 2 * Nn - i;
            c[0][1] = -1;
            nc = 1;
        }
    };
    srcs(i0, N.x, cx, ncx);
    srcs(i1, N.y, cy, ncy);
    srcs(i2, N.z, cz, ncz);
    const int nel = nl * nblk;
    // zero this grid point; each thread later owns elements e = tid + k*bs
    for (int e = threadIdx.x; e < 2 * nel; e += blockDim.x) {
        int p = e / nel, rr = e % nel;
        int64_t idx = ((int64_t)(p * nl) * nblk + rr) * M + g;
        Ks[2 * idx] = 0.0;
        Ks[2 * idx + 1] = 0.0;
    }
    for (int ax = 0; ax < ncx; ax++)
        for (int ay = 0; ay < ncy; ay++)
            for (int az = 0; az < ncz; az++) {
                int nx = cx[ax][0], ny = cy[ay][0], nz = cz[az][0];
                bool inhole = false;
                for (int h = 0; h < nhole; h++)
                    if (hole[3 * h] == nx && hole[3 * h + 1] == ny && hole[3 * h + 2] == nz) inhole = true;
                if (inhole) continue;  // uniform across the block
                double xp[3] = {cx[ax][1] * dcell * nx, cy[ay][1] * dcell * ny, cz[az][1] * dcell * nz};
                double r = sqrt(xp[0] * xp[0] + xp[1] * xp[1] + xp[2] * xp[2]);
                double ct = xp[2] / r;
                __syncthreads();
                if (threadIdx.x == 0) sph_hankel_up(W, r * rimed, hn);
                if (threadIdx.x == 1) {
                    cplx<double> e1 = (xp[0] == 0.0 && xp[1] == 0.0)
                                          ? cmk(1.0, 0.0)
                                          : cmk(xp[0] / sqrt(xp[0] * xp[0] + xp[1] * xp[1]),
                                                xp[1] / sqrt(xp[0] * xp[0] + xp[1] * xp[1]));
                    eph[W] = cmk(1.0, 0.0);
                    for (int m = 1; m <= W; m++) {
                        eph[W + m] = cmul(eph[W + m - 1], e1);
                        eph[W - m] = cconj(eph[W + m]);
                    }
                }
                // normalized Legendre ymn(m, n), one thread per m >= 0
                if (threadIdx.x >= 2 && threadIdx.x < 2 + W + 1) {
                    int m = threadIdx.x - 2;
                    double sbe = sqrt((1.0 + ct) * (1.0 - ct));
                    double bc = 1.0;
                    for (int q = 1; q <= m; q++) bc *= sqrt((double)(2 * q) * (double)(2 * q - 1)) / (double)q;
                    double smm = ((m & 1) ? -1.0 : 1.0) * pow(0.5 * sbe, m) * bc;
                    const int ld = 2 * W + 1;
                    for (int nn = 0; nn <= W; nn++) {
                        ymn[(m + W) + ld * nn] = 0.0;
                        ymn[(-m + W) + ld * nn] = 0.0;
                    }
                    double ym1 = 0.0, y0 = smm;
                    ymn[(m + W) + ld * m] = y0;
                    if (m < W) {
                        double y1 = sqrt((double)(2 * m + 1)) * ct * y0;
                        ymn[(m + W) + ld * (m + 1)] = y1;
                        ym1 = y0;
                        y0 = y1;
                        for (int nn = m + 1; nn <= W - 1; nn++) {
                            double y2 = (-sqrt((double)(nn - m)) * sqrt((double)(nn + m)) * ym1 +
                                         (double)(2 * nn + 1) * ct * y0) /
                                        (sqrt((double)(nn + 1 - m)) * sqrt((double)(nn + 1 + m)));
                            ymn[(m + W) + ld * (nn + 1)] = y2;
                            ym1 = y0;
                            y0 = y2;
                        }
                    }
                    if (m > 0)
                        for (int nn = m; nn <= W; nn++)
                            ymn[(-m + W) + ld * nn] = ((m & 1) ? -1.0 : 1.0) * ymn[(m + W) + ld * nn];
                }
                __syncthreads();
                const int ld = 2 * W + 1;
                for (int e = threadIdx.x; e < nel; e += blockDim.x) {
                    int lr = e / nblk, n = e % nblk;  // target row address (l0+lr), source address n
                    int kl = kl_of[l0 + lr], mn = mn_of[n];
                    int l = (int)floor(sqrt((double)(kl + 1))), k = kl + 1 - l * (l + 1);
                    int nn = (int)floor(sqrt((double)(mn + 1))), m = mn + 1 - nn * (nn + 1);
                    int v = m - k;
                    int wmax = nn + l, wmin = imx(iab(v), iab(nn - l));
                    cplx<double> a1 = cmk(0.0, 0.0), b1 = cmk(0.0, 0.0);
                    const double *t = tc + ((int64_t)kl * nblk + mn) * (W + 1);
                    for (int w = wmin; w <= wmax; w++) {
                        double f = ymn[(v + W) + ld * w] * t[w];
                        if (((wmax - w) & 1) == 0) crfma(a1, f, hn[w]);
                        else crfma(b1, f, hn[w]);
                    }
                    cplx<double> ev = eph[v + W];
                    cplx<double> m1 = cmul(ev, cmk(a1.re - b1.im, a1.im + b1.re));
                    cplx<double> m2 = cmul(ev, cmk(a1.re + b1.im, a1.im - b1.re));
                    int64_t i0s = ((int64_t)(0 * nl + lr) * nblk + n) * M + g;
                    int64_t i1s = ((int64_t)(1 * nl + lr) * nblk + n) * M + g;
                    double *K0 = Ks + 2 * i0s, *K1 = Ks + 2 * i1s;
                    K0[0] += m1.re;
                    K0[1] += m1.im;
                    K1[0] += m2.re;
                    K1[1] += m2.im;
                }
            }
}

// complex FP64 chunk -> FP32 spectrum, scaled by 1/M (the /ncells8 of :898)
__global__ void to_f32_scaled(const cplx<double> *__restrict__ in, cplx<float> *__restrict__ out, int64_t n,
                              double scale)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = cmk((float)(in[q].re * scale), (float)(in[q].im * scale));
}

// ===========================================================================
// host side
// ===========================================================================
struct FftOp {
    std::string err;
    FftGeometry g;
    int nsph = 0, Lp = 0, npad = 0;
    std::vector<double> pos, radius;
    int cd[3] = {0, 0, 0};
    double d = 0, cb[3] = {0, 0, 0};
    int LG = 0, nblk = 0, ncells = 0, M = 0, nnb = 0;
    int nbr[27][3];
    std::vector<int> snode;
    std::vector<int> pairs;
    int npairs = 0;
    // device structures
    int *d_cstart = nullptr, *d_cmem = nullptr, *d_sstart = nullptr, *d_sent = nullptr;
    int64_t *d_off_sph = nullptr, *d_off_node = nullptr, *d_off_pj = nullptr, *d_off_pi = nullptr;
    unsigned char *d_ident_j = nullptr, *d_ident_h = nullptr;
    double *d_rt_j = nullptr, *d_rt_h = nullptr;
    void *jtab[2] = {nullptr, nullptr}, *htab[2] = {nullptr, nullptr};
    bool ready[2] = {false, false};
    cplx<float> *K = nullptr;
    // work buffers per precision
    void *xin[2] = {}, *asph[2] = {}, *bufA[2] = {}, *bufB[2] = {}, *far[2] = {}, *nf[2] = {}, *nr[2] = {},
         *out[2] = {};
    double setup_ms = 0, conv_ms = 0;
    long launches = 0;
    // HIP events around the spectral contraction, per precision (the stage
    // whose HBM rate the benchmark reports)
    std::vector<hipEvent_t> cev;
    std::vector<int> cev_prec;
    size_t cev_used = 0;
    double contract_ms[2] = {0, 0};
    long contract_launches[2] = {0, 0};
};

static void collect_contract(FftOp *f)
{
    for (size_t i = 0; i + 1 < f->cev_used; i += 2) {
        float ms = 0.f;
        if (hipEventSynchronize(f->cev[i + 1]) == hipSuccess &&
            hipEventElapsedTime(&ms, f->cev[i], f->cev[i + 1]) == hipSuccess) {
            f->contract_ms[f->cev_prec[i / 2]] += ms;
            f->contract_launches[f->cev_prec[i / 2]]++;
        }
    }
    f->cev_used = 0;
}

static int ferr(FftOp *f, int code, const char *msg)
{
    f->err = msg;
    return code;
}
#define FCHK(call)                                                                                                     \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) {                                                                                        \
            f->err = std::string(#call) + ": " + hipGetErrorString(e_);                                                \
            return 2;                                                                                                  \
        }                                                                                                              \
    } while (0)

template <typename P> static hipError_t falloc(P **p, size_t bytes)
{
    *p = nullptr;
    if (!bytes) return hipSuccess;
    return hipMalloc((void **)p, bytes);
}
template <typename P> static void ffree(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static bool is235(int n)
{
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

FftOp *fft_create(const FftGeometry &g, double *fva, const double *target_min, const double *target_max,
                  int d_specified, int node_order, int neighbor_model, int *cell_dim, double *d_cell,
                  int *node_order_out, int *err, const char **why)
{
    FftOp *f = new FftOp();
    f->g = g;
    f->nsph = g.nsph;
    f->Lp = g.Lmax;
    f->npad = (g.nsph + MSTM_SLOTS - 1) / MSTM_SLOTS * MSTM_SLOTS;
    f->pos.assign(g.pos, g.pos + 3 * g.nsph);
    f->radius.assign(g.radius, g.radius + g.nsph);
    const int n = g.nsph;
    // node_selection (fft_translation-5.f90:654-825).  Target box: the
    // caller's, else the bounds of the sphere centres (sphere_min_position /
    // sphere_max_position, mstm-modules-33.f90:5237-5244)
    double tmin[3] = {1e10, 1e10, 1e10}, tmax[3] = {-1e10, -1e10, -1e10};
    for (int i = 0; i < n; i++)
        for (int m = 0; m < 3; m++) {
            tmin[m] = std::min(tmin[m], f->pos[3 * i + m]);
            tmax[m] = std::max(tmax[m], f->pos[3 * i + m]);
        }
    for (int m = 0; m < 3; m++) {
        if (target_min) tmin[m] = target_min[m];
        if (target_max) tmax[m] = target_max[m];
    }
    double amean = 0, svol = 0;
    for (int i = 0; i < n; i++) {
        amean += f->radius[i];
        svol += f->radius[i] * f->radius[i] * f->radius[i];
    }
    svol = pow(svol, 1.0 / 3.0);
    amean /= n;
    double tvol = 1.0;
    for (int m = 0; m < 3; m++) tvol *= std::max(tmax[m] - tmin[m], amean);
    const double pi4 = (double)3.14159265358979323846f;  // the reference's real(4) pi
    double d = *d_cell;  // d_specified: the caller's cell size (:708-719 skipped)
    if (!d_specified) {
        double fv = *fva;
        if (fv <= 0.0) fv = std::max(std::min(4.0 * pi4 / 3.0 * svol * svol * svol / tvol, 1.0), 0.02);
        *fva = fv;  // fva is intent(inout) in the reference
        d = pow(4.0 * pi4 / 3.0 / fv / (double)n, 1.0 / 3.0) * svol;
    }
    if (!(d > 0.0)) {
        *err = 1;
        *why = "node_selection: cell size must be positive";
        delete f;
        return nullptr;
    }
    for (int m = 0; m < 3; m++) {
        int nd = (int)ceil((tmax[m] - tmin[m]) / d);
        if (nd < 1) {  // correctn235 would not end on 0 (fft_translation-5.f90:1068-1076)
            *err = 1;
            *why = "node_selection: the target box has no extent along an axis";
            delete f;
            return nullptr;
        }
        while (!is235(nd)) nd++;
        f->cd[m] = nd;
    }
    d = 0.0;
    for (int m = 0; m < 3; m++) d = std::max(d, (tmax[m] - tmin[m]) / (double)f->cd[m]);
    f->d = d;
    for (int m = 0; m < 3; m++) f->cb[m] = tmin[m];
    f->LG = (node_order <= 0) ? -node_order + (int)ceil(d) : node_order;
    if (f->LG > LMAX_LIST || f->Lp > LMAX_LIST) {
        *err = 4;
        *why = "node_selection: node order above the compiled limit (12)";
        delete f;
        return nullptr;
    }
    f->nblk = f->LG * (f->LG + 2);
    f->ncells = f->cd[0] * f->cd[1] * f->cd[2];
    f->M = 8 * f->ncells;
    f->snode.resize(n);
    std::vector<int> sn3(3 * n);
    for (int i = 0; i < n; i++) {
        for (int m = 0; m < 3; m++) {
            int nd = (int)floor((f->pos[3 * i + m] - tmin[m]) / d) + 1;
            nd = std::max(1, std::min(nd, f->cd[m]));
            sn3[3 * i + m] = nd;
        }
        f->snode[i] = (sn3[3 * i] - 1) + f->cd[0] * ((sn3[3 * i + 1] - 1) + f->cd[1] * (sn3[3 * i + 2] - 1));
    }
    int nn = 0;
    for (int iz = -1; iz <= 1; iz++)
        for (int iy = -1; iy <= 1; iy++)
            for (int ix = -1; ix <= 1; ix++) {
                int ir = ix * ix + iy * iy + iz * iz;
                if ((neighbor_model == 0 && ir > 0) || (neighbor_model == 1 && ir > 1) || (neighbor_model == 2 && ir > 2))
                    continue;
                f->nbr[nn][0] = ix;
                f->nbr[nn][1] = iy;
                f->nbr[nn][2] = iz;
                nn++;
            }
    f->nnb = nn;
    // cell membership (insertion order = sphere index)
    std::vector<std::vector<int>> members(f->ncells);
    for (int i = 0; i < n; i++) members[f->snode[i]].push_back(i);
    // near-field pairs (:794-823): for i, stencil cells in order, members j > i
    for (int i = 0; i < n; i++)
        for (int q = 0; q < nn; q++) {
            int c[3];
            bool ok = true;
            for (int m = 0; m < 3; m++) {
                c[m] = sn3[3 * i + m] - f->nbr[q][m];
                if (c[m] < 1 || c[m] > f->cd[m]) ok = false;
            }
            if (!ok) continue;
            int cell = (c[0] - 1) + f->cd[0] * ((c[1] - 1) + f->cd[1] * (c[2] - 1));
            for (int j : members[cell])
                if (j > i) {
                    f->pairs.push_back(i);
                    f->pairs.push_back(j);
                }
        }
    f->npairs = (int)f->pairs.size() / 2;
    cell_dim[0] = f->cd[0];
    cell_dim[1] = f->cd[1];
    cell_dim[2] = f->cd[2];
    *d_cell = f->d;
    *node_order_out = f->LG;
    *err = 0;
    return f;
}

void fft_grid(const FftOp *f, double *cell_boundary, int *sphere_node, int *number_neighbor_nodes, int *npairs)
{
    if (cell_boundary)
        for (int m = 0; m < 3; m++) cell_boundary[m] = f->cb[m];
    if (sphere_node)
        for (int i = 0; i < f->nsph; i++) {
            const int c = f->snode[i];
            sphere_node[3 * i] = c % f->cd[0] + 1;
            sphere_node[3 * i + 1] = (c / f->cd[0]) % f->cd[1] + 1;
            sphere_node[3 * i + 2] = c / (f->cd[0] * f->cd[1]) + 1;
        }
    if (number_neighbor_nodes) *number_neighbor_nodes = f->nnb - 1;
    if (npairs) *npairs = f->npairs;
}

void fft_bind(FftOp *f, const FftGeometry &g)
{
    if (f) f->g = g;
}

int fft_required_vcc(const FftOp *f) { return f ? std::max(f->LG, f->Lp) : 0; }
bool fft_ready(const FftOp *f, int prec) { return f && f->ready[prec]; }
const char *fft_last_error(const FftOp *f) { return f ? f->err.c_str() : "no FFT operator"; }
double fft_stat(FftOp *f, int which)
{
    if (!f) return 0.0;
    if (which >= 3) collect_contract(f);
    switch (which) {
    case 0: return f->setup_ms;
    case 1: return f->conv_ms;
    case 2: return (double)f->launches;
    case 3: return f->contract_ms[1];
    case 4: return (double)f->contract_launches[1];
    case 5: return f->contract_ms[0];
    case 6: return (double)f->contract_launches[0];
    default: return 0.0;
    }
}

void fft_reset_stats(FftOp *f)
{
    if (!f) return;
    collect_contract(f);
    f->contract_ms[0] = f->contract_ms[1] = 0.0;
    f->contract_launches[0] = f->contract_launches[1] = 0;
    f->launches = 0;
}

static void free_prec(FftOp *f, int p)
{
    ffree(f->jtab[p]);
    ffree(f->htab[p]);
    ffree(f->xin[p]);
    ffree(f->asph[p]);
    ffree(f->bufA[p]);
    ffree(f->bufB[p]);
    ffree(f->far[p]);
    ffree(f->nf[p]);
    ffree(f->nr[p]);
    ffree(f->out[p]);
    f->ready[p] = false;
}

void fft_clear(FftOp *f)
{
    if (!f) return;
    free_prec(f, 0);
    free_prec(f, 1);
    ffree(f->K);
}

void fft_destroy(FftOp *f)
{
    if (!f) return;
    collect_contract(f);
    for (auto &e : f->cev) hipEventDestroy(e);
    fft_clear(f);
    ffree(f->d_cstart);
    ffree(f->d_cmem);
    ffree(f->d_sstart);
    ffree(f->d_sent);
    ffree(f->d_off_sph);
    ffree(f->d_off_node);
    ffree(f->d_off_pj);
    ffree(f->d_off_pi);
    ffree(f->d_ident_j);
    ffree(f->d_ident_h);
    ffree(f->d_rt_j);
    ffree(f->d_rt_h);
    delete f;
}

// host Clebsch-Gordan (Racah's sum, long double) for the dense node kernel
static long double lf(int n)
{
    static std::vector<long double> t;
    if (t.empty()) {
        t.resize(200);
        t[0] = 1;
        for (int k = 1; k < 200; k++) t[k] = t[k - 1] * k;
    }
    return t[n];
}
static long double cg(int j1, int m1, int j2, int m2, int J)
{
    int M = m1 + m2;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J || J < std::abs(j1 - j2) || J > j1 + j2) return 0;
    long double pre = (2.0L * J + 1) * lf(J + j1 - j2) * lf(J - j1 + j2) * lf(j1 + j2 - J) / lf(j1 + j2 + J + 1) *
                      lf(J + M) * lf(J - M) * lf(j1 - m1) * lf(j1 + m1) * lf(j2 - m2) * lf(j2 + m2);
    long double s = 0;
    for (int k = std::max(0, std::max(j2 - J - m1, j1 - J + m2)); k <= std::min(j1 + j2 - J, std::min(j1 - m1, j2 + m2));
         k++)
        s += ((k & 1) ? -1.0L : 1.0L) /
             (lf(k) * lf(j1 + j2 - J - k) * lf(j1 - m1 - k) * lf(j2 + m2 - k) * lf(J - j2 + m1 + k) * lf(J - j1 - m2 + k));
    return sqrtl(pre) * s;
}

static int build_kernel(FftOp *f, hipStream_t st)
{
    if (f->K) return 0;
    const int LG = f->LG, nblk = f->nblk, W = 2 * LG, M = f->M;
    // tran_coef of gentrancoefconstants (mstm-modules-33.f90:1724-1760), indexed
    // by natural target kl and source mn: i^w c (-1)^k C(l,-k;n,m|w) C(l,-1;n,1|w),
    // c = -i^(l-n) sqrt((2l+1)(2n+1)); real part for even wmax-w, else imaginary
    std::vector<double> tc((size_t)nblk * nblk * (W + 1), 0.0);
    for (int l = 1; l <= LG; l++)
        for (int k = -l; k <= l; k++)
            for (int n = 1; n <= LG; n++)
                for (int m = -n; m <= n; m++) {
                    int kl = l * (l + 1) + k - 1, mn = n * (n + 1) + m - 1;
                    int wmax = n + l, wmin = std::max(std::abs(m - k), std::abs(n - l));
                    long double c = -sqrtl((long double)(2 * l + 1) * (2 * n + 1));
                    int ipow = ((l - n) % 4 + 4) % 4;  // c carries i^(l-n)
                    for (int w = wmin; w <= wmax; w++) {
                        long double v = c * ((k & 1) ? -1.0L : 1.0L) * cg(l, -k, n, m, w) * cg(l, -1, n, 1, w);
                        int pw = (ipow + w) % 4;  // v * i^pw
                        long double re = pw == 0 ? v : pw == 2 ? -v : 0.0L;
                        long double im = pw == 1 ? v : pw == 3 ? -v : 0.0L;
                        tc[((size_t)kl * nblk + mn) * (W + 1) + w] = (double)(((wmax - w) % 2 == 0) ? re : im);
                    }
                }
    // model-2 address -> natural index maps (K rows/cols use addresses)
    std::vector<int> kl_of(nblk), mn_of(nblk);
    for (int l = 1; l <= LG; l++)
        for (int k = -l; k <= l; k++) {
            kl_of[amn(k, l, LG)] = l * (l + 1) + k - 1;
            mn_of[amn(k, l, LG)] = l * (l + 1) + k - 1;
        }
    std::vector<int> hole;
    for (int q = 0; q < f->nnb; q++)
        if (f->nbr[q][0] >= 0 && f->nbr[q][1] >= 0 && f->nbr[q][2] >= 0) {
            hole.push_back(f->nbr[q][0]);
            hole.push_back(f->nbr[q][1]);
            hole.push_back(f->nbr[q][2]);
        }
    double *d_tc = nullptr;
    int *d_kl = nullptr, *d_mn = nullptr, *d_hole = nullptr;
    FCHK(falloc(&d_tc, sizeof(double) * tc.size()));
    FCHK(falloc(&d_kl, sizeof(int) * nblk));
    FCHK(falloc(&d_mn, sizeof(int) * nblk));
    FCHK(falloc(&d_hole, sizeof(int) * std::max<size_t>(3, hole.size())));
    FCHK(hipMemcpy(d_tc, tc.data(), sizeof(double) * tc.size(), hipMemcpyHostToDevice));
    FCHK(hipMemcpy(d_kl, kl_of.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(d_mn, mn_of.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
    if (!hole.empty()) FCHK(hipMemcpy(d_hole, hole.data(), sizeof(int) * hole.size(), hipMemcpyHostToDevice));
    FCHK(falloc(&f->K, sizeof(cplx<float>) * (size_t)2 * nblk * nblk * M));
    // chunk over target rows so the FP64 staging stays under ~2 GB
    size_t per_row = (size_t)2 * nblk * M * sizeof(cplx<double>);
    int rows = (int)std::max<size_t>(1, std::min<size_t>(nblk, ((size_t)2 << 30) / per_row));
    cplx<double> *ks = nullptr, *kt = nullptr;
    FCHK(falloc(&ks, per_row * rows));
    FCHK(falloc(&kt, per_row * rows));
    int3 Nd = make_int3(f->cd[0], f->cd[1], f->cd[2]);
    int3 D2 = make_int3(2 * f->cd[0], 2 * f->cd[1], 2 * f->cd[2]);
    for (int l0 = 0; l0 < nblk; l0 += rows) {
        int l1 = std::min(nblk, l0 + rows), nl = l1 - l0;
        hipLaunchKernelGGL(node_kernel_build, dim3(M), dim3(256), 0, st, (double *)ks, Nd, f->d, LG, nblk, d_hole,
                           (int)hole.size() / 3, d_tc, d_kl, d_mn, f->g.rimed, l0, l1);
        FCHK(hipGetLastError());
        int batch = 2 * nl * nblk;
        FCHK(fft_axis<double>(ks, kt, batch, D2, D2, 2, D2.z, D2.z, D2.z, -1, st));
        FCHK(fft_axis<double>(kt, ks, batch, D2, D2, 1, D2.y, D2.y, D2.y, -1, st));
        FCHK(fft_axis<double>(ks, kt, batch, D2, D2, 0, D2.x, D2.x, D2.x, -1, st));
        // kt[(p*nl + lr)*nblk + n][M] -> K[p][l0+lr][n][M]
        for (int p = 0; p < 2; p++) {
            int64_t cnt = (int64_t)nl * nblk * M;
            hipLaunchKernelGGL(to_f32_scaled, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                               kt + (int64_t)p * nl * nblk * M, f->K + ((int64_t)p * nblk + l0) * nblk * M, cnt,
                               1.0 / (double)M);
            FCHK(hipGetLastError());
        }
    }
    FCHK(hipStreamSynchronize(st));
    ffree(ks);
    ffree(kt);
    ffree(d_tc);
    ffree(d_kl);
    ffree(d_mn);
    ffree(d_hole);
    (void)Nd;
    return 0;
}

int fft_setup(FftOp *f, int prec, hipStream_t st)
{
    if (!f) return 3;
    // the sphere<->node and near-field tables need the vector-coupling table
    // up to the node order (the context binds it before setup)
    if (!f->g.d_vcc || f->g.vcc_LM < fft_required_vcc(f)) {
        f->err = "FFT set-up without a bound vector-coupling table of the node order";
        return 3;
    }
    auto t0 = std::chrono::steady_clock::now();
    const int n = f->nsph, Lp = f->Lp, LG = f->LG;
    const int NBP = Lp * (Lp + 2);
    if (!f->d_cstart) {
        // device lists: cells, near-field CSR, source offsets, translation vectors
        std::vector<int> cstart(f->ncells + 1, 0), cmem(n);
        for (int i = 0; i < n; i++) cstart[f->snode[i] + 1]++;
        for (int c = 0; c < f->ncells; c++) cstart[c + 1] += cstart[c];
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; i++) cmem[fill[f->snode[i]]++] = i;
        std::vector<std::vector<int>> ent(n);
        for (int q = 0; q < f->npairs; q++) {
            ent[f->pairs[2 * q]].push_back(2 * q);          // i gets forward term
            ent[f->pairs[2 * q + 1]].push_back(2 * q + 1);  // j gets reverse term
        }
        std::vector<int> sstart(n + 1, 0), sent;
        for (int i = 0; i < n; i++) {
            sstart[i + 1] = sstart[i] + (int)ent[i].size();
            for (int e : ent[i]) sent.push_back(e);
        }
        std::vector<int64_t> off_sph(n), off_node(n), off_pj(f->npairs), off_pi(f->npairs);
        std::vector<unsigned char> ij(n), ih(std::max(1, f->npairs), 0);
        std::vector<double> rtj(3 * (size_t)n), rth(3 * (size_t)std::max(1, f->npairs));
        for (int i = 0; i < n; i++) {
            off_sph[i] = (int64_t)i * 2 * NBP;
            off_node[i] = f->snode[i];
            int c = f->snode[i];
            int c3[3] = {c % f->cd[0], (c / f->cd[0]) % f->cd[1], c / (f->cd[0] * f->cd[1])};
            double r2 = 0;
            for (int m = 0; m < 3; m++) {
                rtj[3 * i + m] = f->d * ((double)(c3[m] + 1) - 0.5) - f->pos[3 * i + m] + f->cb[m];
                r2 += rtj[3 * i + m] * rtj[3 * i + m];
            }
            ij[i] = r2 < 1e-12;
        }
        for (int q = 0; q < f->npairs; q++) {
            int i = f->pairs[2 * q], j = f->pairs[2 * q + 1];
            off_pj[q] = (int64_t)j * 2 * NBP;
            off_pi[q] = (int64_t)i * 2 * NBP;
            for (int m = 0; m < 3; m++) rth[3 * q + m] = f->pos[3 * i + m] - f->pos[3 * j + m];
        }
        FCHK(falloc(&f->d_cstart, sizeof(int) * cstart.size()));
        FCHK(falloc(&f->d_cmem, sizeof(int) * std::max(1, n)));
        FCHK(falloc(&f->d_sstart, sizeof(int) * sstart.size()));
        FCHK(falloc(&f->d_sent, sizeof(int) * std::max<size_t>(1, sent.size())));
        FCHK(falloc(&f->d_off_sph, sizeof(int64_t) * n));
        FCHK(falloc(&f->d_off_node, sizeof(int64_t) * n));
        FCHK(falloc(&f->d_off_pj, sizeof(int64_t) * std::max(1, f->npairs)));
        FCHK(falloc(&f->d_off_pi, sizeof(int64_t) * std::max(1, f->npairs)));
        FCHK(falloc(&f->d_ident_j, n));
        FCHK(falloc(&f->d_ident_h, std::max(1, f->npairs)));
        FCHK(falloc(&f->d_rt_j, sizeof(double) * rtj.size()));
        FCHK(falloc(&f->d_rt_h, sizeof(double) * rth.size()));
        FCHK(hipMemcpy(f->d_cstart, cstart.data(), sizeof(int) * cstart.size(), hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_cmem, cmem.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_sstart, sstart.data(), sizeof(int) * sstart.size(), hipMemcpyHostToDevice));
        if (!sent.empty()) FCHK(hipMemcpy(f->d_sent, sent.data(), sizeof(int) * sent.size(), hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_off_sph, off_sph.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_off_node, off_node.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
        if (f->npairs) {
            FCHK(hipMemcpy(f->d_off_pj, off_pj.data(), sizeof(int64_t) * f->npairs, hipMemcpyHostToDevice));
            FCHK(hipMemcpy(f->d_off_pi, off_pi.data(), sizeof(int64_t) * f->npairs, hipMemcpyHostToDevice));
            FCHK(hipMemcpy(f->d_ident_h, ih.data(), f->npairs, hipMemcpyHostToDevice));
        }
        FCHK(hipMemcpy(f->d_ident_j, ij.data(), n, hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_rt_j, rtj.data(), sizeof(double) * rtj.size(), hipMemcpyHostToDevice));
        FCHK(hipMemcpy(f->d_rt_h, rth.data(), sizeof(double) * rth.size(), hipMemcpyHostToDevice));
    }
    int rc = build_kernel(f, st);
    if (rc) return rc;
    if (!f->ready[prec]) {
        const size_t esz = prec == 0 ? 8 : 4;
        const int VW = 16 / (int)esz;
        auto groups = [](int s) { return (size_t)((s + MSTM_SLOTS - 1) / MSTM_SLOTS) * MSTM_SLOTS; };
        size_t jb = groups(n) * slot_len(LG, Lp, VW) * esz;
        size_t hb = groups(std::max(1, f->npairs)) * slot_len(Lp, Lp, VW) * esz;
        FCHK(falloc(&f->jtab[prec], jb));
        FCHK(falloc(&f->htab[prec], hb));
        if (prec == 0) {
            FCHK(launch_list_tables<double>((double *)f->jtab[0], f->d_rt_j, n, LG, Lp, 1, f->g.rimed, f->g.d_vcc,
                                            f->g.vcc_LM, st));
            FCHK(launch_list_tables<double>((double *)f->htab[0], f->d_rt_h, f->npairs, Lp, Lp, 3, f->g.rimed,
                                            f->g.d_vcc, f->g.vcc_LM, st));
        } else {
            FCHK(launch_list_tables<float>((float *)f->jtab[1], f->d_rt_j, n, LG, Lp, 1, f->g.rimed, f->g.d_vcc,
                                           f->g.vcc_LM, st));
            FCHK(launch_list_tables<float>((float *)f->htab[1], f->d_rt_h, f->npairs, Lp, Lp, 3, f->g.rimed,
                                           f->g.d_vcc, f->g.vcc_LM, st));
        }
        const size_t cz = 2 * esz;  // complex element size
        FCHK(falloc(&f->xin[prec], cz * 2 * (size_t)f->npad * 2 * NBP));
        FCHK(falloc(&f->asph[prec], cz * (size_t)n * 4 * f->nblk));
        FCHK(falloc(&f->bufA[prec], cz * (size_t)4 * f->nblk * f->M));
        FCHK(falloc(&f->bufB[prec], cz * (size_t)4 * f->nblk * f->M));
        FCHK(falloc(&f->far[prec], cz * (size_t)n * 4 * NBP));
        FCHK(falloc(&f->nf[prec], cz * (size_t)std::max(1, f->npairs) * 4 * NBP));
        FCHK(falloc(&f->nr[prec], cz * (size_t)std::max(1, f->npairs) * 4 * NBP));
        FCHK(falloc(&f->out[prec], cz * (size_t)2 * f->npad * 2 * NBP));
        FCHK(hipMemsetAsync(f->out[prec], 0, cz * (size_t)2 * f->npad * 2 * NBP, st));
        f->ready[prec] = true;
    }
    FCHK(hipStreamSynchronize(st));
    f->setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

template <typename T> static int matvec_t(FftOp *f, int nrhs, const double *ain, double *aout, const int *flags,
                                          hipStream_t st)
{
    const int prec = sizeof(T) == 8 ? 0 : 1;
    const int n = f->nsph, Lp = f->Lp, LG = f->LG, nblk = f->nblk;
    const int NBP = Lp * (Lp + 2);
    const int3 N = make_int3(f->cd[0], f->cd[1], f->cd[2]);
    cplx<T> *xin = (cplx<T> *)f->xin[prec], *asph = (cplx<T> *)f->asph[prec], *A = (cplx<T> *)f->bufA[prec],
            *B = (cplx<T> *)f->bufB[prec], *far = (cplx<T> *)f->far[prec], *nf = (cplx<T> *)f->nf[prec],
            *nr = (cplx<T> *)f->nr[prec], *out = (cplx<T> *)f->out[prec];
    const T *jt = (const T *)f->jtab[prec], *ht = (const T *)f->htab[prec];
    FCHK(launch_prep<T>(ain, 2 * f->g.neqns, (T *)xin, nrhs, flags, n, f->npad, Lp, f->g.d_order, f->g.d_offset,
                        f->g.d_an, st));
    const int64_t rs_sph = (int64_t)f->npad * 2 * NBP;
    // sphere -> node (J translation, :573-626)
    dim3 lg(n);
    const size_t ljb = list_translate_lds<T>(LG, Lp), lhb = list_translate_lds<T>(Lp, Lp);
    FCHK((launch_list_translate<T, 0>)(lg, ljb, st, jt, n, LG, Lp, xin, f->d_off_sph, rs_sph, (int64_t)NBP, (int64_t)1,
                                     f->d_ident_j, asph));
    int64_t nn4 = (int64_t)4 * nblk * f->ncells;
    hipLaunchKernelGGL(node_reduce_kernel<T>, dim3((unsigned)((nn4 + 255) / 256)), dim3(256), 0, st, asph, f->d_cstart,
                       f->d_cmem, f->ncells, nblk, B);
    FCHK(hipGetLastError());
    // forward zero-padded 3-D FFT: z, y, x
    const int batch = 4 * nblk;
    int3 d0 = N, d1 = make_int3(N.x, N.y, 2 * N.z), d2 = make_int3(N.x, 2 * N.y, 2 * N.z),
         d3 = make_int3(2 * N.x, 2 * N.y, 2 * N.z);
    FCHK(fft_axis<T>(B, A, batch, d0, d1, 2, 2 * N.z, N.z, 2 * N.z, -1, st));
    FCHK(fft_axis<T>(A, B, batch, d1, d2, 1, 2 * N.y, N.y, 2 * N.y, -1, st));
    FCHK(fft_axis<T>(B, A, batch, d2, d3, 0, 2 * N.x, N.x, 2 * N.x, -1, st));
    // per-frequency contraction with the node kernel spectrum
    const int cX = (f->M + 127) / 128, cZ = (nblk + CONTRACT_LC - 1) / CONTRACT_LC;
    dim3 cg_((unsigned)((2 * cX + 7) / 8 * 8 * cZ));
    if (f->cev_used + 2 > 512) collect_contract(f);
    while (f->cev.size() < f->cev_used + 2) {
        hipEvent_t e;
        FCHK(hipEventCreate(&e));
        f->cev.push_back(e);
    }
    f->cev_prec.resize(f->cev.size() / 2 + 1);
    f->cev_prec[f->cev_used / 2] = prec;
    FCHK(hipEventRecord(f->cev[f->cev_used], st));
    hipLaunchKernelGGL(contract_kernel<T>, cg_, dim3(128), 0, st, f->K, A, B, f->M, nblk);
    FCHK(hipGetLastError());
    FCHK(hipEventRecord(f->cev[f->cev_used + 1], st));
    f->cev_used += 2;
    // inverse: x, y, z keeping the first N per axis
    FCHK(fft_axis<T>(B, A, batch, d3, d2, 0, 2 * N.x, 2 * N.x, N.x, 1, st));
    FCHK(fft_axis<T>(A, B, batch, d2, d1, 1, 2 * N.y, 2 * N.y, N.y, 1, st));
    FCHK(fft_axis<T>(B, A, batch, d1, d0, 2, 2 * N.z, 2 * N.z, N.z, 1, st));
    // node -> sphere (reverse J translation) from gnode[r][p][l][cell] in A
    const int64_t nc = f->ncells;
    FCHK((launch_list_translate<T, 1>)(lg, ljb, st, jt, n, LG, Lp, A, f->d_off_node, 2 * nblk * nc, nblk * nc, nc,
                                     f->d_ident_j, far));
    // near-field pairs, both directions (local_sphere_to_sphere_expansion)
    if (f->npairs) {
        dim3 pg(f->npairs);
        FCHK((launch_list_translate<T, 0>)(pg, lhb, st, ht, f->npairs, Lp, Lp, xin, f->d_off_pj, rs_sph, (int64_t)NBP,
                                         (int64_t)1, f->d_ident_h, nf));
        FCHK((launch_list_translate<T, 1>)(pg, lhb, st, ht, f->npairs, Lp, Lp, xin, f->d_off_pi, rs_sph, (int64_t)NBP,
                                         (int64_t)1, f->d_ident_h, nr));
    }
    int64_t ns4 = (int64_t)n * 4 * NBP;
    hipLaunchKernelGGL(sphere_reduce_kernel<T>, dim3((unsigned)((ns4 + 255) / 256)), dim3(256), 0, st, far, nf, nr,
                       f->d_sstart, f->d_sent, n, f->npad, NBP, out);
    FCHK(hipGetLastError());
    FCHK(launch_reduce_post<T>((const T *)out, 1, aout, 2 * f->g.neqns, nrhs, flags, n, f->npad, Lp, f->g.d_order,
                               f->g.d_offset, f->g.d_an, st));
    f->launches++;
    return 0;
}

int fft_matvec(FftOp *f, int nrhs, const double *ain, double *aout, const int *flags, int prec, hipStream_t st)
{
    if (!f) return 3;
    if (!f->ready[prec]) {  // the context sets the operator up (fft_setup after binding its tables)
        f->err = "FFT operator of this precision not set up";
        return 3;
    }
    return prec == 0 ? matvec_t<double>(f, nrhs, ain, aout, flags, st) : matvec_t<float>(f, nrhs, ain, aout, flags, st);
}
