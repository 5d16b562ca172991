// context.hip -- C-ABI entry points (include/mstm_hip.h) and the host-side
// orchestration of the device kernels: sphere state, operator set-up, the
// BiCG driver of cbicg (mstm-solver-8.f90:654-949), the mixed-precision
// refinement around it, and fixedorsoln (mstm-solver-8.f90:350-569).
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mstm_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "specfun.hpp"
#include "fft.hpp"

namespace {

struct DirectOp {
    bool ready = false;
    void *tab = nullptr;
    void *xin = nullptr;
    void *part = nullptr;
    size_t tab_bytes = 0;
};

struct Stats {
    double tables_ms[2] = {0, 0};
    double tile_ms[2] = {0, 0};
    long tile_launches[2] = {0, 0};
    double matvec_ms = 0;
    long matvecs[2] = {0, 0};
    double solve_ms = 0;
    long iters_fp64 = 0, iters_fp32 = 0, refinements = 0;
    double table_bytes[2] = {0, 0};
    double fft_setup_ms = 0, fft_conv_ms = 0;
    long fft_launches = 0;
};

}  // namespace

struct mstm_hip_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    int nsph = 0, npad = 0, Lmax = 0, neqns = 0;
    double rimed = 1.0;
    std::vector<int> order, offset;
    std::vector<double> pos, radius, ri;
    double *d_pos = nullptr, *d_radius = nullptr, *d_mre = nullptr, *d_mim = nullptr;
    int *d_order = nullptr, *d_offset = nullptr, *d_mieoff = nullptr;
    double *d_an = nullptr;  // complex, (npad*Lmax*4), sphere-uniform stride 4*Lmax
    int2 *d_tiles = nullptr;
    int ntiles = 0, nblocks = 0;
    double *d_vcc = nullptr;
    int vcc_LM = 0;
    DirectOp dop[2];
    FftOp *fft = nullptr;
    // Krylov work vectors (complex FP64, neqns each)
    double *d_kv = nullptr;
    size_t kv_n = 0;
    double *d_partials = nullptr;
    std::vector<double> h_partials;
    // event pool for tile-kernel timing
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_prec;
    size_t ev_used = 0;
    bool timing = true;
    Stats stats;
};

static int fail(mstm_hip_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}
#define HIPCHK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(c, MSTM_HIP_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));            \
    } while (0)

template <typename P> static hipError_t dalloc(P **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    return hipMalloc((void **)p, bytes);
}
template <typename P> static void dfree(P *&p)
{
    if (p) hipFree((void *)p);
    p = nullptr;
}

// ---------------------------------------------------------------- constants
// vector-coupling products vcc(n, l, m, w) of axialtrancoefinit
// (mstm-modules-33.f90:1680-1710): -sqrt((2n+1)(2l+1)) (-1)^m
// C(n,-m;l,m|w,0) C(n,-1;l,1|w,0) (Re+Im)(i^(n-l+w)); Clebsch-Gordan
// coefficients from Racah's sum in long double.
static long double lfac(int n)
{
    static std::vector<long double> f;
    if (f.empty()) {
        f.resize(400);
        f[0] = 1.0L;
        for (int k = 1; k < 400; k++) f[k] = f[k - 1] * (long double)k;
    }
    return f[n];
}
static long double clebsch(int j1, int m1, int j2, int m2, int J)
{
    int M = m1 + m2;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J || J < std::abs(j1 - j2) || J > j1 + j2) return 0;
    long double pre = (2.0L * J + 1.0L) * lfac(J + j1 - j2) * lfac(J - j1 + j2) * lfac(j1 + j2 - J) / lfac(j1 + j2 + J + 1);
    pre *= lfac(J + M) * lfac(J - M) * lfac(j1 - m1) * lfac(j1 + m1) * lfac(j2 - m2) * lfac(j2 + m2);
    int kmin = std::max(0, std::max(j2 - J - m1, j1 - J + m2)), kmax = std::min(j1 + j2 - J, std::min(j1 - m1, j2 + m2));
    long double s = 0;
    for (int k = kmin; k <= kmax; k++)
        s += ((k & 1) ? -1.0L : 1.0L) / (lfac(k) * lfac(j1 + j2 - J - k) * lfac(j1 - m1 - k) * lfac(j2 + m2 - k) *
                                          lfac(J - j2 + m1 + k) * lfac(J - j1 - m2 + k));
    return sqrtl(pre) * s;
}
static int ensure_vcc(mstm_hip_ctx *c, int LM)
{
    if (c->d_vcc && c->vcc_LM >= LM) return 0;
    dfree(c->d_vcc);
    size_t n = (size_t)(LM + 1) * (LM + 1) * (2 * LM + 1) * (2 * LM + 1);
    std::vector<double> h(n, 0.0);
    for (int nn = 1; nn <= LM; nn++)
        for (int l = 1; l <= LM; l++) {
            long double c1 = sqrtl((long double)(2 * nn + 1)) * sqrtl((long double)(2 * l + 1));
            for (int m = -std::min(nn, l); m <= std::min(nn, l); m++)
                for (int w = std::abs(nn - l); w <= nn + l; w++) {
                    long double c2 = -c1 * ((std::abs(m) & 1) ? -1.0L : 1.0L);
                    int k4 = (nn - l + w) & 3;
                    long double sg = (k4 == 0 || k4 == 1) ? 1.0L : -1.0L;
                    h[vcc_index(nn, l, m, w, LM)] =
                        (double)(c2 * clebsch(nn, -m, l, m, w) * clebsch(nn, -1, l, 1, w) * sg);
                }
        }
    HIPCHK(dalloc(&c->d_vcc, n * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_vcc, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    c->vcc_LM = LM;
    return 0;
}

static void release_direct(mstm_hip_ctx *c, int p)
{
    dfree(c->dop[p].tab);
    dfree(c->dop[p].xin);
    dfree(c->dop[p].part);
    c->dop[p].ready = false;
}

static void release_spheres(mstm_hip_ctx *c)
{
    release_direct(c, 0);
    release_direct(c, 1);
    if (c->fft) {
        fft_destroy(c->fft);
        c->fft = nullptr;
    }
    dfree(c->d_pos);
    dfree(c->d_radius);
    dfree(c->d_mre);
    dfree(c->d_mim);
    dfree(c->d_order);
    dfree(c->d_offset);
    dfree(c->d_mieoff);
    dfree(c->d_an);
    dfree(c->d_tiles);
    dfree(c->d_kv);
    c->kv_n = 0;
    c->nsph = 0;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

int mstm_hip_create(mstm_hip_ctx **ctx, int device)
{
    if (!ctx) return MSTM_HIP_EARG;
    *ctx = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MSTM_HIP_EHIP;
    mstm_hip_ctx *c = new mstm_hip_ctx();
    c->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(dalloc(&c->d_partials, sizeof(double) * 4 * 2 * krylov_grid()));
    c->h_partials.resize(4 * 2 * krylov_grid());
    *ctx = c;
    return 0;
}

int mstm_hip_destroy(mstm_hip_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->st);
    release_spheres(c);
    dfree(c->d_vcc);
    dfree(c->d_partials);
    for (auto &e : c->ev) hipEventDestroy(e);
    hipStreamDestroy(c->st);
    delete c;
    return 0;
}

const char *mstm_hip_last_error(const mstm_hip_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int finish_sphere_layout(mstm_hip_ctx *c)
{
    int n = c->nsph;
    c->offset.assign(n + 1, 0);
    c->Lmax = 0;
    for (int i = 0; i < n; i++) {
        c->offset[i + 1] = c->offset[i] + 2 * c->order[i] * (c->order[i] + 2);
        c->Lmax = std::max(c->Lmax, c->order[i]);
    }
    c->neqns = c->offset[n];
    if (c->Lmax > 8) return fail(c, MSTM_HIP_EUNSUP, "sphere order %d > 8 not compiled", c->Lmax);
    std::vector<int> mieoff(n);
    for (int i = 0; i < n; i++) mieoff[i] = 4 * c->Lmax * i;
    dfree(c->d_offset);
    dfree(c->d_mieoff);
    HIPCHK(dalloc(&c->d_offset, sizeof(int) * (n + 1)));
    HIPCHK(dalloc(&c->d_mieoff, sizeof(int) * n));
    HIPCHK(hipMemcpy(c->d_order, c->order.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_offset, c->offset.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_mieoff, mieoff.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    // direct-operator tiling over blocks of MSTM_SLOTS spheres
    c->npad = (n + MSTM_SLOTS - 1) / MSTM_SLOTS * MSTM_SLOTS;
    c->nblocks = c->npad / MSTM_SLOTS;
    std::vector<int2> tiles;
    for (int bi = 0; bi < c->nblocks; bi++)
        for (int bj = bi; bj < c->nblocks; bj++) tiles.push_back(make_int2(bi, bj));
    c->ntiles = (int)tiles.size();
    dfree(c->d_tiles);
    HIPCHK(dalloc(&c->d_tiles, sizeof(int2) * tiles.size()));
    HIPCHK(hipMemcpy(c->d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice));
    // Krylov work space: 15 vectors (slot plan in mstm_hip_fixedorsoln)
    c->kv_n = (size_t)c->neqns;
    dfree(c->d_kv);
    HIPCHK(dalloc(&c->d_kv, sizeof(double) * 2 * c->kv_n * 15));
    return 0;
}

int mstm_hip_set_spheres(mstm_hip_ctx *c, int n, const double *pos, const double *radius, const double *ri,
                         double ri_medium, double mie_eps)
{
    if (!c || n <= 0 || !pos || !radius || !ri) return fail(c, MSTM_HIP_EARG, "set_spheres: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    release_spheres(c);
    c->nsph = n;
    c->rimed = ri_medium;
    c->pos.assign(pos, pos + 3 * n);
    c->radius.assign(radius, radius + n);
    c->ri.assign(ri, ri + 2 * n);
    std::vector<double> mre(n), mim(n);
    for (int i = 0; i < n; i++) {
        mre[i] = ri[2 * i];
        mim[i] = ri[2 * i + 1];
    }
    HIPCHK(dalloc(&c->d_pos, sizeof(double) * 3 * n));
    HIPCHK(dalloc(&c->d_radius, sizeof(double) * n));
    HIPCHK(dalloc(&c->d_mre, sizeof(double) * n));
    HIPCHK(dalloc(&c->d_mim, sizeof(double) * n));
    HIPCHK(dalloc(&c->d_order, sizeof(int) * n));
    HIPCHK(hipMemcpy(c->d_pos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_radius, radius, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_mre, mre.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_mim, mim.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    // pass 1: orders
    HIPCHK(launch_mie(c->d_radius, c->d_mre, c->d_mim, ri_medium, n, mie_eps, c->d_order, nullptr, 0, c->st));
    c->order.resize(n);
    HIPCHK(hipMemcpyAsync(c->order.data(), c->d_order, sizeof(int) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    int rc = finish_sphere_layout(c);
    if (rc) return rc;
    // pass 2: Mie matrices at the selected orders, stride 4*Lmax per sphere
    HIPCHK(dalloc(&c->d_an, sizeof(double) * 2 * 4 * (size_t)c->Lmax * c->npad));
    HIPCHK(hipMemsetAsync(c->d_an, 0, sizeof(double) * 2 * 4 * (size_t)c->Lmax * c->npad, c->st));
    HIPCHK(launch_mie(c->d_radius, c->d_mre, c->d_mim, ri_medium, n, mie_eps, c->d_order, c->d_an, 4 * c->Lmax, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return 0;
}

int mstm_hip_set_mie(mstm_hip_ctx *c, const int *order, const double *an_mie)
{
    if (!c || !c->nsph || !order || !an_mie) return fail(c, MSTM_HIP_ESTATE, "set_mie: no spheres");
    int n = c->nsph;
    release_direct(c, 0);
    release_direct(c, 1);
    if (c->fft) {
        fft_destroy(c->fft);
        c->fft = nullptr;
    }
    c->order.assign(order, order + n);
    int rc = finish_sphere_layout(c);
    if (rc) return rc;
    std::vector<double> h((size_t)2 * 4 * c->Lmax * c->npad, 0.0);
    size_t src = 0;
    for (int i = 0; i < n; i++)
        for (int q = 0; q < 4 * c->order[i]; q++, src++) {
            h[2 * ((size_t)i * 4 * c->Lmax + q)] = an_mie[2 * src];
            h[2 * ((size_t)i * 4 * c->Lmax + q) + 1] = an_mie[2 * src + 1];
        }
    dfree(c->d_an);
    HIPCHK(dalloc(&c->d_an, sizeof(double) * h.size()));
    HIPCHK(hipMemcpy(c->d_an, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

int mstm_hip_get_orders(const mstm_hip_ctx *c, int *order, int *offset)
{
    if (!c || !c->nsph) return MSTM_HIP_ESTATE;
    if (order) memcpy(order, c->order.data(), sizeof(int) * c->nsph);
    if (offset) memcpy(offset, c->offset.data(), sizeof(int) * (c->nsph + 1));
    return 0;
}

int mstm_hip_number_eqns(const mstm_hip_ctx *c) { return c ? c->neqns : 0; }

int mstm_hip_get_mie(const mstm_hip_ctx *cc, double *an_mie)
{
    mstm_hip_ctx *c = const_cast<mstm_hip_ctx *>(cc);
    if (!c || !c->nsph || !an_mie) return MSTM_HIP_ESTATE;
    std::vector<double> h((size_t)2 * 4 * c->Lmax * c->npad);
    HIPCHK(hipMemcpy(h.data(), c->d_an, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    size_t dst = 0;
    for (int i = 0; i < c->nsph; i++)
        for (int q = 0; q < 4 * c->order[i]; q++, dst++) {
            an_mie[2 * dst] = h[2 * ((size_t)i * 4 * c->Lmax + q)];
            an_mie[2 * dst + 1] = h[2 * ((size_t)i * 4 * c->Lmax + q) + 1];
        }
    return 0;
}

int mstm_hip_setup_direct(mstm_hip_ctx *c, int prec)
{
    if (!c || !c->nsph) return fail(c, MSTM_HIP_ESTATE, "setup_direct: no spheres");
    if (prec != MSTM_HIP_FP64 && prec != MSTM_HIP_FP32) return fail(c, MSTM_HIP_EARG, "setup_direct: precision");
    HIPCHK(hipSetDevice(c->device));
    release_direct(c, prec);
    int rc = ensure_vcc(c, c->Lmax);
    if (rc) return rc;
    DirectOp &d = c->dop[prec];
    const int L = c->Lmax;
    const size_t esz = prec == MSTM_HIP_FP64 ? sizeof(double) : sizeof(float);
    const int VW = 16 / (int)esz;
    size_t slot = (size_t)slot_len(L, L, VW);
    d.tab_bytes = (size_t)c->ntiles * MSTM_SLOTS * MSTM_SLOTS * slot * esz;
    size_t nbp = (size_t)L * (L + 2);
    HIPCHK(dalloc(&d.tab, d.tab_bytes));
    HIPCHK(dalloc(&d.xin, esz * 2 * 2 * 2 * (size_t)c->npad * nbp));
    HIPCHK(dalloc(&d.part, esz * 2 * (size_t)c->nblocks * 2 * 2 * c->npad * nbp));
    auto t0 = std::chrono::steady_clock::now();
    if (prec == MSTM_HIP_FP64)
        HIPCHK(launch_direct_tables<double>((double *)d.tab, c->d_tiles, c->ntiles, L, c->nsph, c->d_pos, c->rimed,
                                            c->d_vcc, c->vcc_LM, c->st));
    else
        HIPCHK(launch_direct_tables<float>((float *)d.tab, c->d_tiles, c->ntiles, L, c->nsph, c->d_pos, c->rimed,
                                           c->d_vcc, c->vcc_LM, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->stats.tables_ms[prec] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->stats.table_bytes[prec] = (double)d.tab_bytes;
    d.ready = true;
    return 0;
}

int mstm_hip_node_selection(mstm_hip_ctx *c, double fva, int node_order, int neighbor_node_model, int *cell_dim,
                            double *d_cell, int *node_order_out)
{
    if (!c || !c->nsph) return fail(c, MSTM_HIP_ESTATE, "node_selection: no spheres");
    if (!cell_dim || !d_cell || !node_order_out) return fail(c, MSTM_HIP_EARG, "node_selection: null output");
    HIPCHK(hipSetDevice(c->device));
    if (c->fft) {
        fft_destroy(c->fft);
        c->fft = nullptr;
    }
    FftGeometry g{c->nsph,          c->Lmax,    c->neqns,   c->pos.data(), c->radius.data(), c->order.data(),
                  c->offset.data(), c->rimed,   c->d_pos,   c->d_order,    c->d_offset,      c->d_an,
                  c->d_vcc,         c->vcc_LM};
    int err = 0;
    c->fft = fft_create(g, fva, node_order, neighbor_node_model, cell_dim, d_cell, node_order_out, &err);
    if (!c->fft) return fail(c, err ? err : MSTM_HIP_EHIP, "node_selection failed");
    return 0;
}

int mstm_hip_setup_fft(mstm_hip_ctx *c, int prec)
{
    if (!c || !c->fft) return fail(c, MSTM_HIP_ESTATE, "setup_fft: call mstm_hip_node_selection first");
    HIPCHK(hipSetDevice(c->device));
    int rc = ensure_vcc(c, std::max(c->Lmax, fft_required_vcc(c->fft)));
    if (rc) return rc;
    FftGeometry g{c->nsph,          c->Lmax,    c->neqns,   c->pos.data(), c->radius.data(), c->order.data(),
                  c->offset.data(), c->rimed,   c->d_pos,   c->d_order,    c->d_offset,      c->d_an,
                  c->d_vcc,         c->vcc_LM};
    fft_bind(c->fft, g);
    rc = fft_setup(c->fft, prec, c->st);
    if (rc) return fail(c, rc, "setup_fft: %s", fft_last_error(c->fft));
    return 0;
}

}  // extern "C"

// --------------------------------------------------------------- operators
static hipEvent_t next_event(mstm_hip_ctx *c)
{
    if (c->ev_used >= c->ev.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        c->ev.push_back(e);
    }
    return c->ev[c->ev_used++];
}

// Device-side sphereinteraction on the direct operator: ain/aout are
// (neqns, nrhs) complex FP64 device arrays, nrhs <= 2.
static int direct_matvec(mstm_hip_ctx *c, int nrhs, const double *ain, double *aout, const int *flags, int prec)
{
    DirectOp &d = c->dop[prec];
    if (!d.ready) {
        int rc = mstm_hip_setup_direct(c, prec);
        if (rc) return rc;
    }
    const int L = c->Lmax;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        e0 = next_event(c);
        e1 = next_event(c);
    }
    if (prec == MSTM_HIP_FP64) {
        HIPCHK(launch_prep<double>(ain, 2 * c->neqns, (double *)d.xin, nrhs, flags, c->nsph, c->npad, L, c->d_order,
                                   c->d_offset, c->d_an, c->st));
        if (e0) hipEventRecord(e0, c->st);
        HIPCHK(launch_direct_tiles<double>((const double *)d.tab, c->d_tiles, c->ntiles, L, (const double *)d.xin,
                                           (double *)d.part, c->npad, c->st));
        if (e1) hipEventRecord(e1, c->st);
        HIPCHK(launch_reduce_post<double>((const double *)d.part, c->nblocks, aout, 2 * c->neqns, nrhs, flags, c->nsph,
                                          c->npad, L, c->d_order, c->d_offset, c->d_an, c->st));
    } else {
        HIPCHK(launch_prep<float>(ain, 2 * c->neqns, (float *)d.xin, nrhs, flags, c->nsph, c->npad, L, c->d_order,
                                  c->d_offset, c->d_an, c->st));
        if (e0) hipEventRecord(e0, c->st);
        HIPCHK(launch_direct_tiles<float>((const float *)d.tab, c->d_tiles, c->ntiles, L, (const float *)d.xin,
                                          (float *)d.part, c->npad, c->st));
        if (e1) hipEventRecord(e1, c->st);
        HIPCHK(launch_reduce_post<float>((const float *)d.part, c->nblocks, aout, 2 * c->neqns, nrhs, flags, c->nsph,
                                         c->npad, L, c->d_order, c->d_offset, c->d_an, c->st));
    }
    if (e0) c->ev_prec.push_back(prec);
    c->stats.matvecs[prec]++;
    return 0;
}

static int matvec(mstm_hip_ctx *c, int nrhs, const double *ain, double *aout, const int *flags, int prec, int use_fft)
{
    if (use_fft) {
        if (!c->fft) return fail(c, MSTM_HIP_ESTATE, "FFT operator not set up (call node_selection + setup_fft)");
        int rc = fft_matvec(c->fft, nrhs, ain, aout, flags, prec, c->st);
        if (rc) return fail(c, rc, "fft matvec: %s", fft_last_error(c->fft));
        c->stats.matvecs[prec]++;
        return 0;
    }
    return direct_matvec(c, nrhs, ain, aout, flags, prec);
}

static void collect_events(mstm_hip_ctx *c)
{
    hipStreamSynchronize(c->st);
    for (size_t q = 0; q + 1 < c->ev_used && q / 2 < c->ev_prec.size(); q += 2) {
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]);
        int p = c->ev_prec[q / 2];
        c->stats.tile_ms[p] += ms;
        c->stats.tile_launches[p]++;
    }
    c->ev_used = 0;
    c->ev_prec.clear();
}

// sum of per-block complex partials written by the Krylov kernels
static int sum_partials(mstm_hip_ctx *c, int nvals, double *out)
{
    int g = krylov_grid();
    HIPCHK(hipMemcpyAsync(c->h_partials.data(), c->d_partials, sizeof(double) * 2 * nvals * g, hipMemcpyDeviceToHost,
                          c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int v = 0; v < nvals; v++) {
        double re = 0, im = 0;
        for (int b = 0; b < g; b++) {
            re += c->h_partials[2 * (nvals * b + v)];
            im += c->h_partials[2 * (nvals * b + v) + 1];
        }
        out[2 * v] = re;
        out[2 * v + 1] = im;
    }
    return 0;
}

static inline void zdiv(double ar, double ai, double br, double bi, double &cr, double &ci)
{
    double d = br * br + bi * bi;
    cr = (ar * br + ai * bi) / d;
    ci = (ai * br - ar * bi) / d;
}

// cbicg on device vectors: pnp (b), anp (x in/out); returns iterations as
// the Fortran do-variable (niter+1 if the loop ran out).  Work vectors come
// from the context pool starting at slot ws.
static int bicg_dev(mstm_hip_ctx *c, int niter, double eps, const double *pnp, double *anp, int prec, int use_fft,
                    int ws, int *iter_out, double *errmax)
{
    const int n = c->neqns;
    const size_t V = 2 * c->kv_n;
    double *cr = c->d_kv + V * ws, *cq = cr + V, *cpw = cq + V, *caps = cpw + 2 * V;
    double *cp = cpw, *cw = cpw + V, *cap = caps, *caw = caps + V;
    *errmax = 0.0;
    *iter_out = 0;
    double bb[2];
    HIPCHK(launch_zdot(pnp, pnp, n, 1, c->d_partials, nullptr, c->st));
    int rc = sum_partials(c, 1, bb);
    if (rc) return rc;
    if (bb[0] == 0.0) return 0;
    const double enorm = bb[0];
    int f1[2] = {2, 0};
    rc = matvec(c, 1, anp, cr, f1, prec, use_fft);
    if (rc) return rc;
    HIPCHK(launch_bicg_init(cr, cq, cw, cp, pnp, anp, n, c->d_partials, c->st));
    double csk[2];
    rc = sum_partials(c, 1, csk);
    if (rc) return rc;
    if (csk[0] == 0.0 && csk[1] == 0.0) return 0;
    int f2[2] = {2, 3};
    int iter;
    for (iter = 1; iter <= niter; iter++) {
        rc = matvec(c, 2, cpw, caps, f2, prec, use_fft);
        if (rc) return rc;
        HIPCHK(launch_bicg_a(cap, caw, cp, cw, n, c->d_partials, c->st));
        double den[2];
        rc = sum_partials(c, 1, den);
        if (rc) return rc;
        if (den[0] == 0.0 && den[1] == 0.0) break;
        double cak[2];
        zdiv(csk[0], csk[1], den[0], den[1], cak[0], cak[1]);
        HIPCHK(launch_bicg_b(anp, cr, cq, cp, cap, caw, cak[0], cak[1], n, c->d_partials, c->st));
        double d2[4];
        rc = sum_partials(c, 2, d2);
        if (rc) return rc;
        double eerr = d2[2] / enorm;
        *errmax = eerr;
        if (prec == MSTM_HIP_FP64) c->stats.iters_fp64++;
        else c->stats.iters_fp32++;
        if (eerr < eps || (csk[0] == 0.0 && csk[1] == 0.0)) break;
        double cbk[2];
        zdiv(d2[0], d2[1], csk[0], csk[1], cbk[0], cbk[1]);
        csk[0] = d2[0];
        csk[1] = d2[1];
        HIPCHK(launch_bicg_c(cp, cw, cr, cq, cbk[0], cbk[1], n, c->st));
    }
    *iter_out = iter;
    return 0;
}

// Mixed precision: x <- x + d with A d = r solved by FP32 BiCG, r = b - A x
// evaluated with the FP64 operator; stops on the reference criterion
// |r|^2/|b|^2 < eps measured in FP64.
static int refine_dev(mstm_hip_ctx *c, int niter, double eps, const double *pnp, double *anp, int use_fft, int *iters,
                      double *errmax)
{
    const int n = c->neqns;
    const size_t V = 2 * c->kv_n;
    double *y = c->d_kv + V * 6, *r = y + V, *d = r + V;  // slots 6..8; bicg_dev uses ws=0 (slots 0..5)
    double bb[2];
    HIPCHK(launch_zdot(pnp, pnp, n, 1, c->d_partials, nullptr, c->st));
    int rc = sum_partials(c, 1, bb);
    if (rc) return rc;
    *iters = 0;
    *errmax = 0;
    if (bb[0] == 0.0) return 0;
    int f1[2] = {2, 0};
    const int max_ref = 6;
    for (int ref = 0; ref <= max_ref; ref++) {
        rc = matvec(c, 1, anp, y, f1, MSTM_HIP_FP64, use_fft);  // y = T H x
        if (rc) return rc;
        // r = b - x + y
        HIPCHK(hipMemcpyAsync(r, pnp, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(launch_zaxpby(-1.0, 0.0, anp, 1.0, 0.0, r, n, c->st));
        HIPCHK(launch_zaxpby(1.0, 0.0, y, 1.0, 0.0, r, n, c->st));
        double rr[2];
        HIPCHK(launch_zdot(r, r, n, 1, c->d_partials, nullptr, c->st));
        rc = sum_partials(c, 1, rr);
        if (rc) return rc;
        double err = rr[0] / bb[0];
        *errmax = err;
        if (err < eps || ref == max_ref || *iters > niter) break;
        c->stats.refinements++;
        // inner FP32 solve of A d = r from d = r, to half the outer target:
        // the FP32 operator's relative error (~1e-7) is far below the 1e-3
        // relative residual asked for, so one refinement step normally ends
        HIPCHK(hipMemcpyAsync(d, r, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, c->st));
        double eps_in = std::max(0.5 * eps * bb[0] / rr[0], 1e-12);
        int it = 0;
        double ein = 0;
        rc = bicg_dev(c, niter - *iters, eps_in, r, d, MSTM_HIP_FP32, use_fft, 0, &it, &ein);
        if (rc) return rc;
        *iters += std::min(it, niter - *iters);
        HIPCHK(launch_zaxpby(1.0, 0.0, d, 1.0, 0.0, anp, n, c->st));
    }
    return 0;
}

extern "C" {

int mstm_hip_sphereinteraction(mstm_hip_ctx *c, int nrhs, const double *ain, double *aout, const int *con_tran,
                               const int *mie_mult, int use_fft, int prec)
{
    if (!c || !c->nsph) return fail(c, MSTM_HIP_ESTATE, "sphereinteraction: no spheres");
    if (nrhs < 1 || !ain || !aout || !con_tran || !mie_mult) return fail(c, MSTM_HIP_EARG, "sphereinteraction: args");
    HIPCHK(hipSetDevice(c->device));
    const size_t V = 2 * (size_t)c->neqns;
    double *din = c->d_kv + V * 13, *dout = c->d_kv;  // staging slots 13-14 in, 0-1 out
    for (int r0 = 0; r0 < nrhs; r0 += 2) {
        int nr = std::min(2, nrhs - r0);
        int flags[2] = {0, 0};
        for (int q = 0; q < nr; q++) flags[q] = (con_tran[r0 + q] ? 1 : 0) | (mie_mult[r0 + q] ? 2 : 0);
        HIPCHK(hipMemcpyAsync(din, ain + V * r0, sizeof(double) * V * nr, hipMemcpyHostToDevice, c->st));
        int rc = matvec(c, nr, din, dout, flags, prec, use_fft);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(aout + V * r0, dout, sizeof(double) * V * nr, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    collect_events(c);
    return 0;
}

int mstm_hip_cbicg(mstm_hip_ctx *c, int niter, double eps, const double *pnp, double *anp, int use_fft, int *iter,
                   double *errmax)
{
    if (!c || !c->nsph) return fail(c, MSTM_HIP_ESTATE, "cbicg: no spheres");
    if (niter <= 0) return fail(c, MSTM_HIP_EUNSUP, "cbicg: niter <= 0 (order-of-scattering) not supported");
    HIPCHK(hipSetDevice(c->device));
    const size_t V = 2 * (size_t)c->neqns;
    double *db = c->d_kv + V * 12, *dx = c->d_kv + V * 11;
    HIPCHK(hipMemcpyAsync(db, pnp, sizeof(double) * V, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dx, anp, sizeof(double) * V, hipMemcpyHostToDevice, c->st));
    int rc = bicg_dev(c, niter, eps, db, dx, MSTM_HIP_FP64, use_fft, 0, iter, errmax);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(anp, dx, sizeof(double) * V, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_events(c);
    return 0;
}

int mstm_hip_fixedorsoln(mstm_hip_ctx *c, double alpha, double sinc, int dir, double eps, int niter, int solver,
                         int use_fft, double *amnp, double *qeff, double *maxerr, int *maxiter, int *istat)
{
    if (!c || !c->nsph) return fail(c, MSTM_HIP_ESTATE, "fixedorsoln: no spheres");
    if (dir != 1) return fail(c, MSTM_HIP_EUNSUP, "fixedorsoln: only dir = 1 (no plane boundaries)");
    if (niter <= 0) return fail(c, MSTM_HIP_EUNSUP, "fixedorsoln: niter <= 0 not supported");
    if (solver != MSTM_HIP_SOLVE_PLAIN && solver != MSTM_HIP_SOLVE_MIXED)
        return fail(c, MSTM_HIP_EUNSUP, "fixedorsoln: solver %d not available", solver);
    if (!amnp) return fail(c, MSTM_HIP_EARG, "fixedorsoln: amnp is required");
    HIPCHK(hipSetDevice(c->device));
    const int n = c->neqns;
    const size_t V = 2 * (size_t)n;
    // work-vector slots: 0-5 BiCG, 6-8 refinement, 9-10 pmnp, 11 x, 12 b, 13-14 staging
    double *pmnp = c->d_kv + V * 9, *x = c->d_kv + V * 11, *b = c->d_kv + V * 12, *stage = c->d_kv + V * 13;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(launch_planewave(pmnp, 2 * n, c->nsph, c->d_pos, c->d_order, c->d_offset, alpha, sinc, c->rimed, c->st));
    for (int p = 0; p < 2; p++) {
        // istat, maxiter, maxerr are reset per polarization (:471-475)
        *istat = 0;
        *maxerr = 0.0;
        *maxiter = 0;
        HIPCHK(launch_mie_mult(pmnp + V * p, b, c->nsph, c->d_order, c->d_offset, c->d_mieoff, c->d_an, 1, 0, c->Lmax,
                               c->st));
        HIPCHK(hipMemcpyAsync(x, b, sizeof(double) * V, hipMemcpyDeviceToDevice, c->st));
        int it = 0, rc;
        double err = 0;
        if (solver == MSTM_HIP_SOLVE_PLAIN) rc = bicg_dev(c, niter, eps, b, x, MSTM_HIP_FP64, use_fft, 0, &it, &err);
        else rc = refine_dev(c, niter, eps, b, x, use_fft, &it, &err);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(stage + V * p, x, sizeof(double) * V, hipMemcpyDeviceToDevice, c->st));
        *maxiter = it;
        *maxerr = err;
        if (it > niter || err > eps) *istat = 1;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->stats.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    HIPCHK(hipMemcpyAsync(amnp, stage, sizeof(double) * 2 * V, hipMemcpyDeviceToHost, c->st));
    if (qeff) {
        double *dq = nullptr;
        HIPCHK(dalloc(&dq, sizeof(double) * 9 * c->nsph));
        HIPCHK(launch_qeff(stage, pmnp, 2 * n, c->nsph, c->d_radius, c->d_order, c->d_offset, c->d_mieoff, c->d_an,
                           c->rimed, dq, c->st));
        HIPCHK(hipMemcpyAsync(qeff, dq, sizeof(double) * 9 * c->nsph, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(dq);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    collect_events(c);
    return 0;
}

int mstm_hip_clear(mstm_hip_ctx *c, int clear_tables)
{
    if (!c) return MSTM_HIP_EARG;
    if (clear_tables) {
        release_direct(c, 0);
        release_direct(c, 1);
        if (c->fft) fft_clear(c->fft);
    }
    c->stats = Stats();
    return 0;
}

const char *mstm_hip_stats_names(void)
{
    return "tables_ms_fp64,tables_ms_fp32,tile_ms_fp64,tile_ms_fp32,tile_launches_fp64,tile_launches_fp32,"
           "matvecs_fp64,matvecs_fp32,solve_ms,iters_fp64,iters_fp32,refinements,table_bytes_fp64,table_bytes_fp32,"
           "fft_setup_ms,fft_conv_ms,fft_launches,npad,ntiles,nblocks";
}

int mstm_hip_stats(const mstm_hip_ctx *c, double *out, int nout)
{
    if (!c || !out) return 0;
    double v[] = {c->stats.tables_ms[0],
                  c->stats.tables_ms[1],
                  c->stats.tile_ms[0],
                  c->stats.tile_ms[1],
                  (double)c->stats.tile_launches[0],
                  (double)c->stats.tile_launches[1],
                  (double)c->stats.matvecs[0],
                  (double)c->stats.matvecs[1],
                  c->stats.solve_ms,
                  (double)c->stats.iters_fp64,
                  (double)c->stats.iters_fp32,
                  (double)c->stats.refinements,
                  c->stats.table_bytes[0],
                  c->stats.table_bytes[1],
                  c->fft ? fft_stat(c->fft, 0) : 0.0,
                  c->fft ? fft_stat(c->fft, 1) : 0.0,
                  c->fft ? fft_stat(c->fft, 2) : 0.0,
                  (double)c->npad,
                  (double)c->ntiles,
                  (double)c->nblocks};
    int k = std::min(nout, (int)(sizeof v / sizeof v[0]));
    for (int q = 0; q < k; q++) out[q] = v[q];
    return k;
}

}  // extern "C"
