// direct.hip -- direct (pair-by-pair) sphere interaction H.a on MI355X.
//
// Replaces the pair loop of external_to_external_expansion
// (mstm-modules-33.f90:6390-6525) and its per-pair coefficient_translation
// (rot path, :6697-6749).  Work decomposition:
//   * spheres are grouped in blocks of S = MSTM_SLOTS; a tile is an ordered
//     block pair (bI <= bJ); per column pair it has two waves;
//   * the first wave computes I <- J (forward, sop = .false.), the second
//     J <- I (reverse, sop = .true., which adds the m-flip of
//     shiftcoefficient and uses the axial block transposed) from the same
//     tables;
//   * lane = slot k (16) x column (rhs, mode) (4).  In step s lane k of the
//     forward wave targets I[k] and sources J[(k+s)%S]; lane k of the reverse
//     wave targets J[k] and sources I[(k-s)%S].  A lane's target never
//     changes inside a tile, so it accumulates in registers and writes one
//     partial sum per (target, partner block): no atomics, deterministic
//     order;
//   * pair tables are laid out in HBM in exactly this consumption order and
//     interleaved over the 16 slots, so each table load is 16 x 16 B;
//   * up to order 5 (FP32) / 4 (FP64) the tables use the symmetric layout
//     of common.hpp: only blocks m >= 0, half-sum/half-difference Wigner rows
//     and one axial block per mode serve the blocks m and -m, which halves
//     the table bytes and the rotation work; a step then runs in three
//     passes over one vector (forward rotation by degree, axial products by
//     block, back rotation and output phase by degree), so that only the
//     accumulators and that vector live across the passes;
//   * in that range all waves of the tile copy a step's table group (16
//     slots) into LDS once per step and read it there: every table byte
//     then leaves the L2 once per tile instead of once per lane that needs
//     it (4 lanes x 2 or 4 waves);
//   * up to order 4 the tile's two source blocks are copied into LDS once
//     per tile too, instead of every lane reloading its source sphere from
//     L2 in every step; FP32 tiles also load the next step's table group
//     into registers while the current step computes;
//   * FP64 four-column products run as two workgroups per tile, one per
//     column pair, on the same XCD; small clusters split a tile's 16 steps
//     over G workgroups (step groups), each with its own partial sums.
// The reduce kernel sums the partials and applies the Mie/con_tran post-op
// of sphereinteraction (mstm-scatprops-26.f90:158-168).
#include <utility>

#include "common.hpp"
#include "kernels.hpp"

template <typename T, int VW> struct vec_of;
template <> struct vec_of<float, 4> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct vec_of<double, 2> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename F, int... Is>
__device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// Load LEN elements of a slot section starting at element E0 (multiple of
// VW) into r[]; p0 = group base + slot*VW (+ a runtime chunk offset).
template <typename T, int E0, int LEN>
__device__ __forceinline__ void load_sec(const T *__restrict__ p0, T *r)
{
    constexpr int VW = 16 / sizeof(T);
    typedef typename vec_of<T, VW>::type V;
#pragma unroll
    for (int c = 0; c < (LEN + VW - 1) / VW; c++) {
        const V *p = reinterpret_cast<const V *>(p0 + (int64_t)(E0 / VW + c) * MSTM_SLOTS * VW);
#ifdef MSTM_TABLE_NT
        V v = __builtin_nontemporal_load(p);
#else
        // both waves of a tile read the same table group within a few
        // microseconds: plain loads let the second one hit in L2
        V v = *p;
#endif
#pragma unroll
        for (int q = 0; q < VW; q++)
            if (c * VW + q < LEN) r[c * VW + q] = v[q];
    }
}

// Load LEN elements of a slot starting at element E0 (any alignment): the
// 16-byte chunks that cover them, unused lanes of the end chunks dropped
template <typename T, int E0, int LEN>
__device__ __forceinline__ void load_span(const T *__restrict__ p0, T *r)
{
    constexpr int VW = 16 / sizeof(T);
    typedef typename vec_of<T, VW>::type V;
#pragma unroll
    for (int c = E0 / VW; c <= (E0 + LEN - 1) / VW; c++) {
        const V v = *reinterpret_cast<const V *>(p0 + (int64_t)c * MSTM_SLOTS * VW);
#pragma unroll
        for (int q = 0; q < VW; q++) {
            const int e = c * VW + q - E0;
            if (e >= 0 && e < LEN) r[e] = v[q];
        }
    }
}

// LDS staging of a step's table group (see the file header): exactly the
// orders with the symmetric table layout
template <typename T, int L> constexpr bool tile_staged() { return direct_sym(L, (int)sizeof(T)); }
// Tiles up to order 4 keep the tile's two source blocks in LDS (instead of
// every lane reloading its source sphere from L2 in every step).  FP32 tiles
// also software-pipeline the staging: step s+1's group is loaded into
// registers while step s computes (FP64 has no registers to spare for it).
// NR here is the number of columns of one workgroup.
template <typename T, int L, int NR> constexpr bool tile_src_staged() { return L <= 4; }
template <typename T, int L, int NR> constexpr bool tile_pipelined()
{
    return sizeof(T) == 4 && tile_src_staged<T, L, NR>();
}
// occupancy target handed to the compiler: tiles with staged sources are
// scheduled as if alone on their SIMD (the FP32 prefetch then overlaps the
// step, and FP64's 80 KB of LDS per workgroup allows one wave per SIMD
// anyway; FP32 still fits two), other tiles up to order 4 for two waves per
// SIMD, larger orders the full register file
template <typename T, int L, int NR> constexpr int tile_waves()
{
    return tile_src_staged<T, L, NR>() ? 1 : (L <= 4 ? 2 : 1);
}
// register prefetch of a table group for the pipelined form
template <typename T, int L, int NR> struct GroupPrefetch {
    static constexpr int VW = 16 / sizeof(T);
    static constexpr int NV = direct_slot_len(L, sizeof(T)) * MSTM_SLOTS / VW;
    static constexpr int NT = 64 * NR;
    static constexpr int PF = tile_pipelined<T, L, NR>() ? (NV + NT - 1) / NT : 1;
    typedef typename vec_of<T, VW>::type V;
    V r[PF];
    // each group is read once, by its one workgroup: streaming loads keep it
    // from evicting the source vectors from L2 (1.5% fewer fetched bytes)
    __device__ __forceinline__ void load(const T *__restrict__ tab, int tile, int s)
    {
        const V *g = reinterpret_cast<const V *>(tab + ((int64_t)tile * MSTM_SLOTS + s) *
                                                           (int64_t)direct_slot_len(L, sizeof(T)) * MSTM_SLOTS);
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int i = threadIdx.x + q * NT;
            if (i < NV) r[q] = __builtin_nontemporal_load(&g[i]);
        }
    }
    __device__ __forceinline__ void store(T *__restrict__ grp) const
    {
        V *d = reinterpret_cast<V *>(grp);
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int i = threadIdx.x + q * NT;
            if (i < NV) d[i] = r[q];
        }
    }
};
template <typename T, int L> constexpr int tile_lds_elems()
{
    return tile_staged<T, L>() ? direct_slot_len(L, sizeof(T)) * MSTM_SLOTS : 1;
}
// staged sources: the tile's two source blocks (16 spheres x NR columns x 2
// modes x NB complex each) sit in LDS for the whole tile.  A sphere's columns
// are followed by two unused complex values, so that the 16 slots of a wave,
// reading 16 different spheres, spread over the LDS banks (other pads of
// columns and spheres measured no faster).
// complex values per column and per sphere of the staged sources
template <typename T, int L> constexpr int src_col() { return 2 * L * (L + 2); }
template <typename T, int L, int NR> constexpr int src_sph() { return NR * src_col<T, L>() + 2; }
template <typename T, int L, int NR> constexpr int tile_src_elems()
{
    return tile_src_staged<T, L, NR>() ? 2 * MSTM_SLOTS * src_sph<T, L, NR>() * 2 : 1;
}
// copy table group s of a tile into LDS, all threads of the workgroup
template <typename T, int L, int NR>
__device__ __forceinline__ void stage_group(const T *__restrict__ tab, int tile, int s, T *__restrict__ grp)
{
    constexpr int VW = 16 / sizeof(T);
    constexpr int NV = direct_slot_len(L, sizeof(T)) * MSTM_SLOTS / VW;
    typedef typename vec_of<T, VW>::type V;
    constexpr int NT = 64 * NR, NQ = (NV + NT - 1) / NT;
    const V *g = reinterpret_cast<const V *>(tab + ((int64_t)tile * MSTM_SLOTS + s) *
                                                       (int64_t)direct_slot_len(L, sizeof(T)) * MSTM_SLOTS);
    V *d = reinterpret_cast<V *>(grp);
    // all of this thread's loads in flight at once, then the LDS stores
    V r[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int i = threadIdx.x + q * NT;
        if (i < NV) r[q] = g[i];
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int i = threadIdx.x + q * NT;
        if (i < NV) d[i] = r[q];
    }
}
// One wave's work on a tile with the symmetric table layout (staged in LDS);
// DIR = 0 forward (I <- J), 1 reverse (J <- I); the wave serves right-hand
// sides 2*rpair and 2*rpair+1 of NR.  A step runs in three passes over one
// vector Y of NB values per lane: degree by degree, the phased source
// coefficients are folded into u/v and rotated into the blocks m and -m
// (Y(n,m)); block by block, the axial products replace Y(.,m) and Y(.,-m)
// (A_{-m} = A_m of the other mode); degree by degree, the back rotation of
// Y(t,.) gives P/Q and the output phase adds them into W(t,.).  Only W and
// Y live across the passes.
template <typename T, int L, int DIR, int NR, int NRL>
__device__ __forceinline__ void tile_body_sym(int2 bt, int tile, const T *__restrict__ tab, cplx<T> *__restrict__ part,
                                              int npad, int rpair, int r0, T *__restrict__ grp,
                                              const cplx<T> *__restrict__ srcl, const cplx<T> *__restrict__ xin,
                                              int s0, int s1, int G, int g)
{
    constexpr bool PIPE = tile_pipelined<T, L, NRL>();
    constexpr bool SRC = tile_src_staged<T, L, NRL>();
    constexpr int VW = 16 / sizeof(T);
    constexpr int NB = L * (L + 2);
    constexpr int S = MSTM_SLOTS;
    const int lane = threadIdx.x & 63;
    const int kslot = lane >> 2, rhs = 2 * rpair + ((lane >> 1) & 1), mode = lane & 1;
    const bool diag = bt.x == bt.y;
    const int tgt = (DIR ? bt.y : bt.x) * S + kslot;

    cplx<T> W[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) W[q] = cmk((T)0, (T)0);

    GroupPrefetch<T, L, NRL> pre;
    if constexpr (PIPE) pre.load(tab, tile, s0);

#pragma unroll 1
    for (int s = s0; s < s1; s++) {
        int slot, src;
        if (!DIR) {
            slot = kslot;
            src = bt.y * S + ((kslot + s) & (S - 1));
        } else {
            slot = (kslot - s + S) & (S - 1);
            src = bt.x * S + slot;
        }
        const cplx<T> *xs = xin + ((int64_t)(rhs * npad + src) * 2 + mode) * NB;
        // staged copy: block 0 = bt.y (forward sources), 1 = bt.x (reverse sources)
        const cplx<T> *xl = srcl + (DIR * S + (src % S)) * src_sph<T, L, NRL>() + (rhs - r0) * src_col<T, L>() + mode * NB;
        __syncthreads();
        cplx<T> xv[PIPE ? NB : 1];
        if constexpr (PIPE) {
            pre.store(grp);
            if (!(diag && s == 0)) {
#pragma unroll
                for (int n = 1; n <= L; n++)
#pragma unroll
                    for (int k = -n; k <= n; k++) xv[nk(n, k)] = xl[amn(DIR ? -k : k, n, L)];
            }
            if (s + 1 < s1) pre.load(tab, tile, s + 1);
        } else {
            stage_group<T, L, NRL>(tab, tile, s, grp);
        }
        __syncthreads();
        if (diag && s == 0) continue;
        const T *p0 = grp + slot * VW;
        // exp(i sigma k phi), k = 0..L; sigma = +1 forward, -1 reverse (im of :6690)
        auto phases = [&](cplx<T> *ph) {
            T pv[2];
            load_sec<T, sphi_off(L, VW), 2>(p0, pv);
            const cplx<T> e1 = cmk(pv[0], DIR ? -pv[1] : pv[1]);
            ph[0] = cmk((T)1, (T)0);
#pragma unroll
            for (int k = 1; k <= L; k++) ph[k] = cmul(ph[k - 1], e1);
        };
        cplx<T> ph[L + 1];
        phases(ph);
        // source coefficient (n,k) of this lane: m-flip for the reverse direction
        auto xsrc = [&](int n, int k) -> cplx<T> {
            cplx<T> v;
            if constexpr (PIPE) v = xv[nk(n, k)];
            else if constexpr (SRC) v = xl[amn(DIR ? -k : k, n, L)];
            else v = xs[amn(DIR ? -k : k, n, L)];
            if (DIR && (k & 1)) v = cmk(-v.re, -v.im);
            return v;
        };
        cplx<T> Y[NB];
        // ---- forward rotation, degree by degree
        static_for<L>([&](auto nc) {
            constexpr int n = decltype(nc)::value + 1;
            cplx<T> u[n + 1], v[n + 1];
            u[0] = xsrc(n, 0);
#pragma unroll
            for (int k = 1; k <= n; k++) {
                const cplx<T> a = cmul(ph[k], xsrc(n, k)), b = cmul(cconj(ph[k]), xsrc(n, -k));
                u[k] = (k & 1) ? csub(a, b) : cadd(a, b);
                v[k] = (k & 1) ? cadd(a, b) : csub(a, b);
            }
            static_for<n + 1>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                T dr[srow_len(m, n)];
                load_span<T, srow_off(m, n), srow_len(m, n)>(p0, dr);
                cplx<T> yp = cmk((T)0, (T)0);
#pragma unroll
                for (int k = 0; k <= n; k++) crfma(yp, dr[k], u[k]);
                if constexpr (m == 0) {
                    Y[nk(n, 0)] = yp;
                } else {
                    cplx<T> ym = cmk((T)0, (T)0);
#pragma unroll
                    for (int k = 1; k <= n; k++) crfma(ym, dr[n + k], v[k]);
                    Y[nk(n, m)] = cadd(yp, ym);
                    Y[nk(n, -m)] = (m & 1) ? csub(ym, yp) : csub(yp, ym);
                }
            });
        });
        // the back pass reloads its rows from LDS rather than keeping the
        // forward pass's copies live across the axial products
        asm volatile("" ::: "memory");
        // ---- axial products, block by block (transposed for the reverse direction)
        static_for<L + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int m1 = imx(1, m);
            constexpr int AL = sax_len(m, L);
            constexpr int AP = pad_to(AL, VW);
            constexpr int ND = L - m1 + 1;
            static_for<(m == 0 ? 1 : 2)>([&](auto hc) {
                constexpr int h = decltype(hc)::value;  // 0: block m (this mode), 1: block -m (other mode)
                constexpr int mm = h ? -m : m;
                T da[AL];
                const int md = m == 0 ? 0 : (h ? 1 - mode : mode);
                load_sec<T, sax_off(m, L, VW), AL>(p0 + (int64_t)md * (AP / VW) * S * VW, da);
                cplx<T> z[ND];
#pragma unroll
                for (int t = 0; t < ND; t++) {
                    cplx<T> acc = cmk((T)0, (T)0);
#pragma unroll
                    for (int q = 0; q < ND; q++) {
                        const int e = DIR ? 2 * (q + ND * t) : 2 * (t + ND * q);
                        cfma(acc, cmk(da[e], da[e + 1]), Y[nk(q + m1, mm)]);
                    }
                    z[t] = acc;
                }
#pragma unroll
                for (int t = 0; t < ND; t++) Y[nk(t + m1, mm)] = z[t];
            });
        });
        asm volatile("" ::: "memory");
        // ---- back rotation and output phase, degree by degree
        static_for<L>([&](auto tc) {
            constexpr int t = decltype(tc)::value + 1;
            cplx<T> P[t + 1], Q[t + 1];
#pragma unroll
            for (int k = 0; k <= t; k++) P[k] = Q[k] = cmk((T)0, (T)0);
            static_for<t + 1>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                T dr[srow_len(m, t)];
                load_span<T, srow_off(m, t), srow_len(m, t)>(p0, dr);
                if constexpr (m == 0) {
                    const cplx<T> za = Y[nk(t, 0)];
#pragma unroll
                    for (int k = 0; k <= t; k++) crfma(P[k], dr[k], za);
                } else {
                    const cplx<T> z = Y[nk(t, m)], zn = Y[nk(t, -m)];
                    const cplx<T> za = (m & 1) ? csub(z, zn) : cadd(z, zn);
                    const cplx<T> zb = (m & 1) ? cadd(z, zn) : csub(z, zn);
#pragma unroll
                    for (int k = 0; k <= t; k++) crfma(P[k], dr[k], za);
#pragma unroll
                    for (int k = 1; k <= t; k++) crfma(Q[k], dr[t + k], zb);
                }
            });
            // U_0 = P_0, U_k = P_k + Q_k, U_-k = (-1)^k (P_k - Q_k)
            W[nk(t, 0)] = cadd(W[nk(t, 0)], P[0]);
#pragma unroll
            for (int k = 1; k <= t; k++) {
                const cplx<T> up = cadd(P[k], Q[k]);
                const cplx<T> um = (k & 1) ? csub(Q[k], P[k]) : csub(P[k], Q[k]);
                cfma(W[nk(t, k)], cconj(ph[k]), up);
                cfma(W[nk(t, -k)], ph[k], um);
            }
        });
    }
    // partial for (target, partner block); m-flip back for the reverse wave
    const int partner = (DIR ? bt.x : bt.y) * G + g;  // one partial per (partner block, step group)
    cplx<T> *out = part + (((int64_t)partner * NR + rhs) * npad + tgt) * 2 * NB + mode * NB;
#pragma unroll
    for (int n = 1; n <= L; n++)
#pragma unroll
        for (int k = -n; k <= n; k++) {
            cplx<T> v;
            if (!DIR) v = W[nk(n, k)];
            else {
                v = W[nk(n, -k)];
                if (k & 1) v = cmk(-v.re, -v.im);
            }
            out[amn(k, n, L)] = v;
        }
}

// One wave's work on a tile with the full table layout (orders above the
// symmetric range): tables read from L2/HBM directly, output phase folded
// into every term; DIR and rpair as for tile_body_sym.
template <typename T, int L, int DIR, int NR, int NRL>
__device__ __forceinline__ void tile_body(const T *__restrict__ tab, int2 bt, int tile, const cplx<T> *__restrict__ xin,
                                          cplx<T> *__restrict__ part, int npad, int rpair, int r0, T *__restrict__ grp,
                                          const cplx<T> *__restrict__ srcl, int s0, int s1, int G, int g)
{
    constexpr bool STAGED = tile_staged<T, L>();
    constexpr bool PIPE = tile_pipelined<T, L, NRL>();
    constexpr int VW = 16 / sizeof(T);
    constexpr int NB = L * (L + 2);
    constexpr int SL = slot_len(L, L, VW);
    constexpr int S = MSTM_SLOTS;
    const int lane = threadIdx.x & 63;
    const int kslot = lane >> 2, rhs = 2 * rpair + ((lane >> 1) & 1), mode = lane & 1;
    const bool diag = bt.x == bt.y;
    const int tgt = (DIR ? bt.y : bt.x) * S + kslot;

    cplx<T> W[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) W[q] = cmk((T)0, (T)0);

    GroupPrefetch<T, L, NRL> pre;
    if constexpr (PIPE) pre.load(tab, tile, s0);

#pragma unroll 1
    for (int s = s0; s < s1; s++) {
        int slot, src;
        if (!DIR) {
            slot = kslot;
            src = bt.y * S + ((kslot + s) & (S - 1));
        } else {
            slot = (kslot - s + S) & (S - 1);
            src = bt.x * S + slot;
        }
        const cplx<T> *xs = xin + ((int64_t)(rhs * npad + src) * 2 + mode) * NB;
        // staged copy: block 0 = bt.y (forward sources), 1 = bt.x (reverse sources)
        const cplx<T> *xl = srcl + (DIR * S + (src % S)) * src_sph<T, L, NRL>() + (rhs - r0) * src_col<T, L>() + mode * NB;
        // all waves of the tile read the same table group in step s: keep them
        // in step so that the group is fetched from HBM once (and, staged,
        // copied to LDS once)
        __syncthreads();
        cplx<T> xv[PIPE ? NB : 1];
        if constexpr (PIPE) {
            pre.store(grp);
            // this step's source coefficients are loaded before the next
            // group's prefetch, so that waiting for them never waits for it
            if (!(diag && s == 0)) {
#pragma unroll
                for (int n = 1; n <= L; n++)
#pragma unroll
                    for (int k = -n; k <= n; k++) xv[nk(n, k)] = xl[amn(DIR ? -k : k, n, L)];
            }
            if (s + 1 < s1) pre.load(tab, tile, s + 1);
            __syncthreads();
        } else if constexpr (STAGED) {
            stage_group<T, L, NRL>(tab, tile, s, grp);
            __syncthreads();
        }
        if (diag && s == 0) continue;
        const T *p0 = STAGED ? grp + slot * VW : tab + ((int64_t)tile * S + s) * (int64_t)SL * S + slot * VW;
        // exp(i sigma k phi), sigma = +1 forward, -1 reverse (im of :6690)
        T pv[2];
        load_sec<T, phi_off(L, L, VW), 2>(p0, pv);
        cplx<T> e1 = cmk(pv[0], DIR ? -pv[1] : pv[1]);
        cplx<T> ph[2 * L + 1];
        ph[L] = cmk((T)1, (T)0);
#pragma unroll
        for (int k = 1; k <= L; k++) {
            ph[L + k] = cmul(ph[L + k - 1], e1);
            ph[L - k] = cconj(ph[L + k]);
        }
        // source coefficients: m-flip for the reverse direction, then phase
        cplx<T> xp[NB];
#pragma unroll
        for (int n = 1; n <= L; n++)
#pragma unroll
            for (int k = -n; k <= n; k++) {
                cplx<T> v;
                if constexpr (PIPE) v = xv[nk(n, k)];
                else v = xs[amn(DIR ? -k : k, n, L)];
                if (DIR && (k & 1)) v = cmk(-v.re, -v.im);
                xp[nk(n, k)] = cmul(ph[L + k], v);
            }
        // this step's back-rotated sums before the output phase (staged form)
        cplx<T> U[STAGED ? NB : 1];
#pragma unroll
        for (int q = 0; q < (STAGED ? NB : 1); q++) U[q] = cmk((T)0, (T)0);
        static_for<2 * L + 1>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            constexpr int m = mblock_m(b);
            constexpr int m1 = imx(1, iab(m));
            constexpr int E0 = sec_off(b, L, L, VW);
            constexpr int RL = rot_len(m, L, L);
            constexpr int AL = ax_len(m, L, L);
            constexpr int AP = pad_to(AL, VW);
            constexpr int ND = L - m1 + 1;
            T dr[RL];
            load_sec<T, E0, RL>(p0, dr);
            // forward rotation: y(n) = sum_k d^n(m,k) xp(n,k)
            cplx<T> y[ND];
#pragma unroll
            for (int n = m1; n <= L; n++) {
                int ro = 0;
#pragma unroll
                for (int q = m1; q < n; q++) ro += 2 * q + 1;
                cplx<T> acc = cmk((T)0, (T)0);
#pragma unroll
                for (int k = -n; k <= n; k++) crfma(acc, dr[ro + k + n], xp[nk(n, k)]);
                y[n - m1] = acc;
            }
            // axial translation of this m block (mode-specific matrix,
            // transposed for the reverse direction)
            T da[AL];
            load_sec<T, E0 + pad_to(RL, VW), AL>(p0 + (int64_t)mode * (AP / VW) * S * VW, da);
            cplx<T> z[ND];
#pragma unroll
            for (int t = 0; t < ND; t++) {
                cplx<T> acc = cmk((T)0, (T)0);
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    const int e = DIR ? 2 * (q + ND * t) : 2 * (t + ND * q);
                    cfma(acc, cmk(da[e], da[e + 1]), y[q]);
                }
                z[t] = acc;
            }
            // back rotation: staged, U(t,k) += d^t(m,k) z(t); otherwise with
            // the output phase folded in, W(t,k) += d^t(m,k) exp(-i sigma k phi) z(t)
#pragma unroll
            for (int t = m1; t <= L; t++) {
                int ro = 0;
#pragma unroll
                for (int q = m1; q < t; q++) ro += 2 * q + 1;
#pragma unroll
                for (int k = -t; k <= t; k++) {
                    if constexpr (STAGED) {
                        crfma(U[nk(t, k)], dr[ro + k + t], z[t - m1]);
                    } else {
                        cplx<T> f = cscale(ph[L - k], dr[ro + k + t]);
                        cfma(W[nk(t, k)], f, z[t - m1]);
                    }
                }
            }
        });
        if constexpr (STAGED) {
            // output phase once per coefficient: W(t,k) += exp(-i sigma k phi) U(t,k)
#pragma unroll
            for (int t = 1; t <= L; t++)
#pragma unroll
                for (int k = -t; k <= t; k++) cfma(W[nk(t, k)], ph[L - k], U[nk(t, k)]);
        }
    }
    // partial for (target, partner block); m-flip back for the reverse wave
    const int partner = (DIR ? bt.x : bt.y) * G + g;  // one partial per (partner block, step group)
    cplx<T> *out = part + (((int64_t)partner * NR + rhs) * npad + tgt) * 2 * NB + mode * NB;
#pragma unroll
    for (int n = 1; n <= L; n++)
#pragma unroll
        for (int k = -n; k <= n; k++) {
            cplx<T> v;
            if (!DIR) v = W[nk(n, k)];
            else {
                v = W[nk(n, -k)];
                if (k & 1) v = cmk(-v.re, -v.im);
            }
            out[amn(k, n, L)] = v;
        }
}

// NR = 2 (one BiCG product pair): 2 waves, forward and reverse.  NR = 4
// (both polarizations' pairs at once): a workgroup serves NRL of the NR
// columns, two waves per column pair.  NRL = 4: waves 0/1 serve columns 0-1
// and waves 2/3 columns 2-3 from the same staged table group, so every table
// byte read from HBM feeds four columns.  NRL = 2 (FP64): two workgroups per
// tile, one per column pair, each staging its own group and sources; the
// two run on the same XCD (workgroups are dealt to MI355X's 8 XCDs round
// robin), so the second reads the group from that XCD's L2.
// G workgroups share a tile (small clusters, G = 2..16): workgroup g takes
// the steps [g S/G, (g+1) S/G) and writes its own partials.
template <typename T, int L, int NR, int NRL>
__global__ __launch_bounds__(64 * NRL) __attribute__((amdgpu_waves_per_eu(tile_waves<T, L, NRL>()))) void direct_tile_kernel(
    const T *__restrict__ tab, const int2 *__restrict__ tiles, int nunits, const cplx<T> *__restrict__ xin,
    cplx<T> *__restrict__ part, int npad, int G, const int *__restrict__ gate)
{
    if (gate && !*gate) return;  // the issuing solver has stopped every system
    constexpr int H = NR / NRL;  // workgroups per (tile, step group)
    int u = blockIdx.x, h = 0;
    if constexpr (H > 1) {
        u = (int)(blockIdx.x / (8 * H)) * 8 + (int)(blockIdx.x % 8);
        h = (int)(blockIdx.x / 8) % H;
        if (u >= nunits) return;  // the grid is padded to whole groups of 8 H
    }
    const int tile = u / G, g = u - tile * G;
    const int s0 = g * (MSTM_SLOTS / G), s1 = s0 + MSTM_SLOTS / G;
    const int r0 = h * NRL;  // first column of this workgroup
    const int2 bt = tiles[tile];
    const int w = threadIdx.x >> 6;
    __shared__ __attribute__((aligned(16))) T grp[tile_lds_elems<T, L>()];
    __shared__ __attribute__((aligned(16))) T srcs[tile_src_elems<T, L, NRL>()];
    cplx<T> *srcl = reinterpret_cast<cplx<T> *>(srcs);
    if constexpr (tile_src_staged<T, L, NRL>()) {
        // the two source blocks of the tile (this workgroup's columns), all
        // threads; visible after the first step's barrier
        constexpr int NB = L * (L + 2), BLK = MSTM_SLOTS * 2 * NB;  // one column r: 16 spheres x 2 modes
        for (int e = threadIdx.x; e < 2 * NRL * BLK; e += 64 * NRL) {
            const int b = e / (NRL * BLK), rem = e - b * NRL * BLK, r = rem / BLK, o = rem - r * BLK;
            const int blk = b == 0 ? bt.y : bt.x;
            const int i = o / (2 * NB), mo = o - i * 2 * NB;  // sphere in block, (mode, coefficient)
            srcl[(b * MSTM_SLOTS + i) * src_sph<T, L, NRL>() + r * src_col<T, L>() + mo] =
                xin[((int64_t)((r0 + r) * npad + blk * MSTM_SLOTS + i) * 2) * NB + mo];
        }
    }
    const int rpair = (r0 >> 1) + (w >> 1);
    if constexpr (tile_staged<T, L>()) {
        if (!(w & 1))
            tile_body_sym<T, L, 0, NR, NRL>(bt, tile, tab, part, npad, rpair, r0, grp, srcl, xin, s0, s1, G, g);
        else if (bt.x != bt.y)
            tile_body_sym<T, L, 1, NR, NRL>(bt, tile, tab, part, npad, rpair, r0, grp, srcl, xin, s0, s1, G, g);
    } else {
        if (!(w & 1))
            tile_body<T, L, 0, NR, NRL>(tab, bt, tile, xin, part, npad, rpair, r0, grp, srcl, s0, s1, G, g);
        else if (bt.x != bt.y)
            tile_body<T, L, 1, NR, NRL>(tab, bt, tile, xin, part, npad, rpair, r0, grp, srcl, s0, s1, G, g);
    }
    if (bt.x == bt.y && (w & 1))
    {
        // idle reverse wave of a diagonal tile: its share of the staging and the same barriers
        GroupPrefetch<T, L, NRL> pre;
        if constexpr (tile_pipelined<T, L, NRL>()) pre.load(tab, tile, s0);
        for (int s = s0; s < s1; s++) {
            __syncthreads();
            if constexpr (tile_pipelined<T, L, NRL>()) {
                pre.store(grp);
                if (s + 1 < s1) pre.load(tab, tile, s + 1);
                __syncthreads();
            } else if constexpr (tile_staged<T, L>()) {
                stage_group<T, L, NRL>(tab, tile, s, grp);
                __syncthreads();
            }
        }
    }
}

// columns per workgroup of a four-column product: FP64 tiles with staged
// sources split the four columns over two workgroups (the staged sources of
// four FP64 columns would leave one workgroup per CU)
template <typename T, int L> constexpr int tile4_nrl() { return tile_src_staged<T, L, 2>() && sizeof(T) == 8 ? 2 : 4; }

template <typename T>
hipError_t launch_direct_tiles(const T *tab, const int2 *tiles, int ntiles, int L, const T *xin, T *partial, int npad,
                               int nr, int G, hipStream_t st, const int *gate)
{
    const cplx<T> *x = reinterpret_cast<const cplx<T> *>(xin);
    cplx<T> *p = reinterpret_cast<cplx<T> *>(partial);
    if (nr != 2 && nr != 4) return hipErrorInvalidValue;
    if (G < 1 || G > MSTM_SLOTS || MSTM_SLOTS % G) return hipErrorInvalidValue;
    const int nunits = ntiles * G;
    switch (L * 8 + nr) {
#define CASE(LL)                                                                                                       \
    case LL * 8 + 2:                                                                                                   \
        hipLaunchKernelGGL((direct_tile_kernel<T, LL, 2, 2>), dim3((unsigned)nunits), dim3(128), 0, st, tab, tiles,    \
                           nunits, x, p, npad, G, gate);                                                               \
        break;                                                                                                         \
    case LL * 8 + 4: {                                                                                                 \
        constexpr int NRL = tile4_nrl<T, LL>(), H = 4 / NRL;                                                           \
        const unsigned nb = H == 1 ? (unsigned)nunits : (unsigned)((nunits + 7) / 8 * 8 * H);                          \
        hipLaunchKernelGGL((direct_tile_kernel<T, LL, 4, NRL>), dim3(nb), dim3(64 * NRL), 0, st, tab, tiles, nunits,   \
                           x, p, npad, G, gate);                                                                       \
        break;                                                                                                         \
    }
#ifdef MSTM_TILE_L_ONLY
        CASE(MSTM_TILE_L_ONLY)
#else
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#endif
#undef CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_direct_tiles<float>(const float *, const int2 *, int, int, const float *, float *, int, int,
                                               int, hipStream_t, const int *);
template hipError_t launch_direct_tiles<double>(const double *, const int2 *, int, int, const double *, double *, int,
                                                int, int, hipStream_t, const int *);

// ---------------------------------------------------------------------------
// Input preparation (sphereinteraction :88-97): for a con_tran column
// x = S T^T conj(a) (S = m-flip with (-1)^m, T^T the transposed 2x2 Mie
// matrix when mie_mult); otherwise x = a.  The tile kernels work on a
// uniform-order copy (order Lp = max order, npad spheres); coefficients of
// degree n > order(i) and padding spheres are zero, which is exact because
// a translation table of order Lp contains those of lower orders.
// flags[r] (host array, r < nrhs <= 4) bit0 = con_tran, bit1 = mie_mult.
// One thread per (r, sphere, e), r < nr (2 or 4 columns in xin).
// an: Mie matrices, 4 complex per degree, at (sphere*Lp + n-1)*4.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode_mn(int e, int L, int &m, int &n)
{
    int row = e % (L + 2), colm = e / (L + 2) + 1;
    if (row <= colm) {
        m = row;
        n = colm;
    } else {
        m = -colm;
        n = row - 1;
    }
}

template <typename T>
__global__ void prep_kernel(const cplx<double> *__restrict__ ain, int ldain, cplx<T> *__restrict__ xin, int nrhs,
                            int4 flags, int nr, int nsph, int npad, int Lp, const int *__restrict__ order,
                            const int *__restrict__ offset, const cplx<double> *__restrict__ an,
                            const int *__restrict__ gate)
{
    if (gate && !*gate) return;
    const int NBP = Lp * (Lp + 2);
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nr * npad * NBP) return;
    int e = (int)(q % NBP);
    int64_t t = q / NBP;
    int sph = (int)(t % npad);
    int r = (int)(t / npad);
    cplx<T> *o = xin + ((int64_t)(r * npad + sph) * 2) * NBP;
    int m, n;
    decode_mn(e, Lp, m, n);
    if (r >= nrhs || sph >= nsph || n > order[sph]) {
        o[e] = cmk((T)0, (T)0);
        o[e + NBP] = cmk((T)0, (T)0);
        return;
    }
    const int L = order[sph], NB = L * (L + 2);
    const int f = r == 0 ? flags.x : (r == 1 ? flags.y : (r == 2 ? flags.z : flags.w));
    const cplx<double> *a = ain + (int64_t)r * ldain + offset[sph];
    if (!(f & 1)) {
        int es = amn(m, n, L);
        cplx<double> v0 = a[es], v1 = a[es + NB];
        o[e] = cmk((T)v0.re, (T)v0.im);
        o[e + NBP] = cmk((T)v1.re, (T)v1.im);
        return;
    }
    // flipped value at (m,n) is (-1)^m (T^T conj a)(-m,n)
    int es = amn(-m, n, L);
    cplx<double> v0 = cconj(a[es]), v1 = cconj(a[es + NB]);
    if (f & 2) {
        const cplx<double> *t4 = an + ((int64_t)sph * Lp + (n - 1)) * 4;  // t4[p + 2q] = T(p,q)
        cplx<double> w0 = cadd(cmul(t4[0], v0), cmul(t4[1], v1));
        cplx<double> w1 = cadd(cmul(t4[2], v0), cmul(t4[3], v1));
        v0 = w0;
        v1 = w1;
    }
    double sg = (m & 1) ? -1.0 : 1.0;
    o[e] = cmk((T)(sg * v0.re), (T)(sg * v0.im));
    o[e + NBP] = cmk((T)(sg * v1.re), (T)(sg * v1.im));
}

static int4 flags4(const int *flags, int nrhs)
{
    return make_int4(flags[0], nrhs > 1 ? flags[1] : 0, nrhs > 2 ? flags[2] : 0, nrhs > 3 ? flags[3] : 0);
}

template <typename T>
hipError_t launch_prep(const double *ain, int ldain, T *xin, int nrhs, const int *flags, int nsph, int npad, int Lp,
                       const int *order, const int *offset, const double *an, hipStream_t st, const int *gate)
{
    if (nrhs < 1 || nrhs > 4) return hipErrorInvalidValue;
    const int nr = nrhs > 2 ? 4 : 2;
    int NBP = Lp * (Lp + 2);
    int64_t n = (int64_t)nr * npad * NBP;
    int bs = 256;
    hipLaunchKernelGGL(prep_kernel<T>, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st,
                       reinterpret_cast<const cplx<double> *>(ain), ldain / 2, reinterpret_cast<cplx<T> *>(xin), nrhs,
                       flags4(flags, nrhs), nr, nsph, npad, Lp, order, offset,
                       reinterpret_cast<const cplx<double> *>(an), gate);
    return hipGetLastError();
}
template hipError_t launch_prep<float>(const double *, int, float *, int, const int *, int, int, int, const int *,
                                       const int *, const double *, hipStream_t, const int *);
template hipError_t launch_prep<double>(const double *, int, double *, int, const int *, int, int, int, const int *,
                                        const int *, const double *, hipStream_t, const int *);

// ---------------------------------------------------------------------------
// Sum of the per-block partials + post-op (sphereinteraction :158-168):
// normal column  out = T h  (mie_mult) or h;  con_tran column out = conj(S h).
// Written back in the reference (unpadded) layout.
// ---------------------------------------------------------------------------
// KS lanes share an output (r, sphere, e): lane j sums the partials
// b = j, j + KS, ... and the KS sums are then added in a fixed tree order
// within the wave, so the result is deterministic; many partials per output
// (small clusters, whose tiles are split in step groups) then keep many
// loads in flight instead of one long dependent loop per thread.
template <typename T, int KS>
__global__ __launch_bounds__(256) void reduce_post_kernel(const cplx<T> *__restrict__ part, int nparts,
                                                          cplx<double> *__restrict__ aout, int ldaout, int nrhs,
                                                          int4 flags, int nr, int nsph, int npad, int Lp,
                                                          const int *__restrict__ order, const int *__restrict__ offset,
                                                          const cplx<double> *__restrict__ an,
                                                          const int *__restrict__ gate)
{
    if (gate && !*gate) return;  // uniform over the grid
    const int NBP = Lp * (Lp + 2);
    const int j = threadIdx.x % KS;
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / KS;
    const int64_t nout = (int64_t)nrhs * nsph * NBP;
    // out-of-range and padding items still take part in the shuffles
    const bool live = q < nout;
    int e = 0, sph = 0, r = 0, m = 0, n = 1, L = 0, f = 0, es = 0;
    if (live) {
        e = (int)(q % NBP);
        const int64_t t = q / NBP;
        sph = (int)(t % nsph);
        r = (int)(t / nsph);
        decode_mn(e, Lp, m, n);
        L = order[sph];
        f = r == 0 ? flags.x : (r == 1 ? flags.y : (r == 2 ? flags.z : flags.w));
        es = (f & 1) ? amn(-m, n, Lp) : e;
    }
    const bool use = live && n <= L;
    cplx<double> h0 = cmk(0.0, 0.0), h1 = cmk(0.0, 0.0);
    if (use) {
        const int64_t stride = (int64_t)nr * npad * 2 * NBP;
        const cplx<T> *p = part + ((int64_t)r * npad + sph) * 2 * NBP + es;
#pragma unroll 4
        for (int b = j; b < nparts; b += KS) {
            const cplx<T> a0 = p[(int64_t)b * stride], a1 = p[(int64_t)b * stride + NBP];
            h0.re += (double)a0.re;
            h0.im += (double)a0.im;
            h1.re += (double)a1.re;
            h1.im += (double)a1.im;
        }
    }
#pragma unroll
    for (int w = 1; w < KS; w <<= 1) {
        h0.re += __shfl_xor(h0.re, w, KS);
        h0.im += __shfl_xor(h0.im, w, KS);
        h1.re += __shfl_xor(h1.re, w, KS);
        h1.im += __shfl_xor(h1.im, w, KS);
    }
    if (!use || j) return;
    const int NB = L * (L + 2);
    cplx<double> *o = aout + (int64_t)r * ldaout + offset[sph];
    int eo = amn(m, n, L);
    if (f & 1) {
        double sg = (m & 1) ? -1.0 : 1.0;
        o[eo] = cmk(sg * h0.re, -sg * h0.im);
        o[eo + NB] = cmk(sg * h1.re, -sg * h1.im);
    } else if (f & 2) {
        const cplx<double> *t4 = an + ((int64_t)sph * Lp + (n - 1)) * 4;
        o[eo] = cadd(cmul(t4[0], h0), cmul(t4[2], h1));
        o[eo + NB] = cadd(cmul(t4[1], h0), cmul(t4[3], h1));
    } else {
        o[eo] = h0;
        o[eo + NB] = h1;
    }
}

template <typename T>
hipError_t launch_reduce_post(const T *partial, int nparts, double *aout, int ldaout, int nrhs, const int *flags,
                              int nsph, int npad, int Lp, const int *order, const int *offset, const double *an,
                              hipStream_t st, const int *gate)
{
    if (nrhs < 1 || nrhs > 4) return hipErrorInvalidValue;
    const int nr = nrhs > 2 ? 4 : 2;
    int NBP = Lp * (Lp + 2);
    const int64_t n = (int64_t)nrhs * nsph * NBP;
    const int bs = 256;
#define RP(KS)                                                                                                         \
    hipLaunchKernelGGL((reduce_post_kernel<T, KS>), dim3((unsigned)((n * KS + bs - 1) / bs)), dim3(bs), 0, st,         \
                       reinterpret_cast<const cplx<T> *>(partial), nparts, reinterpret_cast<cplx<double> *>(aout),     \
                       ldaout / 2, nrhs, flags4(flags, nrhs), nr, nsph, npad, Lp, order, offset,                       \
                       reinterpret_cast<const cplx<double> *>(an), gate)
    // lanes per output: with many outputs one lane each keeps the loads of
    // a wave on consecutive elements of one partial; with few (small
    // clusters, many step-group partials) several lanes per output keep
    // enough loads in flight
    if (n >= 65536) RP(1);
    else if (nparts >= 128) RP(8);
    else if (nparts >= 32) RP(4);
    else RP(1);
#undef RP
    return hipGetLastError();
}
template hipError_t launch_reduce_post<float>(const float *, int, double *, int, int, const int *, int, int, int,
                                              const int *, const int *, const double *, hipStream_t, const int *);
template hipError_t launch_reduce_post<double>(const double *, int, double *, int, int, const int *, int, int, int,
                                               const int *, const int *, const double *, hipStream_t, const int *);
