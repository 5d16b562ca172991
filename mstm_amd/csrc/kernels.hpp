// kernels.hpp -- launcher declarations shared by the .hip translation units.
#pragma once
#include <vector>

#include "common.hpp"

// context.hip: vector-coupling constants of axialtrancoefinit for orders <= LM
void vcc_table(int LM, std::vector<double> &h);

// tables.hip
template <typename T>
hipError_t launch_direct_tables(T *tab, const int2 *tiles, int ntiles, int L, int nsph, const double *pos, double rimed,
                                const double *vcc, int LM, hipStream_t st);
template <typename T>
hipError_t launch_list_tables(T *tab, const double *rtran, int nslots, int LT, int LS, int itype, double rimed,
                              const double *vcc, int LM, hipStream_t st);
// the FP64 and FP32 direct tables in one pass: orders whose layout kind is
// the same at both precisions (direct_tables_both_ok)
bool direct_tables_both_ok(int L);
hipError_t launch_direct_tables_both(double *tab64, float *tab32, const int2 *tiles, int ntiles, int L, int nsph,
                                     const double *pos, double rimed, const double *vcc, int LM, hipStream_t st);

// direct.hip
// gate (optional, device int): the kernels of a product return at once
// when *gate == 0 (every system of the Krylov loop that issued the product
// has stopped), so that products enqueued ahead of the host's next read of
// the solver state cost little.
template <typename T>
// G (1, 2, 4, 8 or 16) workgroups per tile share its 16 steps; the partials
// are then (partner block x G) per target
hipError_t launch_direct_tiles(const T *tab, const int2 *tiles, int ntiles, int L, const T *xin, T *partial, int npad,
                               int nr, int G, hipStream_t st, const int *gate = nullptr);
template <typename T>
hipError_t launch_prep(const double *ain, int ldain, T *xin, int nrhs, const int *flags, int nsph, int npad, int Lp,
                       const int *order, const int *offset, const double *an, hipStream_t st,
                       const int *gate = nullptr);
template <typename T>
hipError_t launch_reduce_post(const T *partial, int nparts, double *aout, int ldaout, int nrhs, const int *flags,
                              int nsph, int npad, int Lp, const int *order, const int *offset, const double *an,
                              hipStream_t st, const int *gate = nullptr);

// mie.hip
hipError_t launch_mie(const double *x, const double *mre, const double *mim, double mmed, int nsph, double qeps,
                      int *order, double *an, int ldan, hipStream_t st);
hipError_t launch_mie_mult(const double *ain, double *aout, int nsph, const int *order, const int *offset,
                           const int *mieoff, const double *an, int idir, int inverse, int Lmax, hipStream_t st);
hipError_t launch_planewave(double *pmnp, int ld, int nsph, const double *pos, const int *order, const int *offset,
                            double alpha, double sinc, int dir, double rimed, hipStream_t st,
                            const int *excited = nullptr);
hipError_t launch_qeff(const double *amnp, const double *pmnp, int ld, int nsph, const double *radius, const int *order,
                       const int *offset, const int *mieoff, const double *an, double rimed, double *qeff,
                       hipStream_t st);

// krylov.hip (all vectors complex FP64, lengths in complex elements)
int krylov_grid();
// Scalars of one BiCG system kept on the device (cbicg's csk, cak, cbk,
// enorm and the convergence state), so that an iteration needs no host
// reduction: the host reads the array every few iterations to steer.
struct BicgScal {
    cplx<double> csk, cak, cbk;
    double enorm, err, eps;
    int active, it, brk;
    int cur;  // iterations begun while active (the loop's iteration number)
};
// step 0: enorm = |b|^2 (active = enorm != 0); 1: csk = r.r (inactive if 0);
// 2: a new iteration (cur + 1), den = <w, A p>, cak = csk / den (breakdown:
// inactive, it = cur, brk = 1); 3: <q, r>, <r, r>: err, stop test (it = cur),
// cbk, csk.  The iteration number is counted on the device, so the launches
// of an iteration are the same every time (they can be replayed from a
// graph).  partials: region s of the system s at partials + s * region
// (complex), one or two values per block.  gate (optional): set to 1 while
// any system is active, else 0.
hipError_t launch_bicg_scal(int step, const double *partials, int64_t region, BicgScal *st, int ns, hipStream_t st_,
                            int *gate = nullptr);
// the vector steps of ns systems in one launch: system s's vectors at the
// given pointers plus s * sv (x, r, q) or s * sp (p, w, Ap, A^H w) doubles,
// its partial sums at partials + s * region complex values
hipError_t launch_bicg_a_dev(double *cap, double *caw, const double *cp, const double *cw, int64_t sp, int n,
                             double *partials, int64_t region, const BicgScal *sc, int ns, hipStream_t st);
hipError_t launch_bicg_b_dev(double *anp, double *cr, double *cq, const double *cp, const double *cap,
                             const double *caw, int64_t sv, int64_t sp, const BicgScal *sc, int n, double *partials,
                             int64_t region, int ns, hipStream_t st);
hipError_t launch_bicg_c_dev(double *cp, double *cw, const double *cr, const double *cq, int64_t sv, int64_t sp,
                             const BicgScal *sc, int n, int ns, hipStream_t st);
hipError_t launch_zdot(const double *a, const double *b, int n, int conj_a, double *partials, double *result,
                       hipStream_t st);
hipError_t launch_bicg_a(double *cap, double *caw, const double *cp, const double *cw, int n, double *partials,
                         hipStream_t st);
hipError_t launch_bicg_b(double *anp, double *cr, double *cq, const double *cp, const double *cap, const double *caw,
                         double cakr, double caki, int n, double *partials, hipStream_t st);
hipError_t launch_bicg_c(double *cp, double *cw, const double *cr, const double *cq, double cbkr, double cbki, int n,
                         hipStream_t st);
hipError_t launch_bicg_init(double *cr, double *cq, double *cw, double *cp, const double *pnp, const double *anp, int n,
                            double *partials, hipStream_t st);
hipError_t launch_zaxpby(double ar, double ai, const double *x, double br, double bi, double *y, int n, hipStream_t st);
// block Gram-Schmidt (GMRES): V_j at V + j*stride (doubles), out/h device arrays of nvec complex
int multi_vec_max();
hipError_t launch_multi_zdot(const double *V, int64_t stride, int nvec, const double *w, int n, double *partials,
                             double *out, hipStream_t st);
hipError_t launch_multi_axpy(const double *V, int64_t stride, int nvec, const double *h, double sign, double *w, int n,
                             hipStream_t st);

// observe.hip (complex FP64 throughout)
hipError_t launch_merge(const double *amnp, int64_t ldc, const int *order, const int *offset, const int *torder,
                        const double *pos, int nsph, int T, int LS, const double *r0, double rimed, const double *vccm,
                        double *part, int G, double *a0, hipStream_t st);
hipError_t launch_scatmat(const double *a0, int T, int nang, const double *ct, const double *phi, double rimed,
                          double *sm, hipStream_t st);
hipError_t launch_distribute(int l, const int *order, const int *offset, const double *pos, int nsph, int LS,
                             const double *r0, double rimed, const double *vccd, double *out, int64_t ldc, int G,
                             hipStream_t st);
int s00_moment_blocks(int T);

// lu.hip (complex FP64, column-major, leading dimension n)
hipError_t launch_lu_unit(double *x, int64_t ld, int n, int j0, int ncol, hipStream_t st);
hipError_t launch_lu_expand(const double *tab, int slot_elems, const int2 *pairs, int npairs, int L, int n,
                            const double *an, double *A, hipStream_t st);
hipError_t launch_lu_identity(double *A, int n, hipStream_t st);
hipError_t launch_lu_store(double *A, int n, int j0, int ncol, const double *th, int64_t ld, hipStream_t st);
size_t lu_factor_scratch_bytes();
hipError_t launch_lu_factor(double *LUm, int n, double *vv, int *piv, int *err, void *scratch, hipStream_t st);
hipError_t launch_lu_solve(const double *LUm, int n, const int *piv, double *b, int64_t ld, int nrhs, hipStream_t st);
hipError_t launch_lu_resid(const double *A, int n, const double *x, const double *p, double *r, hipStream_t st);
hipError_t launch_s00_moments(const double *a0, int T, const double *cgm, const double *cg1, double *out,
                              hipStream_t st);
// fosmexpansion's eight sequences over w = 0..2T (out[e*(2T+1) + w]); part:
// fosm_blocks(T) * 8 * (2T+1) doubles; bc: bcof table, leading dimension ldb >= 4T+2
int fosm_blocks(int T);
hipError_t launch_fosm(const double *a0, int T, const double *bc, int ldb, double *part, double *out, hipStream_t st);
