/*
 * mstm_hip.h -- C-ABI of the MI355X implementation of MSTM's solution hot
 * path (libmstm_hip.so).  Plain pointers and sizes only; every complex
 * array is interleaved (re, im) doubles, i.e. Fortran complex(8), laid out
 * exactly as MSTM's coefficient vectors: per sphere a block a(0:L+1,1:L,1:2)
 * at sphere_offset(i) (mstm-modules-33.f90:2351-2377, 5354-5376).
 *
 * Each entry point names the reference procedure whose contract it takes
 * over (paths relative to MSTM's december2023/ tree).  The Fortran module
 * mstm_amd/fortran/mstm_hip_c.f90 binds them with ISO_C_BINDING, and
 * mstm_amd/fortran/mstm_hip_solver.f90 wraps them in procedures with the
 * argument lists of cbicg, fixedorsoln and sphereinteraction.
 *
 * Return codes: 0 = success, otherwise an MSTM_HIP_E* code and
 * mstm_hip_last_error() holds the message.  Calls on one context must come
 * from one host thread; all device work runs on the context's stream.
 */
#ifndef MSTM_HIP_H
#define MSTM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSTM_HIP_OK 0
#define MSTM_HIP_EARG 1    /* invalid argument or size                    */
#define MSTM_HIP_EHIP 2    /* HIP runtime error (allocation, launch, ...)  */
#define MSTM_HIP_ESTATE 3  /* call out of order (e.g. no spheres yet)      */
#define MSTM_HIP_EUNSUP 4  /* configuration not supported by this build    */

#define MSTM_HIP_FP64 0 /* plain: FP64 tables and arithmetic               */
#define MSTM_HIP_FP32 1 /* mixed: FP32 tables and arithmetic               */
#define MSTM_HIP_FP64_FP32 2 /* mstm_hip_setup_direct only: both tables      */

#define MSTM_HIP_SOLVE_PLAIN 0   /* cbicg in FP64, iteration-identical         */
#define MSTM_HIP_SOLVE_MIXED 1   /* FP32 BiCG inside FP64 iterative refinement */
#define MSTM_HIP_SOLVE_LEARNED 2 /* MIXED started from the learned guess       */
/* FP64 flexible GMRES(8) with FP32 BiCG inner solves; a loaded learned
 * network is the inner preconditioner of the first Arnoldi vector */
#define MSTM_HIP_SOLVE_FGMRES 3
/* direct solution (solution_method other than 'i'): dense A = I - T H and its
 * LU factors (up to 16384 equations), both polarizations back-substituted;
 * iterations 0 */
#define MSTM_HIP_SOLVE_LU 4

typedef struct mstm_hip_ctx mstm_hip_ctx;

/* Context lifecycle.  One context = one GPU, one geometry at a time. */
int mstm_hip_create(mstm_hip_ctx **ctx, int device);
int mstm_hip_destroy(mstm_hip_ctx *ctx);
const char *mstm_hip_last_error(const mstm_hip_ctx *ctx);

/* Sphere set + Lorenz-Mie coefficients and truncation orders computed on the
 * device.  Replaces miecoefcalc(mie_epsilon) (mstm-modules-33.f90:5301-5407)
 * for external spheres in a real medium.  pos[3n] (x,y,z per sphere),
 * radius[n] (size parameters), ri[2n] (re, im), ri_medium real. */
int mstm_hip_set_spheres(mstm_hip_ctx *ctx, int nsphere, const double *pos, const double *radius, const double *ri,
                         double ri_medium, double mie_epsilon);

/* Override orders and Mie matrices with the host's (sphere_order, an_mie
 * after miecoefcalc, mstm-modules-33.f90:5393-5404): the drop-in path when the
 * Fortran host keeps computing its Mie data. */
int mstm_hip_set_mie(mstm_hip_ctx *ctx, const int *order, const double *an_mie);

/* sphere_order(n), sphere_offset(n+1) (0-based offsets), number_eqns. */
int mstm_hip_get_orders(const mstm_hip_ctx *ctx, int *order, int *offset);
int mstm_hip_number_eqns(const mstm_hip_ctx *ctx);
/* an_mie in the reference packing (4 complex per degree per sphere). */
int mstm_hip_get_mie(const mstm_hip_ctx *ctx, double *an_mie);

/* Build the pair translation tables of the direct operator on the device
 * (the stored form of external_to_external_expansion,
 * mstm-modules-33.f90:6437-6462 with store_translation_matrix).  precision:
 * MSTM_HIP_FP64 or MSTM_HIP_FP32; both may coexist.  MSTM_HIP_FP64_FP32
 * builds both, in one pass where the orders allow (each pair's FP64 values
 * worked out once, stored at both precisions; the same tables as two calls). */
int mstm_hip_setup_direct(mstm_hip_ctx *ctx, int precision);

/* Node grid of the FFT operator: node_selection(fva, target_min, target_max,
 * d_specified, local_host) (fft_translation-5.f90:654-825) followed by the
 * node_order rule of main_calling_program (mstm-input-37.f90:1238-1262).
 *   fva          in/out: <= 0 selects the volume fraction automatically
 *                (clamped to [0.02, 1]); set to the fraction used unless
 *                d_specified
 *   target_min/_max  the target box (3 each); NULL: the bounds of the sphere
 *                centres (sphere_min/max_position)
 *   d_specified, d_cell  d_specified != 0 takes *d_cell as the cell size;
 *                *d_cell returns the final cell size (d_cell of the module)
 *   local_host   must be 0 (spheres inside a host sphere are not on this path)
 *   node_order   <= 0 gives -node_order + ceiling(d_cell)
 *   neighbor_node_model  0..3 (stencil of 1, 7, 19 or 27 cells)
 * Outputs cell_dim[3] and the node order.  Replaces any earlier grid. */
int mstm_hip_node_selection(mstm_hip_ctx *ctx, double *fva, const double *target_min, const double *target_max,
                            int d_specified, double *d_cell, int local_host, int node_order, int neighbor_node_model,
                            int *cell_dim, int *node_order_out);
/* The private grid state node_selection keeps (fft_translation-5.f90:22-27):
 * cell_boundary[3], sphere_node[3 nsphere] (1-based cell indices per sphere)
 * and number_neighbor_nodes; each output may be NULL.  A Fortran host can
 * compare it with its own grid. */
int mstm_hip_fft_grid(const mstm_hip_ctx *ctx, double *cell_boundary, int *sphere_node, int *number_neighbor_nodes);
/* Build the FFT operator (fft_translation_initialization, :901-974, plus the
 * sphere<->node and near-field tables) at the given precision. */
int mstm_hip_setup_fft(mstm_hip_ctx *ctx, int precision);

/* aout(:,r) = A_r ain(:,r): sphereinteraction (mstm-scatprops-26.f90:25-180)
 * for external spheres.  ain/aout: (number_eqns, nrhs) complex, host memory.
 * con_tran[r] != 0 selects conj(H^T T^T conj(.)), mie_mult[r] != 0 applies
 * the T-matrix.  use_fft selects fft_external_to_external_expansion
 * (fft_translation-5.f90:72-336) instead of the direct pair sum.
 * rhs_list (NULL = all) and skip_external_translation as the reference's
 * optional arguments: a column left out of the list, or every column when
 * skip_external_translation != 0, gets no translation, i.e. a zero column
 * (external spheres have no host-sphere or plane-boundary part). */
int mstm_hip_sphereinteraction(mstm_hip_ctx *ctx, int nrhs, const double *ain, double *aout, const int *con_tran,
                               const int *mie_mult, int use_fft, int precision, const int *rhs_list,
                               int skip_external_translation);

/* The product of mstm_hip_sphereinteraction on the direct operator without
 * any device allocation.  No counterpart in MSTM: for callers that cannot
 * allocate device memory (serving loops, code under stream capture, harnesses
 * that forbid hipMalloc).  Same arguments and results as
 * mstm_hip_sphereinteraction with use_fft = 0 (use_fft != 0 is
 * MSTM_HIP_EUNSUP); the pair tables of `precision` must already exist
 * (mstm_hip_setup_direct, which allocates them, else MSTM_HIP_ESTATE).
 * Staging, padded input and partial sums live in one 32 MiB array per GPU,
 * held from the library's load to the end of the process; a cluster whose
 * product needs more is MSTM_HIP_EARG (about 570 spheres of order 4 at FP32,
 * 400 at FP64, four columns).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (several threads or contexts) race. */
int mstm_hip_sphereinteraction_noalloc(mstm_hip_ctx *ctx, int nrhs, const double *ain, double *aout,
                                       const int *con_tran, const int *mie_mult, int use_fft, int precision,
                                       const int *rhs_list, int skip_external_translation);

/* cbicg (mstm-solver-8.f90:654-949): solve (I - T H) anp = pnp with the
 * reference's complex biconjugate gradient, FP64, starting from anp.
 * Stops when |r|^2/|b|^2 < eps.  iter: iterations done (niter+1 when not
 * converged, as the Fortran do-variable), errmax: last |r|^2/|b|^2.
 * niter < 0: the order-of-scattering series (:785-825), anp += c_k with
 * c_1 = T H anp, c_{k+1} = T H c_k, until |c_k|^2/|pnp|^2 < eps (errmax =
 * that ratio, iter = k) or -niter terms (iter = -niter + 1, errmax = 0).
 * niter = 0 (:768-781): anp becomes the host-sphere part of the product,
 * zero for external spheres.  pnp = 0 returns with anp unchanged. */
int mstm_hip_cbicg(mstm_hip_ctx *ctx, int niter, double eps, const double *pnp, double *anp, int use_fft, int *iter,
                   double *errmax);

/* fixedorsoln (mstm-solver-8.f90:350-569): plane-wave RHS for incidence
 * (alpha, sin beta, dir = 1 from below / 2 from above), both polarizations,
 * solve, and the efficiencies of qefficiencyfactors
 * (mstm-scatprops-26.f90:782-870).  amnp(number_eqns,2) complex,
 * qeff(3,3,nsphere) (may be NULL).  As in the reference, maxerr/maxiter/istat
 * report the second polarization's solve.  solver: MSTM_HIP_SOLVE_*.
 * niter = 0: no solve, amnp = T f (:513-522); niter < 0: cbicg's
 * order-of-scattering series from amnp = T f for every solver but
 * MSTM_HIP_SOLVE_LU (istat is then 1, as in the reference).
 * excited_spheres (NULL = all): nsphere flags; a sphere with flag 0 gets no
 * incident field (sphereplanewavecoef, mstm-scatprops-26.f90:217-284). */
int mstm_hip_fixedorsoln(mstm_hip_ctx *ctx, double alpha, double sinc, int dir, double eps, int niter, int solver,
                         int use_fft, double *amnp, double *qeff, double *maxerr, int *maxiter, int *istat,
                         const int *excited_spheres);

/* Inner solver of the mixed refinement (MSTM_HIP_SOLVE_MIXED / _LEARNED; no
 * counterpart in MSTM): 0 = BiCG (the reference's recurrences in FP32),
 * 1 = GMRES (one product column per system and step; the default). */
int mstm_hip_set_inner_solver(mstm_hip_ctx *ctx, int kind);

/* Largest number of FP64 refinement cycles of MSTM_HIP_SOLVE_MIXED and
 * _LEARNED (no counterpart in MSTM; default 6, one normally suffices).  A
 * polarization still above eps after them is handed to the plain FP64
 * solver from b; 0 hands it over at once, after the first FP64 residual. */
int mstm_hip_set_refinement_cycles(mstm_hip_ctx *ctx, int cycles);

/* BiCG iterations of small clusters (up to 500 spheres, direct operator) are
 * replayed four at a time from a HIP graph captured once per set of buffers
 * (no counterpart in MSTM).  on = 0 issues the same launches one by one, with
 * identical results: for tools that do not follow graph launches.  Default 1;
 * a context whose stream cannot be captured switches to 0 by itself. */
int mstm_hip_set_graphs(mstm_hip_ctx *ctx, int on);

/* ---------------------------------------------------------------------------
 * Post-solve observables of a single-origin expansion (a cluster without
 * plane boundaries, single_origin_expansion = .true., the reference default;
 * mstm-input-37.f90:1159-1197 and 1503-1535).  All sums are FP64 on the GPU.
 * ------------------------------------------------------------------------- */

/* Translation orders of the spheres about the origin r0 and the common-origin
 * order T, as main_calling_program sets them (mstm-input-37.f90:1159-1197)
 * with tranordertest (mstm-modules-33.f90:1896-1942):
 *   torder_i = min(tranordertest(|r_i - r0|, m_medium, L_i, eps), maxt),
 *   T = max(min(max_mie_order, maxt), max_i torder_i).
 * origin NULL: the mean sphere position (cluster_origin).  eps is
 * translation_epsilon (reference default 1e-5), maxt max_t_matrix_order
 * (default 100; at most 127 here).  torder (nsphere ints) and tmo may be
 * NULL.  The orders are kept for mstm_hip_merge_to_common_origin. */
int mstm_hip_translation_orders(mstm_hip_ctx *ctx, const double *origin, double eps, int maxt, int *torder,
                                int *tmo);

/* The same state from a host that computed translation_order, t_matrix_order
 * and cluster_origin itself (mstm-input-37.f90:1159-1197): torder holds
 * nsphere orders (1..127), T the common-origin order (1..127). */
int mstm_hip_set_translation_orders(mstm_hip_ctx *ctx, const double *origin, const int *torder, int T);

/* merge_to_common_origin (mstm-scatprops-26.f90:447-559) of both columns of
 * a solution: each sphere's expansion is carried to r0 by the regular
 * rotate-axial-rotate translation of order min(T, torder_i) and summed.
 * amnp: neqns x 2 complex (host), or NULL for the solution the last
 * mstm_hip_fixedorsoln left on the device.  amnp0 (may be NULL): 2 columns of
 * 2 T(T+2) complex in the reference's E/M layout (0:T+1, T, 2) per column
 * (a0(.,.,1) = aL + aR, a0(.,.,2) = aL - aR).  The merged coefficients stay
 * on the device for the two calls below. */
int mstm_hip_merge_to_common_origin(mstm_hip_ctx *ctx, const double *amnp, double *amnp0);

/* scatteringmatrix (mstm-scatprops-26.f90:1498-1573; rotate_plane = .false.,
 * normalize_s11 = .false.) at nang directions (cos theta, phi), i.e. the loop
 * of scattering_matrix_calculation (mstm-input-37.f90:3036-3062).  amnp0 NULL:
 * the last merge (T ignored); else the caller's common-origin coefficients
 * of order T in the layout above.  sm: nang x 16, the Fortran sm(4,4)
 * (column-major) of each direction. */
int mstm_hip_scattering_matrix(mstm_hip_ctx *ctx, int T, const double *amnp0, int nang, const double *ct,
                               const double *phi, double *sm);

/* s00(1,1,0) and s00(1,1,1) of fosmexpansion (mstm-scatprops-26.f90:1820-1951)
 * and the asymmetry parameter g = s00(1,1,1) / (3 s00(1,1,0)) (the first
 * Legendre moment of S11, mstm-scatprops-26.f90:1669).  4 pi s00(1,1,0) is the
 * scattering cross section (k^-2 units).  amnp0/T as for the scattering
 * matrix.  s00 (2 doubles) or g may be NULL. */
int mstm_hip_asymmetry(mstm_hip_ctx *ctx, int T, const double *amnp0, double *s00, double *g);

/* fosmexpansion (mstm-scatprops-26.f90:1820-1951): the generalized-spherical
 * expansion coefficients s00, s02, sp22, sm22 of the scattering matrix, each
 * 4 x 4 x (2T+1) doubles in the Fortran order of s(4,4,0:2T) (element
 * (i,j,w) at (w*4 + j-1)*4 + i-1), with the reference's sign patch and
 * 1/(2 m_medium) scale.  amnp0 NULL: the last merge, whose order must be T;
 * else the caller's common-origin coefficients of order T. */
int mstm_hip_fosmexpansion(mstm_hip_ctx *ctx, int T, const double *amnp0, double *s00, double *s02, double *sp22,
                           double *sm22);

/* direct_solver (mstm-solver-8.f90:571-642): anp = A^{-1} pnp with the dense
 * interaction matrix A = I - T H (general_interaction_matrix with mie_mult,
 * mstm-modules-33.f90:5854-5985, built from the FP64 operator) and its LU
 * factors (lu_decomposition's scaled partial pivoting, :951-1020), made when
 * initialize != 0 or none are held (they stay until the spheres change or
 * mstm_hip_clear(ctx, 1)); then niter refinement steps r = A x - p,
 * x -= LU^{-1} r.  solution_error (may be NULL): sqrt(sum |r|^2) / neqns of
 * the last step, 0 when niter = 0.  Up to 16384 equations. */
int mstm_hip_direct_solver(mstm_hip_ctx *ctx, const double *pnp, double *anp, int initialize, int niter,
                           double *solution_error);

/* tmatrix_solution (mstm-solver-8.f90:22-332), iterative, one GPU: the
 * cluster T-matrix about the origin of the last translation-order call
 * (mstm_hip_translation_orders / _set_translation_orders; T = its
 * t_matrix_order).  For l = 1..T and each column (k, q), k = -l..l, q = 1, 2,
 * the unit TE/TM excitation of degree l at the origin (:169-175) is
 * distributed to the spheres (distribute_from_common_origin,
 * mstm-scatprops-26.f90:561-669, on the GPU), multiplied by the Mie matrices,
 * solved (solver MSTM_HIP_SOLVE_PLAIN: cbicg's FP64 iterates from x = b, the
 * two columns of one k together; MSTM_HIP_SOLVE_MIXED: the mixed solver with
 * its plain-FP64 hand-over) and merged back at order l.  Stops after order l
 * when the change of the orientation-averaged Q_ext is below conveps
 * (conveps > 0, :322).
 *   tmat        complex, the T-matrix file body in the reference's write order
 *               (:243-256): per column, n = 1..l, m = -n..n, p = 1, 2;
 *               capacity tmat_cap complex values >= sum_l 4(2l+1) l(l+2)
 *   sphere_qeff nsphere x 3: per-sphere efficiency sums (ext, abs, sca) over
 *               all columns (sphere_qeff of the reference); may be NULL
 *   qteff       3: the T-matrix sums of :233-237; may be NULL
 *   order_log   T x 4 per order: cumulative max iterations, Q_ext, Q_abs and
 *               the change of Q_ext (the reference's progress lines); may be NULL
 *   t_order     the order reached; maxiter / maxerr / istat as the reference. */
int mstm_hip_tmatrix_solution(mstm_hip_ctx *ctx, double eps, int niter, double conveps, int solver, double *tmat,
                              int64_t tmat_cap, double *sphere_qeff, double *qteff, double *order_log, int *t_order,
                              int *maxiter, double *maxerr, int *istat);

/* clear_fft_matrix (fft_translation-5.f90:42-70) / table release. */
int mstm_hip_clear(mstm_hip_ctx *ctx, int clear_tables);

/* Timing and counters of the last calls (milliseconds, counts); see
 * mstm_hip_stats_names() for the meaning of each slot.  Returns the number
 * of slots written. */
int mstm_hip_stats(const mstm_hip_ctx *ctx, double *out, int nout);
const char *mstm_hip_stats_names(void);

/* ---------------------------------------------------------------------------
 * Learned initial guess.  No counterpart in MSTM: the reference always starts
 * BiCG from x0 = b (mstm-solver-8.f90:483).  A per-sphere FP32 MLP over the
 * neighbour geometry (LIG_NFEAT = 64 features, `nlayers` SiLU layers of
 * width `hidden`, linear output) predicts the solution block of every sphere
 * for incidence along +z:
 *     x0_i = exp(i kappa d_i) (b_i + exp(i k z_i) NN_i)
 * with d_i the depth below the cluster surface and kappa a learned complex
 * wavenumber shift.  Parameters: one flat float array, per layer W [out x in]
 * row-major then bias [out], then kappa (re, im); Lnet is the sphere order
 * the output layer covers (8 Lnet (Lnet+2) outputs, both polarizations).
 * MSTM_HIP_SOLVE_LEARNED checks the prediction's residual against the zero
 * start, runs the mixed solver from the better one (every answer is accepted
 * on its FP64 residual), and hands over to the plain FP64 solver from b if
 * the mixed solve does not reach eps; other incidence directions start from
 * zero.  MSTM_HIP_SOLVE_MIXED starts from zero and hands over the same way.
 * ------------------------------------------------------------------------- */
int mstm_hip_lig_nparams(int hidden, int nlayers, int Lnet);
/* load (params != NULL) or unload the network of the learned solver */
int mstm_hip_set_learned(mstm_hip_ctx *ctx, const float *params, int nparams, int hidden, int nlayers, int Lnet);
/* per-sphere features feat[nsphere][64] and depths depth[nsphere] */
int mstm_hip_lig_features(mstm_hip_ctx *ctx, float *feat, float *depth);
/* training pairs from a solution amnp(number_eqns,2) for incidence along +z:
 * target = a_i exp(-i k z_i), base = b_i exp(-i k z_i) in the order-Lnet
 * layout, rows of 8 Lnet (Lnet+2) floats per sphere */
int mstm_hip_lig_samples(mstm_hip_ctx *ctx, const double *amnp, int Lnet, float *target, float *base);
/* the learned starting vectors x0(number_eqns,2) for incidence along +z */
int mstm_hip_lig_guess(mstm_hip_ctx *ctx, double *x0);

/* Trainer: the network's forward, backward and Adam steps on the GPU. */
typedef struct mstm_hip_trainer mstm_hip_trainer;
int mstm_hip_trainer_create(mstm_hip_trainer **t, int device, int hidden, int nlayers, int Lnet, unsigned seed);
int mstm_hip_trainer_destroy(mstm_hip_trainer *t);
const char *mstm_hip_trainer_last_error(const mstm_hip_trainer *t);
/* get (set = 0) or set (set = 1, also resets the Adam moments) the parameters */
int mstm_hip_trainer_params(mstm_hip_trainer *t, float *params, int nparams, int set);
/* training set: feat[n][64], target/base[n][8 Lnet (Lnet+2)], depth[n] */
int mstm_hip_trainer_set_data(mstm_hip_trainer *t, int nsamp, const float *feat, const float *target,
                              const float *base, const float *depth);
/* one shuffled pass with Adam; mean_loss = mean |yhat - y|^2 / mean |y|^2 */
int mstm_hip_trainer_epoch(mstm_hip_trainer *t, int batch, float lr, unsigned seed, double *mean_loss);
/* loss over the training set without an update */
int mstm_hip_trainer_loss(mstm_hip_trainer *t, int batch, double *loss);
/* gradient of the mean loss over the given samples (kappa last) */
int mstm_hip_trainer_grad(mstm_hip_trainer *t, int nsamp, const float *feat, const float *target, const float *base,
                          const float *depth, float *grad, double *loss);
/* raw network outputs out[n][8 Lnet (Lnet+2)] */
int mstm_hip_trainer_forward(mstm_hip_trainer *t, int nsamp, const float *feat, float *out);

/* ---------------------------------------------------------------------------
 * Batched solve: many independent clusters, each with the same ndir
 * incidence directions x 2 polarizations (R = 2 ndir <= 64 right-hand sides
 * per cluster).  MSTM does one fixedorsoln (mstm-solver-8.f90:350-569) per
 * cluster and direction; here each cluster's H is built once as dense FP64
 * blocks (stored FP64 and FP32) and every BiCG product of all clusters and
 * right-hand sides is one grouped complex GEMM on MFMA (FP32), inside FP64
 * iterative refinement.  Spheres of one cluster must share one order.
 * Host vectors (batch_vectors, batch_product): per cluster an (R, neqns_c)
 * complex array in MSTM layout, clusters concatenated; column 2d+q is
 * direction d, polarization q.
 * ------------------------------------------------------------------------- */
typedef struct mstm_hip_batch mstm_hip_batch;
int mstm_hip_batch_create(mstm_hip_batch **b, int device);
int mstm_hip_batch_destroy(mstm_hip_batch *b);
const char *mstm_hip_batch_last_error(const mstm_hip_batch *b);
/* counts[ncluster]; spheres concatenated: pos[3n], radius[n], ri[2n];
 * directions alpha[ndir], beta[ndir] (radians, beta in [0, pi]) */
int mstm_hip_batch_set(mstm_hip_batch *b, int ncluster, const int *counts, const double *pos, const double *radius,
                       const double *ri, double ri_medium, double mie_eps, int ndir, const double *alpha,
                       const double *beta);
/* sphere order of each cluster (Lorenz-Mie criterion of miecoefcalc) */
int mstm_hip_batch_orders(const mstm_hip_batch *b, int *order);
/* dense operators, right-hand sides b = T f, and (tune != 0) the GEMM autotune */
int mstm_hip_batch_setup(mstm_hip_batch *b, int tune);
/* solve all to |r|^2/|b|^2 < eps; iters/err per (cluster, column) */
int mstm_hip_batch_solve(mstm_hip_batch *b, double eps, int niter, int use_graph, int *iters, double *err);
/* which: 0 solution, 1 right-hand side T f, 2 incident field f */
int mstm_hip_batch_vectors(mstm_hip_batch *b, int which, double *out);
/* y = H x (op 0) or H^H x (op 1); precision MSTM_HIP_FP32 (MFMA) or MSTM_HIP_FP64 (op 0) */
int mstm_hip_batch_product(mstm_hip_batch *b, int precision, int op, const double *x, double *y);
/* qtot per (cluster, direction): q[(c ndir + d) 9 + 3k + j], k unpol/pol1/pol2, j ext/abs/sca */
int mstm_hip_batch_qeff(mstm_hip_batch *b, double *q);
int mstm_hip_batch_stats(const mstm_hip_batch *b, double *out, int nout);
const char *mstm_hip_batch_stats_names(void);
int mstm_hip_batch_tune_table(const mstm_hip_batch *b, double *out, int nout);

#ifdef __cplusplus
}
#endif
#endif
