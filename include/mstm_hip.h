/*
 * mstm_hip.h -- C-ABI of the MI355X implementation of MSTM's solution hot
 * path (libmstm_hip.so).  Plain pointers and sizes only; every complex
 * array is interleaved (re, im) doubles, i.e. Fortran complex(8), laid out
 * exactly as MSTM's coefficient vectors: per sphere a block a(0:L+1,1:L,1:2)
 * at sphere_offset(i) (mstm-modules-33.f90:2351-2377, 5354-5376).
 *
 * Each entry point names the reference procedure whose contract it takes
 * over (paths relative to MSTM's december2023/ tree).  The Fortran shim
 * module (mstm_amd/fortran/mstm_hip_mod.f90) binds all of them with
 * ISO_C_BINDING; INTEGRATION.md shows the reference-side call sites.
 *
 * Return codes: 0 = success, otherwise an MSTM_HIP_E* code and
 * mstm_hip_last_error() holds the message.  Calls on one context must come
 * from one host thread; all device work runs on the context's stream.
 */
#ifndef MSTM_HIP_H
#define MSTM_HIP_H

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

#define MSTM_HIP_SOLVE_PLAIN 0   /* cbicg in FP64, iteration-identical         */
#define MSTM_HIP_SOLVE_MIXED 1   /* FP32 BiCG inside FP64 iterative refinement */
#define MSTM_HIP_SOLVE_LEARNED 2 /* MIXED started from the learned guess       */

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
 * MSTM_HIP_FP64 or MSTM_HIP_FP32; both may coexist. */
int mstm_hip_setup_direct(mstm_hip_ctx *ctx, int precision);

/* Node grid of the FFT operator: node_selection + node_order rule
 * (fft_translation-5.f90:654-825, mstm-input-37.f90:1238-1262).  fva <= 0
 * selects the volume fraction automatically; node_order <= 0 gives
 * -node_order + ceiling(d_cell).  Outputs cell_dim[3], d_cell, node order. */
int mstm_hip_node_selection(mstm_hip_ctx *ctx, double fva, int node_order, int neighbor_node_model, int *cell_dim,
                            double *d_cell, int *node_order_out);
/* Build the FFT operator (fft_translation_initialization, :901-974, plus the
 * sphere<->node and near-field tables) at the given precision. */
int mstm_hip_setup_fft(mstm_hip_ctx *ctx, int precision);

/* aout(:,r) = A_r ain(:,r): sphereinteraction (mstm-scatprops-26.f90:25-180)
 * for external spheres.  ain/aout: (number_eqns, nrhs) complex, host memory.
 * con_tran[r] != 0 selects conj(H^T T^T conj(.)), mie_mult[r] != 0 applies
 * the T-matrix.  use_fft selects fft_external_to_external_expansion
 * (fft_translation-5.f90:72-336) instead of the direct pair sum. */
int mstm_hip_sphereinteraction(mstm_hip_ctx *ctx, int nrhs, const double *ain, double *aout, const int *con_tran,
                               const int *mie_mult, int use_fft, int precision);

/* cbicg (mstm-solver-8.f90:654-949): solve (I - T H) anp = pnp with the
 * reference's complex biconjugate gradient, FP64, starting from anp.
 * Stops when |r|^2/|b|^2 < eps.  iter: iterations done (niter+1 when not
 * converged, as the Fortran do-variable), errmax: last |r|^2/|b|^2. */
int mstm_hip_cbicg(mstm_hip_ctx *ctx, int niter, double eps, const double *pnp, double *anp, int use_fft, int *iter,
                   double *errmax);

/* fixedorsoln (mstm-solver-8.f90:350-569), iterative branch: plane-wave RHS
 * for incidence (alpha, sin beta, dir=1), both polarizations, solve, and the
 * efficiencies of qefficiencyfactors (mstm-scatprops-26.f90:782-870).
 * amnp(number_eqns,2) complex, qeff(3,3,nsphere).  As in the reference,
 * maxerr/maxiter/istat report the second polarization's solve.
 * solver: MSTM_HIP_SOLVE_*. */
int mstm_hip_fixedorsoln(mstm_hip_ctx *ctx, double alpha, double sinc, int dir, double eps, int niter, int solver,
                         int use_fft, double *amnp, double *qeff, double *maxerr, int *maxiter, int *istat);

/* clear_fft_matrix (fft_translation-5.f90:42-70) / table release. */
int mstm_hip_clear(mstm_hip_ctx *ctx, int clear_tables);

/* Timing and counters of the last calls (milliseconds, counts); see
 * mstm_hip_stats_names() for the meaning of each slot.  Returns the number
 * of slots written. */
int mstm_hip_stats(const mstm_hip_ctx *ctx, double *out, int nout);
const char *mstm_hip_stats_names(void);

#ifdef __cplusplus
}
#endif
#endif
