/*
 * mstm_oracle.h -- CPU FP64 restatement of the MSTM solution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (mstm_amd/csrc).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Conventions (all follow the reference, file:line relative to the MSTM tree):
 *  - coefficient block of a sphere of order L: complex a(0:L+1,1:L,1:2),
 *    column-major, m>=0 at (m,n), m<0 at (n+1,-m), last index = L/R mode
 *    (december2023/mstm-modules-33.f90:2351-2377, amnpaddress model 2).
 *  - Mie 2x2 L/R matrices an(p,q,n) packed at mie_offset+p+2q+4(n-1)
 *    (mstm-modules-33.f90:5393-5404).
 *  - every complex is C99 double complex (= Fortran complex(8)).
 */
#ifndef MSTM_ORACLE_H
#define MSTM_ORACLE_H
#include <complex.h>
typedef double complex zc;

#ifdef __cplusplus
extern "C" {
#endif

/* special functions / tables */
double orc_bcof(int n, int l);
double orc_cg(int j1, int m1, int j2, int m2, int J);
void orc_rotcoef(double cbe, int kmax, int nmax, double *dc);
void orc_normalizedlegendre(double cbe, int mmax, int nmax, double *dc);
int orc_atcdim(int ntot, int ltot);
int orc_moffset(int m, int ntot, int ltot);
void orc_axialtrancoef(int itype, double r, const zc *ri, int nmax, int lmax, zc *ac);
void orc_gentranmatrix(int nodr_s, int nodr_t, const double *xp, const zc *ri, int itype, zc *ac);
void orc_riccati_psi(int n, double x, double *psi);
void orc_riccati_xi(int n, double x, zc *xi);
void orc_logderiv(int n, double zr, double zi, zc *d);

/* single pair translation (rot path or dense path) */
void orc_pair_translate(int itype, const double *rtran, const zc *ri, int noa, int nog,
                        int sop, int top, int force_rot, const zc *ain, zc *gout);

/* Lorenz-Mie */
int orc_mie_order(double x, double mre, double mim, double mmed, double qeps);
void orc_mie_coef(double x, double mre, double mim, double mmed, int nodr, zc *anp, double *qext,
                  double *qsca);

/* cluster state */
typedef struct orc_cluster orc_cluster;
orc_cluster *orc_cluster_new(int n, const double *pos, const double *radius, const zc *ri,
                             double ri_medium, double mie_eps, int switch_order);
void orc_cluster_free(orc_cluster *c);
int orc_cluster_neqns(const orc_cluster *c);
void orc_cluster_orders(const orc_cluster *c, int *order, int *offset);
void orc_cluster_mie(const orc_cluster *c, zc *an);
void orc_mie_mult(const orc_cluster *c, int idir, int inverse, const zc *ain, zc *aout);
void orc_external_to_external(const orc_cluster *c, int nrhs, const int *contran, const zc *ain,
                              zc *gout);
void orc_sphereinteraction(const orc_cluster *c, int nrhs, const int *contran, const int *miemult,
                           int use_fft, const zc *ain, zc *aout);
void orc_planewave(const orc_cluster *c, double alpha, double sinc, int dir, zc *pmnp);
void orc_planewave_ex(const orc_cluster *c, double alpha, double sinc, int dir, const int *excited, zc *pmnp);
int orc_cbicg(const orc_cluster *c, int use_fft, int niter, double eps, const zc *pnp, zc *anp,
              double *errmax);
void orc_qeff(const orc_cluster *c, const zc *amnp, const zc *pmnp, double *qeff);
void orc_qtot(const orc_cluster *c, const double *qeff, double *qtot);
int orc_fixedorsoln(const orc_cluster *c, int use_fft, double alpha, double sinc, int dir,
                    double eps, int niter, zc *amnp, double *qeff, double *maxerr, int *maxiter);
int orc_fixedorsoln_ex(const orc_cluster *c, int use_fft, double alpha, double sinc, int dir, double eps, int niter,
                       const int *excited, zc *amnp, double *qeff, double *maxerr, int *maxiter);

/* FFT node grid */
int orc_node_selection(orc_cluster *c, double fva, int node_order_in, int neighbor_model,
                       int *cell_dim, double *d_cell, int *node_order);
int orc_node_selection_ex(orc_cluster *c, double *fva_io, const double *target_min, const double *target_max,
                          int d_specified, double *d_cell, int node_order_in, int neighbor_model, int *cell_dim,
                          int *node_order);
void orc_fft_external_to_external(const orc_cluster *c, int nrhs, const int *contran,
                                  const zc *ain, zc *gout);
void orc_node_scatter(const orc_cluster *c, const zc *asphere, zc *anode);
void orc_node_gather(const orc_cluster *c, const zc *anode, zc *asphere);
void orc_node_convolution(const orc_cluster *c, const zc *anode, zc *gnode);
int orc_local_pairs(const orc_cluster *c, int *pairs);

/* post-solve observables (single-origin expansion) */
int orc_tranorder(double r, double rimed, int lmax, double eps);
int orc_translation_orders(const orc_cluster *c, const double *r0, double eps, int maxt, int *torder);
void orc_merge(const orc_cluster *c, int T, const int *torder, const double *r0, int nrhs, const zc *amnp,
               zc *amnp0);
void orc_scatmat(int T, const zc *a0, double ct, double phi, double rimed, double *sm, zc *sa_out);
void orc_fosm(int T, const zc *amnp, double rimed, double *s00, double *s02, double *sp22, double *sm22);

/* T-matrix solution (random-orientation path) */
void orc_distribute(const orc_cluster *c, int l, const double *r0, const zc *a0, zc *amnp);
void orc_ranorient(int T, const zc *tmat, double rimed, double cbeam, double *sm, double *smcf);
void orc_ranorient_coef(int T, const zc *dm, const zc *dmcf, double rimed, double *sm, double *smcf);
int orc_tmatrix(const orc_cluster *c, int T, const int *torder, const double *r0, double vol_radius, double eps,
                int niter, double conveps, zc *tmat, double *sqeff, double *qteff, double *qlog, int *maxiter,
                double *maxerr, int *istat);

#ifdef __cplusplus
}
#endif
#endif
