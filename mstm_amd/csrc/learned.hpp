// learned.hpp -- learned initial guess: per-sphere neighbour features, an
// FP32 multilayer perceptron (inference and training kernels) and the
// mapping between network outputs and MSTM coefficient vectors.
//
// Model.  For plane-wave incidence along +z the solution block of sphere i is
// predicted as
//     x0_i = exp(i kappa d_i) * ( b_i + exp(i k z_i) * NN(f_i) )
// b_i = T_i f_i is the default starting vector of fixedorsoln
// (mstm-solver-8.f90:483), d_i the depth of the sphere below the cluster
// surface along the incidence direction, kappa a learned complex wavenumber
// shift (the coherent-field phase and attenuation per unit depth) and NN a
// per-sphere MLP over the neighbour geometry f_i.  The network output holds
// both polarizations of a sphere of order LNET in MSTM's packed layout,
// 2 x 2 x LNET(LNET+2) complex values.
#pragma once
#include "common.hpp"

#define LIG_NFEAT 64    // features per sphere
#define LIG_NSHELL 4    // radial shells inside the cutoff
#define LIG_SHELLF 14   // features per shell
#define LIG_GLOBALF 6   // features before the shells

// feature/sample kernels
hipError_t launch_lig_features(int nsph, const double *pos, const double *radius, const double *mre,
                               const double *mim, double rimed, float *feat, float *depth, hipStream_t st);
hipError_t launch_lig_samples(int nsph, int pol, const int *order, const int *offset, const double *pos, double rimed,
                              const double *amnp, const double *b, int Lnet, float *target, float *base,
                              hipStream_t st);
hipError_t launch_lig_combine(int nsph, int pol, const int *order, const int *offset, const double *pos, double rimed,
                              const float *out, int Lnet, const float *depth, const float *kappa, const double *b,
                              double *x, hipStream_t st);

// dense layers: C = op(A) op(B) + bias, optional SiLU (Z keeps the pre-activation)
hipError_t launch_gemm(bool ta, bool tb, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                       const float *bias, float *C, int ldc, float *Z, int act, hipStream_t st);
hipError_t launch_silu_bwd(int n, const float *z, float *dh, hipStream_t st);
hipError_t launch_colsum(int rows, int cols, const float *a, float *out, hipStream_t st);
hipError_t launch_lig_loss(int B, int NC, const float *O, const float *Y, const float *base, const float *depth,
                           const float *kappa, float *dO, float inv_norm, double *partials, int nparts,
                           hipStream_t st);
hipError_t launch_lig_loss_finish(int nparts, const double *partials, float gscale, double lscale, float *dkappa,
                                  double *loss_acc, hipStream_t st);
hipError_t launch_adam(int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st);
hipError_t launch_gather_rows(int rows, int cols, const int *idx, const float *src, float *dst, hipStream_t st);
int lig_loss_parts();

// host-side MLP drivers (learned.hip)
struct MlpShape {
    int nin, hidden, nlayers, nout;
};
size_t mlp_nparams(const MlpShape &s);
hipError_t mlp_forward(const MlpShape &s, const float *params, const float *X, int B, float *const *H, float *const *Z,
                       float *out, hipStream_t st);
