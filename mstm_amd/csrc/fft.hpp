// fft.hpp -- FFT-accelerated sphere interaction (node grid) interface.
//
// MI355X form of fft_translation (fft_translation-5.f90:1-1076): spheres
// are J-translated to the centre of their node cell, node-to-node H is a
// 3-D convolution evaluated with zero-padded FFTs against a precomputed
// node-kernel spectrum, node fields are J-translated back, and sphere pairs
// in neighbouring cells are translated directly.
#pragma once
#include <hip/hip_runtime.h>

struct FftOp;
struct FftGeometry {
    int nsph, Lmax, neqns;
    const double *pos;     // host, 3*nsph
    const double *radius;  // host
    const int *order, *offset;
    double rimed;
    // device arrays owned by the context
    const double *d_pos;
    const int *d_order, *d_offset;
    const double *d_an;
    const double *d_vcc;
    int vcc_LM;
};
// node_selection: fva in/out (<= 0: automatic), target box (null: sphere
// centre bounds), d_cell in (d_specified) / out
FftOp *fft_create(const FftGeometry &g, double *fva, const double *target_min, const double *target_max,
                  int d_specified, int node_order, int neighbor_model, int *cell_dim, double *d_cell,
                  int *node_order_out, int *err, const char **why);
// the grid of node_selection: cell_boundary[3] (the target box minimum),
// sphere_node[3 nsph] (1-based cell indices), number of stencil cells - 1
// (number_neighbor_nodes) and near-field pairs
void fft_grid(const FftOp *f, double *cell_boundary, int *sphere_node, int *number_neighbor_nodes, int *npairs);
// (re)binds the context's device pointers (vcc may have been regrown)
void fft_bind(FftOp *f, const FftGeometry &g);
int fft_setup(FftOp *f, int prec, hipStream_t st);
int fft_matvec(FftOp *f, int nrhs, const double *ain, double *aout, const int *flags, int prec, hipStream_t st);
void fft_clear(FftOp *f);
void fft_destroy(FftOp *f);
const char *fft_last_error(const FftOp *f);
// 0 setup ms, 1 convolution ms, 2 products, 3/4 FP32 contraction ms and
// launches (HIP events), 5/6 the same for FP64
double fft_stat(FftOp *f, int which);
void fft_reset_stats(FftOp *f);
int fft_required_vcc(const FftOp *f);
bool fft_ready(const FftOp *f, int prec);
