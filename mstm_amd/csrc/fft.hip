// fft.hip -- FFT node-grid operator (placeholder until the node kernels land).
#include "fft.hpp"
struct FftOp {
    int dummy;
};
FftOp *fft_create(const FftGeometry &, double, int, int, int *, double *, int *, int *err)
{
    *err = 4;
    return nullptr;
}
int fft_setup(FftOp *, int, hipStream_t) { return 4; }
int fft_matvec(FftOp *, int, const double *, double *, const int *, int, hipStream_t) { return 4; }
void fft_clear(FftOp *) {}
void fft_destroy(FftOp *f) { delete f; }
const char *fft_last_error(const FftOp *) { return "FFT operator not built"; }
double fft_stat(const FftOp *, int) { return 0.0; }
int fft_required_vcc(const FftOp *) { return 0; }
