// Host build of the AO kernels' e^x (qmckl_amd/csrc/exp_neg.hpp) for the accuracy test:
// the device code computes the same bits, every step being an exactly rounded FP64 operation.
#include "../qmckl_amd/csrc/exp_neg.hpp"

extern "C" void exp_neg_host(const double* x, double* y, long n, int deg) {
  for (long i = 0; i < n; ++i) y[i] = deg == 13 ? qmckl_amd::exp_neg(x[i]) : qmckl_amd::exp_neg_f32(x[i]);
}
