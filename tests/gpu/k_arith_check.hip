// Test library (not part of the product): element-wise FP64 operations as the
// device code compiles them (same flags as libkoral_hip.so), so a GPU test can
// compare them bit for bit with the host's IEEE results: square root,
// division and the portable exp/log/pow of kd_math.h.
#include <hip/hip_runtime.h>
#include "kd_math.h"

namespace {

__global__ void k_arith(int op, long n, const double *__restrict__ x, const double *__restrict__ y,
                        double *__restrict__ out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i], b = y[i];
  double r = 0.;
  switch (op) {
    case 0: r = sqrt(a); break;
    case 1: r = a / b; break;
    case 2: r = kd_exp(a); break;
    case 3: r = kd_log(a); break;
    case 4: r = kd_pow(a, b); break;
    case 5: r = 1. / a; break;
    default: r = 0.; break;
  }
  out[i] = r;
}

}  // namespace

extern "C" int arith_check(int op, long n, const double *x, const double *y, double *out, void *stream)
{
  if (n <= 0) return 0;
  k_arith<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(op, n, x, y, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
