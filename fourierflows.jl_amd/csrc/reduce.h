// Double-double (compensated FP64) reduction helpers shared by the Parseval
// sums and the operators' diagnostics.
#pragma once

#include "common.h"

namespace ffhip {

// Double-double accumulation: error-free transforms must not be contracted,
// hence the function-scoped pragmas.
__device__ __forceinline__ double2 dd_add(double2 a, double2 b) {
#pragma clang fp contract(off)
  const double s = a.x + b.x;
  const double bb = s - a.x;
  double e = (a.x - (s - bb)) + (b.x - bb);
  e += a.y + b.y;
  const double hi = s + e;
  return make_double2(hi, e - (hi - s));
}

__device__ __forceinline__ double2 dd_sq(double x) {  // x*x exactly as hi + lo
#pragma clang fp contract(off)
  const double p = x * x;
  return make_double2(p, fma(x, x, -p));
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double2 block_dd_sum(double2 v) {
  __shared__ double2 sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    double2 o;
    o.x = __shfl_down(v.x, off, 64);
    o.y = __shfl_down(v.y, off, 64);
    v = dd_add(v, o);
  }
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kRedThreads / 64; ++i) v = dd_add(v, sh[i]);
  }
  return v;
}


int parseval_launch(int mode, const void* uh, int prec, int ndim, const int64_t* dims, int rfft, int64_t nx,
                    double norm, double* out, hipStream_t st);
int red_buffer(double2** p);
int red_scratch_setup(int device);  // device address of the allocation-free reduction's scratch array

}  // namespace ffhip
