// Host-side FFT plan objects shared by the transform code and the operators.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace ffhip {

// One transformed axis: complex FFT length, radix schedule, twiddle table
// exp(-2 pi i k/n) (k < n) in the plan precision, and the execution path.
struct Axis {
  int64_t n = 1;
  std::vector<int> rad;
  void* tw = nullptr;
  bool lds = false;  // whole transform in one LDS tile
  int nb = 1;        // transforms (rows) or columns per block on the LDS path
  bool fast = false; // power-of-two fast kernels (fft_fast.hip) run this axis
};

void unit_root(int64_t k, int64_t n, double* c, double* s);
std::vector<int> factor_radices(int64_t n);
int upload_roots(int64_t n, int64_t count, int prec, void** dptr);

// fast kernels of power-of-two axes (fft_fast.hip); FFHIP_FFT_FAST=0 turns them off
bool fast_rows_ok(int64_t m);  // r2c/c2r rows of half length m
bool fast_cols_ok(int64_t n);  // c2c along a strided axis of length n
int fast_rows_r2c(int prec, const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw,
                  hipStream_t st);
int fast_rows_c2r(int prec, const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, double scale,
                  hipStream_t st);
int fast_cols_c2c(int prec, const void* in, void* out, int64_t ncols, int64_t stride, int64_t nslabs,
                  int64_t slab_dist, const Axis& ax, int dir, double scale, hipStream_t st);

}  // namespace ffhip

struct ffhip_plan {
  int kind = 0;  // 0 rfft, 1 fft
  int ndim = 1;
  int64_t n[3] = {1, 1, 1};
  int64_t batch = 1;
  int prec = FFHIP_F64;
  ffhip::Axis ax[3];  // ax[0] is x; for rfft its length is nx/2
  void* rtw = nullptr;  // rfft split twiddles exp(-2 pi i k/nx), k <= nx/2
  void* ws = nullptr;   // spectral-size workspace (inverse keeps its input)
  void* g0 = nullptr;   // global-path ping-pong buffers
  void* g1 = nullptr;
  void* g2 = nullptr;
  void* adj = nullptr;  // spectral-size buffer of the transposed inverse (allocated on first use)
};

namespace ffhip {
// ldiv! with an extra factor: out = extra * ldiv!(in) (extra = prod(n) gives the
// unnormalised inverse); the input is preserved
int plan_inverse_scaled(ffhip_plan* p, const void* in, void* out, double extra, hipStream_t st);
}  // namespace ffhip

