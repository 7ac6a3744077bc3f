// Time-stepper objects (host side) shared by the pointwise kernels and the
// fused operators.
#pragma once

#include <cstdint>

#include "common.h"

struct ffhip_etdrk4 {
  int ndim = 1;
  int64_t dims[3] = {1, 1, 1};
  int64_t count = 0;  // prod(dims)
  int prec = FFHIP_F64;
  int cplx_coef = 0;  // coefficients complex (L complex)
  double dt = 0;      // dt rounded to the state precision (src/timesteppers.jl:457)
  // FP64 coefficient arrays: zeta, alpha, beta, Gamma, expLdt, exp_half_Ldt
  void* coef[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // state-precision scratch: sol1, sol2, N1, N2, N3, N4
  void* scratch[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // operator-private compacted copies of the coefficients (fused paths)
  void* packed = nullptr;
  const void* packed_owner = nullptr;
  uint64_t packed_key = 0;  // geometry the packed copy was made for
  // host copy of the linear operator (FP64) for the operators that need it
  double* L_host_re = nullptr;
  double* L_host_im = nullptr;
  void* filter = nullptr;  // FilteredETDRK4: FP64 real spectral filter, or null
};

struct ffhip_rk4 {
  int ndim = 1;
  int64_t dims[3] = {1, 1, 1};
  int64_t count = 0;
  int prec = FFHIP_F64;
  int cplx_coef = 0;
  void* L = nullptr;  // FP64 real or complex
  void* scratch[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // sol1, RHS1..RHS4
  void* filter = nullptr;  // Filtered* explicit steppers: FP64 real spectral filter, or null
};

namespace ffhip {

// Pointwise launches used by both the C ABI and the operators' generic paths.
int etd_substep12(const ffhip_etdrk4* ts, void* out, const void* sol, const void* N, hipStream_t st);
int etd_substep3(const ffhip_etdrk4* ts, void* out, const void* sol1, const void* N1, const void* N3,
                 hipStream_t st);
int etd_update(const ffhip_etdrk4* ts, void* sol, const void* N1, const void* N2, const void* N3, const void* N4,
               hipStream_t st);
int rk4_addlinear(const ffhip_rk4* ts, void* RHS, const void* sol, hipStream_t st);
int rk4_substepsol(const ffhip_rk4* ts, void* newsol, const void* sol, const void* RHS, double dt, hipStream_t st);
int rk4_update(const ffhip_rk4* ts, void* sol, const void* R1, const void* R2, const void* R3, const void* R4,
               double dt, hipStream_t st);
// explicit steppers sharing the RK4 object (L, filter, scratch)
int apply_filter(int prec, int64_t count, const void* filter, void* sol, hipStream_t st);
int euler_update(const ffhip_rk4* ts, void* sol, const void* N, double dt, hipStream_t st);
int ab3_update(const ffhip_rk4* ts, void* sol, const void* RHS, const void* RHSm1, const void* RHSm2, double dt,
               int64_t step, hipStream_t st);
int lsrk54_stage(const ffhip_rk4* ts, void* sol, void* S2, const void* RHS, double A, double B, double dt,
                 hipStream_t st);
int etd_scratch(ffhip_etdrk4* t);
int rk4_scratch(ffhip_rk4* t);
int dealias_launch(void* fh, int prec, int ndim, const int64_t* dims, int64_t batch, const int64_t* lo,
                   const int64_t* hi, hipStream_t st);

}  // namespace ffhip
