// Built-in operator object shared by the generic and fused paths.
#pragma once

#include <cstdint>

#include "plan.h"
#include "stepper.h"

struct ffhip_op {
  int kind = 0;
  int prec = FFHIP_F64;
  int ndim = 2;
  int64_t n[3] = {1, 1, 1};   // nx, ny, nz
  double box[3] = {1, 1, 1};  // Lx, Ly, Lz
  double af = 1.0 / 3.0;      // aliased_fraction
  double params[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int nparams = 0;
  int64_t nkr = 1;
  int64_t spec_count = 0, phys_count = 0;
  int64_t lo[3] = {1, 1, 1}, hi[3] = {0, 0, 0};  // 1-based aliased ranges on the rfft layout
  ffhip_plan* plan = nullptr;
  void* spec[3] = {nullptr, nullptr, nullptr};
  void* phys[4] = {nullptr, nullptr, nullptr, nullptr};
  void* tables = nullptr;  // tracer: sin/cos of the grid coordinates (FP64)
  void* fused = nullptr;   // fused implementation state (power-of-two vorticity grids)
  void* fused3 = nullptr;  // fused RK4 state of the tracer (power-of-two 3D grids)
};

namespace ffhip {

int op_calcN(ffhip_op* op, void* N, const void* sol, hipStream_t st);
// fused path (vort2d_fused.hip): create sets op->fused when the grid qualifies
int fused_create(ffhip_op* op);
void fused_destroy(ffhip_op* op);
int fused_etdrk4_run(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, hipStream_t st);
// fused tracer RK4 (tracer3d_fused.hip)
int tracer_fused_create(ffhip_op* op);
void tracer_fused_destroy(ffhip_op* op);
int tracer_fused_rk4_run(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, hipStream_t st);
int tracer_fused_rk4_adjoint(ffhip_op* op, ffhip_rk4* ts, void* lam, double dt, int64_t nsteps, hipStream_t st);
// discrete adjoint (vort2d_adjoint.hip)
int fused_forward_ckpt(ffhip_op* op, ffhip_etdrk4* ts, const void* sol0, void* solT, int64_t nsteps,
                       int64_t ncheck, hipStream_t st);
int fused_adjoint(ffhip_op* op, ffhip_etdrk4* ts, const void* lamT, void* lam0, int64_t nsteps, int64_t ncheck,
                  hipStream_t st);
int64_t fused_checkpoint_id(const ffhip_op* op);

}  // namespace ffhip
