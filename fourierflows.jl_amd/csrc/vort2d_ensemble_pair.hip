// Stage-pair kernels of the ensembles: the row pass and the column stage of
// one ETDRK4 stage of every member in one launch (four launches per step
// instead of eight).  Synchronisation between a member's row and column
// items is per member (EnsSync, vort2d_ens.h), so one member's column work
// overlaps other members' row work and no workgroup waits at a launch
// boundary.  Kept in its own translation unit: its instantiations compile in
// parallel with the per-pass kernels.
#include "vort2d_ens.h"

namespace ffhip {

template <typename R, int BD, int TH, int STAGE>
__global__ void __launch_bounds__(BD) k_ens_pair(const EnsMember* __restrict__ mem, const int2* __restrict__ items,
                                                 int ncol, int nrow, EnsSync sy) {
  extern __shared__ unsigned char vsm[];
  const int b = blockIdx.x;
  if (b < nrow) {  // row item
    const int2 it = items[ncol + b];
    const VortGeom g = mem[it.x].g;
    const VortRowArgs a = mem[it.x].ra;
    switch (g.nx) {
      case 64: ens_row_n<R, BD, ens_pts(64, TH), 64>(g, a, it.y, vsm); break;
      case 128: ens_row_n<R, BD, ens_pts(128, TH), 128>(g, a, it.y, vsm); break;
      case 256: ens_row_n<R, BD, ens_pts(256, TH), 256>(g, a, it.y, vsm); break;
      case 512: ens_row_n<R, BD, ens_pts(512, TH), 512>(g, a, it.y, vsm); break;
      default: ens_row_n<R, BD, ens_pts(1024, TH), 1024>(g, a, it.y, vsm); break;
    }
    __threadfence();  // this thread's stores are visible device-wide ...
    __syncthreads();  // ... for every thread of the workgroup, before the count
    if (threadIdx.x == 0) atomicAdd(&sy.rows_done[it.x], 1);
    return;
  }
  const int2 it = items[b - nrow];  // column item
  if (threadIdx.x == 0) {
    const int need = sy.counts[it.x].x;
    int n = 0;
    while (__hip_atomic_load(&sy.rows_done[it.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
      if (++n > kEnsSpinLimit) {
        atomicOr(sy.err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every thread sees the member's row outputs
  const VortGeom g = mem[it.x].g;
  const VortColArgs a = mem[it.x].ca;
  switch (g.ny) {
    case 64: ens_col_n<R, BD, ens_pts(64, TH), 64, STAGE>(g, a, it.y, vsm); break;
    case 128: ens_col_n<R, BD, ens_pts(128, TH), 128, STAGE>(g, a, it.y, vsm); break;
    case 256: ens_col_n<R, BD, ens_pts(256, TH), 256, STAGE>(g, a, it.y, vsm); break;
    case 512: ens_col_n<R, BD, ens_pts(512, TH), 512, STAGE>(g, a, it.y, vsm); break;
    default: ens_col_n<R, BD, ens_pts(1024, TH), 1024, STAGE>(g, a, it.y, vsm); break;
  }
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&sy.cols_done[it.x], 1) == sy.counts[it.x].y - 1) {
    // the member's last column item: every row and column workgroup of it
    // has passed, so the counts can start over for the next launch
    atomicExch(&sy.rows_done[it.x], 0);
    atomicExch(&sy.cols_done[it.x], 0);
  }
}

template <typename R, int BD, int TH, int STAGE>
static int pair_launch(const EnsPairArgs& pa, hipStream_t st, bool attr_only) {
  auto k = k_ens_pair<R, BD, TH, STAGE>;
  if (attr_only) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)pa.lds));
    return FFHIP_OK;
  }
  static const char* names[4] = {"ens_pair_stage1", "ens_pair_stage2", "ens_pair_stage3", "ens_pair_stage4"};
  prof_launch(names[STAGE - 1], pa.bytes, k, dim3(pa.nrow + pa.ncol), dim3(BD), pa.lds, st, pa.mem, pa.items,
              pa.ncol, pa.nrow, pa.sy);
  FF_LAUNCHED("k_ens_pair");
  return FFHIP_OK;
}

template <typename R, int BD, int TH>
static int pair_stage(int stage, const EnsPairArgs& pa, hipStream_t st, bool attr_only) {
  switch (stage) {
    case 1: return pair_launch<R, BD, TH, 1>(pa, st, attr_only);
    case 2: return pair_launch<R, BD, TH, 2>(pa, st, attr_only);
    case 3: return pair_launch<R, BD, TH, 3>(pa, st, attr_only);
    default: return pair_launch<R, BD, TH, 4>(pa, st, attr_only);
  }
}

template <typename R, int BD>
static int pair_thr(int thr, int stage, const EnsPairArgs& pa, hipStream_t st, bool attr_only) {
  switch (thr) {
    case 256: return pair_stage<R, BD, 256>(stage, pa, st, attr_only);
    case 2048: return pair_stage<R, BD, 2048>(stage, pa, st, attr_only);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "stage-pair kernels: no threshold %d", thr);
  }
}

int ens_pair_launch(int prec, int bd, int thr, int stage, const EnsPairArgs& pa, hipStream_t st, bool attr_only) {
  FF_CHECK(stage >= 1 && stage <= 4, FFHIP_ERR_ARG, "stage-pair kernels: stage %d", stage);
  if (prec == FFHIP_F32)
    return bd == 128 ? pair_thr<float, 128>(thr, stage, pa, st, attr_only)
                     : pair_thr<float, 256>(thr, stage, pa, st, attr_only);
  return bd == 128 ? pair_thr<double, 128>(thr, stage, pa, st, attr_only)
                   : pair_thr<double, 256>(thr, stage, pa, st, attr_only);
}

}  // namespace ffhip
