// Device code shared by the ensemble kernels (vort2d_ensemble.hip: one
// launch per pass; vort2d_ensemble_pair.hip: one launch per stage pair).
#pragma once

#include "vort2d_body.h"

namespace ffhip {

struct EnsMember {  // device-resident descriptor of one member
  VortGeom g;
  VortColArgs ca;
  VortRowArgs ra;
};

// points per thread of size n when sizes from t up take 16
constexpr int ens_pts(int n, int t) { return n >= t ? 16 : 8; }
constexpr bool ens_ok(int n, int t, int bd) { return n / ens_pts(n, t) <= bd; }
constexpr int kEnsThresholds[4] = {128, 256, 512, 2048};  // 2048: 8 points at every ensemble size
constexpr int kEnsDefaultThr = 256;                       // the single-grid rule
// per-size mask of a threshold (bit s: size 64 << s takes 16 points)
constexpr int ens_mask(int t) {
  int m = 0;
  for (int s = 0; s < 5; ++s)
    if ((64 << s) >= t) m |= 1 << s;
  return m;
}

template <typename R, int BD, int P, int N, int STAGE>
__device__ __forceinline__ void ens_col_n(const VortGeom& g, const VortColArgs& a, int first, unsigned char* sm) {
  constexpr int T = N / P;
  if constexpr (T <= BD) {
    constexpr int CPB = BD / T;
    const int t = threadIdx.x;
    const int cc = t % CPB;
    const int col = first + cc;
    constexpr int SL = team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), CPB);
    cplx<R>* base = reinterpret_cast<cplx<R>*>(sm);
    const cplx<R>* tw = team_tw_load<N, P, R>(base + CPB * SL, a.tw);
    vcol_body<R, N, P, STAGE>(g, a, col / kVCB, col % kVCB, t / CPB, base + cc * SL, tw, col / kVCB < g.TX);
  }
}

template <typename R, int BD, int P, int N>
__device__ __forceinline__ void ens_row_n(const VortGeom& g, const VortRowArgs& a, int first, unsigned char* sm) {
  constexpr int T = N / P;
  if constexpr (T <= BD) {
    constexpr int RPB = BD / T;
    const int t = threadIdx.x;
    const int r = t % RPB;
    constexpr int SL = team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), RPB);
    cplx<R>* base = reinterpret_cast<cplx<R>*>(sm);
    const cplx<R>* tw = team_tw_load<N, P, R>(base + RPB * SL, a.tw);
    vrow_body<R, N, P>(g, a, first + r, t / RPB, base + r * SL, tw);
  }
}

// The member's size picks the body (uniform over the workgroup: every item
// is one member); TH (threshold) fixes each size's points per thread.
template <typename R, int BD, int TH, int STAGE>
__global__ void __launch_bounds__(BD) k_ens_col(const EnsMember* __restrict__ mem, const int2* __restrict__ items) {
  extern __shared__ unsigned char vsm[];
  const int2 it = items[blockIdx.x];
  const VortGeom g = mem[it.x].g;
  const VortColArgs a = mem[it.x].ca;
  switch (g.ny) {
    case 64: ens_col_n<R, BD, ens_pts(64, TH), 64, STAGE>(g, a, it.y, vsm); break;
    case 128: ens_col_n<R, BD, ens_pts(128, TH), 128, STAGE>(g, a, it.y, vsm); break;
    case 256: ens_col_n<R, BD, ens_pts(256, TH), 256, STAGE>(g, a, it.y, vsm); break;
    case 512: ens_col_n<R, BD, ens_pts(512, TH), 512, STAGE>(g, a, it.y, vsm); break;
    default: ens_col_n<R, BD, ens_pts(1024, TH), 1024, STAGE>(g, a, it.y, vsm); break;
  }
}

template <typename R, int BD, int TH>
__global__ void __launch_bounds__(BD) k_ens_row(const EnsMember* __restrict__ mem, const int2* __restrict__ items) {
  extern __shared__ unsigned char vsm[];
  const int2 it = items[blockIdx.x];
  const VortGeom g = mem[it.x].g;
  const VortRowArgs a = mem[it.x].ra;
  switch (g.nx) {
    case 64: ens_row_n<R, BD, ens_pts(64, TH), 64>(g, a, it.y, vsm); break;
    case 128: ens_row_n<R, BD, ens_pts(128, TH), 128>(g, a, it.y, vsm); break;
    case 256: ens_row_n<R, BD, ens_pts(256, TH), 256>(g, a, it.y, vsm); break;
    case 512: ens_row_n<R, BD, ens_pts(512, TH), 512>(g, a, it.y, vsm); break;
    default: ens_row_n<R, BD, ens_pts(1024, TH), 1024>(g, a, it.y, vsm); break;
  }
}

// ----------------------------------------------------------------- stage pairs
// The row pass and the column stage of one ETDRK4 stage in one launch:
// workgroups 0..nrow-1 take the row items, the others the column items.  A
// member's column items wait until every row item of that member has ended:
// each row workgroup makes its stores visible device-wide and counts itself
// in rows_done[member]; a column workgroup's first thread polls that count
// (bounded: after kEnsSpinLimit rounds it sets *err and goes on, so every
// wavefront ends) and the workgroup then reads with device-scope acquire.
// Workgroups are dispatched in index order on each XCD, so a column
// workgroup can only be resident once every row workgroup of its XCD has
// been dispatched: the rows it waits for always run.  The member's last
// column workgroup resets both counts for the next launch.
struct EnsSync {
  int* rows_done;       // [count] row workgroups of this launch that ended
  int* cols_done;       // [count] column workgroups of this launch that ended
  int* err;             // set when a wait ran out
  const int2* counts;   // [count] (row items, column items) per member
};

constexpr int kEnsSpinLimit = 1 << 20;  // polls (about a second) before a wait gives up

// EnsPairArgs: what one stage-pair launch needs
struct EnsPairArgs {
  const EnsMember* mem;
  const int2* items;  // column items, then row items
  int ncol, nrow;
  EnsSync sy;
  size_t lds;
  double bytes;
};

// launch (or only set the LDS limit of) the stage-pair kernel of this
// precision, workgroup size, threshold and stage
int ens_pair_launch(int prec, int bd, int thr, int stage, const EnsPairArgs& pa, hipStream_t st, bool attr_only);

}  // namespace ffhip
