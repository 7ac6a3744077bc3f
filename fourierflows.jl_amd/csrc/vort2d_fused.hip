// Fused ETDRK4 step of 2D vorticity on power-of-two grids.
//
// One ETDRK4 step is 4 stages; each stage is one row kernel and one column
// kernel, 8 launches per step (the reference issues ~100: 4 calcN! of 3
// ldiv!, 2 mul!, ~6 broadcasts and 2 dealias! each, plus 3 stage updates).
//
//   column kernel (y):  forward y-FFT of (u zeta)^ and (v zeta)^
//                       -> N = -i kr (u zeta)^ - i l (v zeta)^ on the kept band
//                       -> the stage combination of src/timesteppers.jl:501-516
//                          in FP64 with FP64 coefficients
//                       -> inverse y-FFTs of zeta^, i l/|k|^2 zeta^, -i kr/|k|^2 zeta^
//   row kernel (x):     c2r of (u + i v) and of zeta (complex-packed), products
//                       u zeta, v zeta, one forward x-FFT of (u zeta + i v zeta)
//                       and the even/odd split back to two half spectra
//
// Only the kept (dealiased) band is ever read or written: KX = nkr - |kralias|
// columns and KY = ny - |lalias| rows (2/3 x 2/3 of the modes at the nominal
// aliased_fraction).  Work arrays:
//   intermediates  [TX][ny][CB]  (column blocks of CB kx, all y; complex)
//   spectral state [TX][KY][CB]  (same blocks, kept rows only)
// so the column kernel streams contiguous blocks and the row kernel reads
// CB x RR tiles of 128 B (FP32).  The user-visible state keeps the
// reference's (nkr, ny) layout; it is gathered into the blocked layout at
// the start of a run and scattered back at the end, where modes outside the
// kept band get exp(dt L)^nsteps exactly as the unfused stepper gives them.
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "vort2d_body.h"

namespace ffhip {


#ifndef FF_COL_PAIR
#define FF_COL_PAIR 1  // FP32 column pass up to 1024: paired transforms (LDS has room there)
#endif
template <typename R, int N>
constexpr bool col_pair() {
  return sizeof(R) == 4 && N <= 1024 && FF_COL_PAIR;
}

#ifndef FF_ROW_PAIR
#define FF_ROW_PAIR 1  // FP32 single-grid row pass: the two inverse transforms side by side
#endif
template <typename R>
constexpr bool row_pair() {
  return sizeof(R) == 4 && FF_ROW_PAIR;
}

#ifndef FF_ROW_WAVES
#define FF_ROW_WAVES 1  // minimum FP32 waves per SIMD the single-team row pass is compiled for
#endif
#ifndef FF_COL_WAVES
#define FF_COL_WAVES 1  // same for the single-team column pass
#endif
#ifndef FF_ROW_WAVES_F64
#define FF_ROW_WAVES_F64 1  // the same minimums for FP64
#endif
#ifndef FF_COL_WAVES_F64
#define FF_COL_WAVES_F64 1
#endif

// Columns per workgroup: kCBK, but one for FP64 4096-point columns (a
// column's buffer and twiddle table take 135 KB of LDS).
template <typename R, int NY>
constexpr int col_cbk() {
  return sizeof(R) == 8 && NY == 4096 ? 1 : kCBK;
}

// STAGE 0: y-inverse of U only.  STAGE 1..4: ETDRK4 stages.  col_cbk columns
// per workgroup; the workgroups of one 4-column tile run on the same XCD.
template <typename R, int NY, int P, int STAGE>
__global__ void __launch_bounds__((col_cbk<R, NY>() * (NY / P)), (sizeof(R) == 4 ? FF_COL_WAVES : FF_COL_WAVES_F64))
    k_vcol(VortGeom g, VortColArgs a) {
  extern __shared__ unsigned char vsm[];
  constexpr int CB = col_cbk<R, NY>();
  int tx, sub;
  xcd_split(blockIdx.x, kVCB / CB, &tx, &sub);
  if (tx >= g.TX) return;  // grid is rounded up to whole XCD groups
  const int t = threadIdx.x;
  const int cc = t % CB;
  constexpr int SL = team_slice(TeamSched<NY, P>::LDS, sizeof(cplx<R>), CB);
  constexpr bool pair = col_pair<R, NY>();  // second buffers after the first, then the table
  cplx<R>* base = reinterpret_cast<cplx<R>*>(vsm);
  const cplx<R>* tw = team_tw_load<NY, P, R>(base + (pair ? 2 : 1) * CB * SL, a.tw);
  vcol_body<R, NY, P, STAGE, pair>(g, a, tx, sub * CB + cc, t / CB, base + cc * SL, tw, true,
                                   pair ? base + (CB + cc) * SL : nullptr);
}

// One column per workgroup of three teams (vcol3_body); the four columns of a
// tile run on the same XCD.
template <typename R, int NY, int P, int STAGE>
__global__ void __launch_bounds__(3 * (NY / P)) k_vcol3(VortGeom g, VortColArgs a) {
  extern __shared__ unsigned char vsm[];
  int tx, c;
  xcd_split(blockIdx.x, kVCB, &tx, &c);
  if (tx >= g.TX) return;  // grid is rounded up to whole XCD groups
  constexpr int T = NY / P;
  const int t = threadIdx.x;
  cplx<R>* base = reinterpret_cast<cplx<R>*>(vsm);
  const cplx<R>* tw = team_tw_load<NY, P, R>(base + 3 * TeamSched<NY, P>::LDS, a.tw);
  vcol3_body<R, NY, P, STAGE>(g, a, tx, c, t / T, t % T, base, tw);
}

// Two rows per workgroup of three teams (vrow3_body); the two workgroups of
// one 4-row group run on the same XCD.
#ifndef FF_ROW3_WAVES
#define FF_ROW3_WAVES 3  // FP32 waves per SIMD the three-team row pass is compiled for
#endif
template <typename R, int NX, int P>
__global__ void __launch_bounds__(3 * (NX / P), sizeof(R) == 4 ? FF_ROW3_WAVES : 2) k_vrow3(VortGeom g,
                                                                                           VortRowArgs a) {
  extern __shared__ unsigned char vsm[];
  int grp, sub;
  xcd_split(blockIdx.x, 2, &grp, &sub);
  if (grp >= g.ny / 4) return;  // grid is rounded up to whole XCD groups
  constexpr int T = NX / P;
  const int t = threadIdx.x;
  cplx<R>* base = reinterpret_cast<cplx<R>*>(vsm);
  const cplx<R>* tw = team_tw_load<NX, P, R>(base + 3 * TeamSched<NX, P>::LDS, a.tw);
  vrow3_body<R, NX, P>(g, a, grp * 4 + sub * 2, t / T, t % T, base, tw);
}

// kRRK rows per workgroup; the kRSUB workgroups of one 4-row group run on the same XCD.
template <typename R, int NX, int P>
__global__ void __launch_bounds__(kRRK*(NX / P), sizeof(R) == 4 ? FF_ROW_WAVES : FF_ROW_WAVES_F64) k_vrow(VortGeom g, VortRowArgs a) {
  extern __shared__ unsigned char vsm[];
  int grp, sub;
  xcd_split(blockIdx.x, kRSUB, &grp, &sub);
  if (grp >= g.ny / 4) return;  // grid is rounded up to whole XCD groups
  const int t = threadIdx.x;
  const int r = t % kRRK;
  constexpr int SL = team_slice(TeamSched<NX, P>::LDS, sizeof(cplx<R>), kRRK);
  cplx<R>* base = reinterpret_cast<cplx<R>*>(vsm);
  if constexpr (row_pair<R>()) {  // second buffers after the first, twiddles in registers
    vrow_body<R, NX, P, true>(g, a, grp * 4 + sub * kRRK + r, t / kRRK, base + r * SL,
                              reinterpret_cast<const cplx<R>*>(a.tw), base + (kRRK + r) * SL);
  } else {
    const cplx<R>* tw = team_tw_load<NX, P, R>(base + kRRK * SL, a.tw);
    vrow_body<R, NX, P>(g, a, grp * 4 + sub * kRRK + r, t / kRRK, base + r * SL, tw);
  }
}

// FP32 rows by wavefronts with half-length transforms (vrowh_body): one
// 4-row group (whole 128 B lines of the blocked arrays) per workgroup of four
// wavefronts, one row each, two LDS buffers per wavefront.
#ifndef FF_ROWH
#define FF_ROWH 1  // FP32 1024 and 2048 row passes use the wavefront kernel
#endif
template <int NX>
constexpr bool rowh_on() {
  return FF_ROWH && (NX == 1024 || NX == 2048);
}
template <int NX>
__global__ void __launch_bounds__(256) k_vrowh(VortGeom g, VortRowArgs a) {
  extern __shared__ unsigned char vsm[];
  constexpr int M = NX / 2, Q = M / 64;
  constexpr int SL = TeamSched<M, Q>::LDS;
  const int grp = blockIdx.x;  // grid = ny / 4
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  cplx<float>* base = reinterpret_cast<cplx<float>*>(vsm) + 2 * w * SL;
  vrowh_body<float, NX, Q>(g, a, grp * 4 + w, lane, base, base + SL);
}

// Rows with half-length transforms by teams of two wavefronts (128 threads:
// FP64 2048-point rows at 8 points per thread, FP32 4096-point rows at 16),
// four rows per workgroup, the M-point twiddle table in LDS after the rows'
// buffers.
#ifndef FF_ROWH_TEAM
#define FF_ROWH_TEAM 1  // FP64 2048 and FP32 4096 row passes use the two-wavefront half-length kernel
#endif
template <typename R, int NX>
constexpr bool rowht_on() {
  return FF_ROWH_TEAM && (NX == 4096 || (sizeof(R) == 8 && NX == 2048));
}
// team size, rows per workgroup and LDS buffers per row: FP64 4096-point
// rows take teams of four wavefronts (2048-point transforms at 8 points per
// thread), two rows per workgroup with one buffer each (their paired
// transforms run one after the other: 102 KB of LDS, two waves per SIMD);
// the others two-wavefront teams, four rows with two buffers each
template <typename R, int NX>
constexpr bool rowht_big() {
  return sizeof(R) == 8 && NX == 4096;
}
template <typename R, int NX>
constexpr int rowht_team() {
  return rowht_big<R, NX>() ? 256 : 128;
}
template <typename R, int NX>
constexpr int rowht_rows() {
  return rowht_big<R, NX>() ? 2 : 4;
}
template <typename R, int NX>
constexpr int rowht_bufs() {
  return rowht_big<R, NX>() ? 1 : 2;
}
template <typename R, int NX>
__global__ void __launch_bounds__((rowht_team<R, NX>() * rowht_rows<R, NX>()), 2) k_vrowht(VortGeom g, VortRowArgs a) {
  extern __shared__ unsigned char vsm[];
  constexpr int M = NX / 2, T = rowht_team<R, NX>(), Q = M / T, RPW = rowht_rows<R, NX>(), NB = rowht_bufs<R, NX>();
  constexpr int SL = TeamSched<M, Q>::LDS;
  cplx<R>* base = reinterpret_cast<cplx<R>*>(vsm);
  const cplx<R>* ltw = team_tw_load<M, Q, R, 2>(base + NB * RPW * SL, a.tw);  // every other NX root
  const int r = threadIdx.x / T, tid = threadIdx.x % T;
  cplx<R>* l1 = base + NB * r * SL;
  vrowh_body<R, NX, Q>(g, a, blockIdx.x * RPW + r, tid, l1, NB == 2 ? l1 + SL : nullptr, ltw);
}

// ----------------------------------------------------------------- host side
static bool pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename R, int N>
static size_t col_lds() {
  constexpr int CB = col_cbk<R, N>();
  constexpr size_t sl = team_slice(TeamSched<N, vcol_pts<R, N>()>::LDS, sizeof(cplx<R>), CB);
  return ((col_pair<R, N>() ? 2 : 1) * CB * sl + tw_slots<N, vcol_pts<R, N>()>()) * sizeof(cplx<R>);
}

template <typename R, int N>
static size_t row_lds() {
  constexpr size_t sl = team_slice(TeamSched<N, vrow_pts<R, N>()>::LDS, sizeof(cplx<R>), kRRK);
  if constexpr (row_pair<R>()) return 2 * kRRK * sl * sizeof(cplx<R>);
  return (kRRK * sl + tw_slots<N, vrow_pts<R, N>()>()) * sizeof(cplx<R>);
}

template <typename R, int N, int STAGE>
static int launch_col(const VortGeom& g, const VortColArgs& a, hipStream_t st) {
  constexpr int P = vcol_pts<R, N>();
  const size_t sh = col_lds<R, N>();
  auto kern = k_vcol<R, N, P, STAGE>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  static const char* names[5] = {"vort_col_stage0", "vort_col_stage1", "vort_col_stage2", "vort_col_stage3",
                                 "vort_col_stage4"};
  const int groups = (g.TX + kXCD - 1) / kXCD * kXCD;
  constexpr int CB = col_cbk<R, N>();
  prof_launch(names[STAGE], col_bytes<R>(g, STAGE, a.do_inverse), kern,
              dim3(groups * (kVCB / CB)), dim3(CB * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_vcol");
  return FFHIP_OK;
}

// Three-team passes (two transforms side by side per column or row pair).
// Measured on MI355X at 2048^2 with the register-twiddle single-team
// kernels: slower in FP32 (the single-team kernels keep two waves per SIMD
// busy) and in FP64 (1953 vs 1891 steps/s with both off), so off by default.
// FFHIP_COL3 / FFHIP_ROW3: 0 off (default), 1 FP64 only, 2 FP32 and FP64.
static bool use3(const char* var, bool f32) {
  const char* e = getenv(var);
  const int mode = e ? atoi(e) : 0;
  return f32 ? mode >= 2 : mode >= 1;
}
// FFHIP_ROWH=0 selects the full-length team kernels for the rows where a
// half-length kernel applies (FP32 1024/2048/4096, FP64 2048; default on)
static bool rowh_enabled() {
  static const bool on = [] {
    const char* e = getenv("FFHIP_ROWH");
    return !e || atoi(e) != 0;
  }();
  return on;
}
// read at every launch (eager steps and graph capture), so a switch set
// before a problem's first run applies to that problem
static bool col3_on(bool f32) { return use3("FFHIP_COL3", f32); }
static bool row3_on(bool f32) { return use3("FFHIP_ROW3", f32); }

template <typename R, int N, int STAGE>
static int launch_col3(const VortGeom& g, const VortColArgs& a, hipStream_t st) {
  constexpr int P = vpts_col(N);
  const size_t sh = ((size_t)3 * TeamSched<N, P>::LDS + tw_slots<N, P>()) * sizeof(cplx<R>);
  auto kern = k_vcol3<R, N, P, STAGE>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  static const char* names[5] = {"vort_col_stage0", "vort_col_stage1", "vort_col_stage2", "vort_col_stage3",
                                 "vort_col_stage4"};
  const int groups = (g.TX + kXCD - 1) / kXCD * kXCD;
  prof_launch(names[STAGE], col_bytes<R>(g, STAGE, a.do_inverse), kern,
              dim3(groups * kVCB), dim3(3 * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_vcol3");
  return FFHIP_OK;
}

template <typename R, int N>
static int launch_row3(const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  constexpr int P = vpts_row(N);
  const size_t sh = ((size_t)3 * TeamSched<N, P>::LDS + tw_slots<N, P>()) * sizeof(cplx<R>);
  auto kern = k_vrow3<R, N, P>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  const int groups = (g.ny / 4 + kXCD - 1) / kXCD * kXCD;
  prof_launch("vort_row", row_bytes<R>(g), kern, dim3(groups * 2), dim3(3 * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_vrow3");
  return FFHIP_OK;
}

template <int N>
static int launch_rowh(const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  constexpr int M = N / 2, Q = M / 64;
  const size_t sh = (size_t)8 * TeamSched<M, Q>::LDS * sizeof(cplx<float>);
  auto kern = k_vrowh<N>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  FF_CHECK(g.ny % 4 == 0, FFHIP_ERR_UNSUPPORTED, "wavefront row kernel: ny = %d", g.ny);
  // the kept band's 16 B chunks fit the lanes' prefetch registers (KXp <= N/3 + 4)
  FF_CHECK(g.KXp <= 128 * (((N / 3 + 4) / 2 + 63) / 64), FFHIP_ERR_UNSUPPORTED, "wavefront row kernel: KXp = %d",
           g.KXp);
  prof_launch("vort_row", row_bytes<float>(g), kern, dim3(g.ny / 4), dim3(256), sh, st, g, a);
  FF_LAUNCHED("k_vrowh");
  return FFHIP_OK;
}

template <typename R, int N>
static int launch_rowht(const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  constexpr int M = N / 2, T = rowht_team<R, N>(), Q = M / T, RPW = rowht_rows<R, N>(), NB = rowht_bufs<R, N>();
  const size_t sh = ((size_t)NB * RPW * TeamSched<M, Q>::LDS + tw_slots<M, Q>()) * sizeof(cplx<R>);
  auto kern = k_vrowht<R, N>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  FF_CHECK(g.ny % RPW == 0, FFHIP_ERR_UNSUPPORTED, "half-length team row kernel: ny = %d", g.ny);
  prof_launch("vort_row", row_bytes<R>(g), kern, dim3(g.ny / RPW), dim3(RPW * T), sh, st, g, a);
  FF_LAUNCHED("k_vrowht");
  return FFHIP_OK;
}

template <typename R, int N>
static int launch_row(const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  if constexpr (sizeof(R) == 4 && rowh_on<N>()) {
    if (rowh_enabled()) return launch_rowh<N>(g, a, st);
  }
  if constexpr (rowht_on<R, N>()) {
    if (rowh_enabled()) return launch_rowht<R, N>(g, a, st);
  }
  if constexpr (sizeof(R) == 8 && N == 4096)  // the full-length team kernel needs more LDS than a workgroup has
    return fail(FFHIP_ERR_UNSUPPORTED, "fused FP64 row pass at nx = 4096 needs the half-length kernel (FFHIP_ROWH)");
  else {
    constexpr int P = vrow_pts<R, N>();
    const size_t sh = row_lds<R, N>();
    auto kern = k_vrow<R, N, P>;
    static bool attr_set = false;
    if (!attr_set) {
      FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)sh));
      attr_set = true;
    }
    const int groups = (g.ny / 4 + kXCD - 1) / kXCD * kXCD;
    prof_launch("vort_row", row_bytes<R>(g), kern, dim3(groups * kRSUB), dim3(kRRK * (N / P)), sh, st, g, a);
    FF_LAUNCHED("k_vrow");
    return FFHIP_OK;
  }
}

template <typename R, int STAGE>
static int dispatch_col(const VortGeom& g, const VortColArgs& a, hipStream_t st) {
  switch (g.ny) {
    case 64: return launch_col<R, 64, STAGE>(g, a, st);
    case 128: return launch_col<R, 128, STAGE>(g, a, st);
    case 256: return launch_col<R, 256, STAGE>(g, a, st);
    case 512: return launch_col<R, 512, STAGE>(g, a, st);
    case 1024:
      if (col3_on(sizeof(R) == 4)) return launch_col3<R, 1024, STAGE>(g, a, st);
      return launch_col<R, 1024, STAGE>(g, a, st);
    case 2048:
      if (col3_on(sizeof(R) == 4)) return launch_col3<R, 2048, STAGE>(g, a, st);
      return launch_col<R, 2048, STAGE>(g, a, st);
    case 4096: return launch_col<R, 4096, STAGE>(g, a, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fused column kernel: ny = %d", g.ny);
  }
}

template <typename R>
static int dispatch_row(const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  switch (g.nx) {
    case 64: return launch_row<R, 64>(g, a, st);
    case 128: return launch_row<R, 128>(g, a, st);
    case 256: return launch_row<R, 256>(g, a, st);
    case 512: return launch_row<R, 512>(g, a, st);
    case 1024:
      if (row3_on(sizeof(R) == 4)) return launch_row3<R, 1024>(g, a, st);
      return launch_row<R, 1024>(g, a, st);
    case 2048:
      if (row3_on(sizeof(R) == 4)) return launch_row3<R, 2048>(g, a, st);
      return launch_row<R, 2048>(g, a, st);
    case 4096: return launch_row<R, 4096>(g, a, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fused row kernel: nx = %d", g.nx);
  }
}

template <typename R>
static int dispatch_stage(int stage, const VortGeom& g, const VortColArgs& a, hipStream_t st) {
  switch (stage) {
    case 0: return dispatch_col<R, 0>(g, a, st);
    case 1: return dispatch_col<R, 1>(g, a, st);
    case 2: return dispatch_col<R, 2>(g, a, st);
    case 3: return dispatch_col<R, 3>(g, a, st);
    default: return dispatch_col<R, 4>(g, a, st);
  }
}

static void drop_graph(VortState* s) {
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  s->exec = nullptr;
  s->graph_packed = nullptr;
}

static void vstate_free(VortState* s) {
  if (!s) return;
  drop_graph(s);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  dev_free(s->U);
  dev_free(s->N1);
  dev_free(s->N2);
  dev_free(s->A);
  for (auto& p : s->I) dev_free(p);
  for (auto& p : s->O) dev_free(p);
  dev_free(s->twx);
  dev_free(s->twy);
  void* adj[] = {s->UB, s->N1B, s->N2B, s->AB, s->LAM, s->IA[0], s->IA[1], s->OA[0], s->OA[1], s->OA[2],
                 s->CK, s->FS};
  for (void* p : adj) dev_free(p);
  delete s;
}

int fused_create(ffhip_op* op) {
  op->fused = nullptr;
  if (op->kind != FFHIP_OP_VORTICITY2D) return FFHIP_OK;
  const char* dis = getenv("FFHIP_DISABLE_FUSED");
  if (dis && atoi(dis) != 0) return FFHIP_OK;
  const int64_t nx = op->n[0], ny = op->n[1];
  const int64_t nmax = 4096;
  if (!pow2(nx) || !pow2(ny) || nx < 64 || ny < 64 || nx > nmax || ny > nmax) return FFHIP_OK;
  if (!(op->af > 0)) return FFHIP_OK;  // the fused kernels assume the Nyquist modes are aliased
  VortGeom g;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.nkr = (int)op->nkr;
  g.KX = (int)op->lo[0] - 1;
  g.KXp = (g.KX + kVCB - 1) / kVCB * kVCB;
  g.TX = g.KXp / kVCB;
  g.ylo = (int)op->lo[1] - 1;
  g.yhi = (int)op->hi[1] - 1;
  g.KY = g.ny - (g.yhi - g.ylo + 1);
  const double twopi = 6.283185307179586476925286766559;
  g.dk = twopi / op->box[0];
  g.dl = twopi / op->box[1];
  g.s2 = 1.0 / ((double)nx * (double)ny * (double)nx * (double)ny);
  // the band must be a proper sub-band containing the Nyquist row/column
  if (g.KX < 1 || g.KX > g.nkr - 1 || g.ylo < 1 || g.yhi < g.ylo || g.yhi >= g.ny || g.ylo > g.ny / 2 ||
      g.yhi < g.ny / 2 || g.KX > g.nx / 2)
    return FFHIP_OK;
  VortState* s = new (std::nothrow) VortState();
  FF_CHECK(s, FFHIP_ERR_OOM, "host allocation failed");
  s->g = g;
  const size_t cs = cplx_size(op->prec);
  const size_t sb = (size_t)g.TX * g.KY * kVCB * cs;
  const size_t ib = (size_t)g.TX * g.ny * kVCB * cs;
  int rc = dev_alloc(&s->U, sb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->N1, sb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->N2, sb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->A, sb);
  for (int i = 0; i < 3 && rc == FFHIP_OK; ++i) rc = dev_alloc(&s->I[i], ib);
  for (int i = 0; i < 2 && rc == FFHIP_OK; ++i) rc = dev_alloc(&s->O[i], ib);
  if (rc == FFHIP_OK) rc = upload_roots(nx, nx, op->prec, &s->twx);
  if (rc == FFHIP_OK) rc = upload_roots(ny, ny, op->prec, &s->twy);
  if (rc != FFHIP_OK) {
    vstate_free(s);
    return rc;
  }
  op->fused = s;
  return FFHIP_OK;
}

void fused_destroy(ffhip_op* op) {
  vstate_free(reinterpret_cast<VortState*>(op->fused));
  op->fused = nullptr;
}

// FFHIP_VORT_GRAPHS=0 keeps every step eager (graph replays are the default)
static bool vort_graphs_on() {
  static const bool on = [] {
    const char* e = getenv("FFHIP_VORT_GRAPHS");
    return !e || atoi(e) != 0;
  }();
  return on;
}

// packed coefficient arrays (blocked kept band), cached on the stepper
static uint64_t geom_key(const VortGeom& g) {
  uint64_t k = 1469598103934665603ULL;
  const int v[7] = {g.nx, g.ny, g.KX, g.ylo, g.yhi, g.TX, g.KY};
  for (int x : v) k = (k ^ (uint64_t)(uint32_t)x) * 1099511628211ULL;
  return k;
}

int vort_ensure_packed(ffhip_op* op, ffhip_etdrk4* ts, const VortGeom& g, hipStream_t st) {
  const uint64_t key = geom_key(g);
  if (ts->packed && ts->packed_owner == op && ts->packed_key == key) return FFHIP_OK;
  dev_free(ts->packed);
  ts->packed = nullptr;
  const size_t n = (size_t)g.TX * g.KY * kVCB;
  const int ncoef = ts->filter ? 7 : 6;  // zeta, alpha, beta, Gamma, exp(dt L), exp(dt L/2)[, filter]
  FF_TRY(dev_alloc(&ts->packed, ncoef * n * sizeof(double)));
  double* dst = reinterpret_cast<double*>(ts->packed);
  for (int c = 0; c < ncoef; ++c) {
    const void* src = c < 6 ? ts->coef[c] : ts->filter;
    hipLaunchKernelGGL(k_vpack, dim3(blocks_for((int64_t)n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const double*>(src), dst + c * n, g);
    FF_LAUNCHED("k_vpack");
  }
  ts->packed_owner = op;
  ts->packed_key = key;
  return FFHIP_OK;
}

template <typename R>
static int vrun(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, hipStream_t st) {
  VortState* s = reinterpret_cast<VortState*>(op->fused);
  const VortGeom& g = s->g;
  FF_TRY(vort_ensure_packed(op, ts, g, st));
  const size_t n = (size_t)g.TX * g.KY * kVCB;
  const double* pc = reinterpret_cast<const double*>(ts->packed);
  hipLaunchKernelGGL((k_vimport<R>), dim3(blocks_for((int64_t)n, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(sol), reinterpret_cast<cplx<R>*>(s->U), g);
  FF_LAUNCHED("k_vimport");
  VortColArgs ca;
  ca.O0 = s->O[0];
  ca.O1 = s->O[1];
  ca.I0 = s->I[0];
  ca.I1 = s->I[1];
  ca.I2 = s->I[2];
  ca.U = s->U;
  ca.N1 = s->N1;
  ca.N2 = s->N2;
  ca.A = s->A;
  ca.Z = pc + 0 * n;
  ca.Al = pc + 1 * n;
  ca.Be = pc + 2 * n;
  ca.Ga = pc + 3 * n;
  ca.E = pc + 4 * n;
  ca.Eh = pc + 5 * n;
  ca.tw = s->twy;
  ca.do_inverse = 1;
  ca.Fl = ts->filter ? pc + 6 * n : nullptr;
  VortRowArgs ra;
  ra.I0 = s->I[0];
  ra.I1 = s->I[1];
  ra.I2 = s->I[2];
  ra.O0 = s->O[0];
  ra.O1 = s->O[1];
  ra.tw = s->twx;
  ra.F = nullptr;
  FF_TRY(dispatch_stage<R>(0, g, ca, st));
  // one step: 4 x (row pass, column stage); the last step of the run skips
  // its final inverse y-FFTs (nothing reads them)
  auto step = [&](hipStream_t s_, bool last) -> int {
    for (int stage = 1; stage <= 4; ++stage) {
      FF_TRY(dispatch_row<R>(g, ra, s_));
      ca.do_inverse = (stage < 4 || !last) ? 1 : 0;
      FF_TRY(dispatch_stage<R>(stage, g, ca, s_));
    }
    return FFHIP_OK;
  };
  int64_t left = nsteps;
  // Whole chunks of kVortChunk steps replay a captured graph.  The graph is
  // captured at the first run of the operator with this stepper, whatever its
  // length, right after one eager step (so every kernel's LDS attribute is
  // set outside the capture); later runs find it ready.  The last step runs
  // eagerly because it differs.  Profiled runs stay eager: they time each
  // launch.
  const bool graphs = !g_prof_on && vort_graphs_on();
  if (graphs && left >= 1 && (!s->exec || s->graph_packed != ts->packed || s->graph_filter != (ts->filter != nullptr))) {
    FF_TRY(step(st, left == 1));
    --left;
    {
      drop_graph(s);
      if (!s->cap) FF_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
      FF_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
      int rc = FFHIP_OK;
      for (int k = 0; k < kVortChunk && rc == FFHIP_OK; ++k) rc = step(s->cap, false);
      hipGraph_t graph = nullptr;
      const hipError_t ee = hipStreamEndCapture(s->cap, &graph);
      if (rc != FFHIP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (ee != hipSuccess) return fail(FFHIP_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ee));
      const hipError_t ei = hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        s->exec = nullptr;
        return fail(FFHIP_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
      }
      s->graph_packed = ts->packed;
      s->graph_filter = ts->filter != nullptr;
      FF_HIP(hipGraphUpload(s->exec, st));  // the first replay then carries no upload
    }
  }
  if (graphs)
    for (; left > kVortChunk; left -= kVortChunk) FF_HIP(hipGraphLaunch(s->exec, st));
  for (int64_t k = 0; k < left; ++k) FF_TRY(step(st, k + 1 == left));
  const int64_t total = (int64_t)g.ny * g.nkr;
  hipLaunchKernelGGL((k_vexport<R>), dim3(blocks_for(total, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<cplx<R>*>(sol), reinterpret_cast<const cplx<R>*>(s->U),
                     reinterpret_cast<const double*>(ts->coef[4]), reinterpret_cast<const double*>(ts->filter),
                     nsteps, g);
  FF_LAUNCHED("k_vexport");
  return FFHIP_OK;
}

int vort_row(int prec, const VortGeom& g, const VortRowArgs& a, hipStream_t st) {
  return prec == FFHIP_F32 ? dispatch_row<float>(g, a, st) : dispatch_row<double>(g, a, st);
}

int vort_col(int prec, int stage, const VortGeom& g, const VortColArgs& a, hipStream_t st) {
  return prec == FFHIP_F32 ? dispatch_stage<float>(stage, g, a, st) : dispatch_stage<double>(stage, g, a, st);
}

int vort_import(int prec, const VortGeom& g, const void* sol, void* U, hipStream_t st) {
  const int64_t n = (int64_t)g.TX * g.KY * kVCB;
  if (prec == FFHIP_F32)
    hipLaunchKernelGGL((k_vimport<float>), dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const float2*>(sol), reinterpret_cast<float2*>(U), g);
  else
    hipLaunchKernelGGL((k_vimport<double>), dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const double2*>(sol), reinterpret_cast<double2*>(U), g);
  FF_LAUNCHED("k_vimport");
  return FFHIP_OK;
}

int vort_export(int prec, const VortGeom& g, void* sol, const void* U, const double* E, const double* F,
                int64_t nsteps, hipStream_t st) {
  const int64_t total = (int64_t)g.ny * g.nkr;
  if (prec == FFHIP_F32)
    hipLaunchKernelGGL((k_vexport<float>), dim3(blocks_for(total, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<float2*>(sol), reinterpret_cast<const float2*>(U), E, F, nsteps, g);
  else
    hipLaunchKernelGGL((k_vexport<double>), dim3(blocks_for(total, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<double2*>(sol), reinterpret_cast<const double2*>(U), E, F, nsteps, g);
  FF_LAUNCHED("k_vexport");
  return FFHIP_OK;
}

int fused_etdrk4_run(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, hipStream_t st) {
  FF_CHECK(op->fused, FFHIP_ERR_UNSUPPORTED, "no fused path for this operator");
  return op->prec == FFHIP_F32 ? vrun<float>(op, ts, sol, nsteps, st) : vrun<double>(op, ts, sol, nsteps, st);
}

}  // namespace ffhip
