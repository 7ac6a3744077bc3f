// Shared definitions of the fused 2D vorticity kernels (forward and adjoint).
#pragma once

#include <cstdint>
#include <type_traits>

#include "ops.h"
#include "prof.h"
#include "team_fft.h"

namespace ffhip {

#ifndef FF_ROW_PREFETCH_F32
#define FF_ROW_PREFETCH_F32 0
#endif
#ifndef FF_ROW_TWREG
#define FF_ROW_TWREG 1  // FP32 row pass: twiddles in registers across its three transforms
#endif
#ifndef FF_COL_TWREG
#define FF_COL_TWREG 1  // FP32 column pass: first-pass twiddles in registers across its five transforms
#endif
#ifndef FF_ROW_TWREG_F64
#define FF_ROW_TWREG_F64 0  // the same for FP64 rows (measured no faster at 2048^2)
#endif
#ifndef FF_COL_TWREG_F64
#define FF_COL_TWREG_F64 1  // FP64 column pass: first-pass twiddles in registers (2048^2: stages ~1% faster)
#endif
#ifndef FF_ROW_PREFETCH_F64
#define FF_ROW_PREFETCH_F64 1
#endif
#ifndef FF_ROWH_UVPRE
#define FF_ROWH_UVPRE 1  // wavefront row pass: u^, v^ loaded before the zeta transform
#endif
#ifndef FF_COL_O1PRE
#define FF_COL_O1PRE 1  // FP64 column pass: second forward input loaded before the first transform
#endif
#ifndef FF_STAGE_UNROLL
#define FF_STAGE_UNROLL 16  // stage-combination loads in flight per thread
#endif

constexpr int kVCB = 4;  // kx columns per layout block (tile width)
#ifndef FF_CBK
#define FF_CBK 2  // kx columns per column-kernel workgroup (a sub-block of the tile)
#endif
#ifndef FF_RRK
#define FF_RRK 2  // y rows per row-kernel workgroup
#endif
constexpr int kCBK = FF_CBK;
constexpr int kRRK = FF_RRK;
constexpr int kCSUB = kVCB / kCBK;  // column workgroups sharing one tile
constexpr int kRSUB = 4 / kRRK;     // row workgroups sharing the 4-row, 128 B chunks
constexpr int kXCD = 8;             // workgroups are dispatched round-robin over 8 XCDs

// Workgroups that share cache lines are placed on the same XCD (same L2):
// block b runs on XCD b % 8, so sub-block s of group q gets
// b = (q % 8) + 8 * ((q / 8) * NSUB + s).
__device__ __forceinline__ void xcd_split(int b, int nsub, int* group, int* sub) {
  const int x = b % kXCD;
  const int slot = b / kXCD;
  *sub = slot % nsub;
  *group = (slot / nsub) * kXCD + x;
}

// Element idx of an array addressed by a 32-bit byte offset from its base
// (valid for arrays under 4 GiB): the load or store takes the uniform base in
// scalar registers and one 32-bit offset register, with no 64-bit address
// arithmetic per element.
template <typename T>
__device__ __forceinline__ T& at32(T* base, uint32_t idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
}

// points per thread of the y (column) and x (row) transforms
constexpr int vpts_col(int n) { return n >= 256 ? 16 : 8; }
constexpr int vpts_row(int n) { return n >= 256 ? 16 : 8; }

// Points per thread of the single-team column passes (forward and
// adjoint).  FP32 takes 8 instead of 16 where that adds no LDS exchange pass
// (512 = 8^3): twice the threads per column for the same passes and fewer
// registers (measured on MI355X: 512^2 forward 46.0 -> 37.4 ms per 500
// steps).  Where 8 adds a pass (1024 and 2048: 4 vs 3) it stays at 16
// (2048^2 measured 34.5 vs 27.5 us per stage).  FP64 at 2048 takes 8
// although it adds a pass: 16 FP64 points per thread need more than 256
// registers, one wave per SIMD, and 8 fit two (2048^2 FP64 measured: column
// stages 65 -> 55 us, rows 59 -> 57 us, 1.96k -> 2.16k steps/s).
#ifndef FF_COL_P8_ALWAYS
#define FF_COL_P8_ALWAYS 0  // 1: FP32 columns take 8 points per thread at every size (measurement switch)
#endif
#ifndef FF_COL_P8_F64
#define FF_COL_P8_F64 1  // FP64 columns of 2048 points take 8 points per thread (0: 16)
#endif
#ifndef FF_ROW_P8_F64
#define FF_ROW_P8_F64 1  // FP64 rows of 2048 points take 8 points per thread (0: 16)
#endif
template <typename R, int N>
constexpr int vcol_pts() {
  if constexpr (sizeof(R) == 8) return N == 4096 || (FF_COL_P8_F64 && N == 2048) ? 8 : vpts_col(N);
  return N >= 256 && (FF_COL_P8_ALWAYS || TeamSched<N, 8>::NPASS == TeamSched<N, 16>::NPASS) ? 8 : vpts_col(N);
}
// The same rule for the single-team row passes, forward and adjoint (512^2
// forward, measured: 37.3 -> 33.0 ms per 500 steps).
template <typename R, int N>
constexpr int vrow_pts() {
  if constexpr (sizeof(R) == 8) return FF_ROW_P8_F64 && N == 2048 ? 8 : vpts_row(N);
  return N >= 256 && TeamSched<N, 8>::NPASS == TeamSched<N, 16>::NPASS ? 8 : vpts_row(N);
}

struct VortGeom {
  int nx, ny, nkr;
  int KX, KXp, TX;  // kept kx count, padded to CB, column blocks
  int ylo, yhi, KY; // aliased rows [ylo, yhi] (0-based) and kept row count
  double dk, dl;    // 2 pi / Lx, 2 pi / Ly
  double s2;        // 1 / (nx ny)^2: normalisation of the product of two unnormalised inverses
};

struct VortState {
  VortGeom g;
  void* U = nullptr;   // state (blocked, kept band)
  void* N1 = nullptr;  // N(u)
  void* N2 = nullptr;  // N(a), then N(a) + N(b)
  void* A = nullptr;   // stage state a
  void* I[3] = {nullptr, nullptr, nullptr};  // row-kernel inputs: zeta^, u^, v^ (y-inverted)
  void* O[2] = {nullptr, nullptr};           // row-kernel outputs: (u zeta)^, (v zeta)^ (x-forward)
  void* twx = nullptr;  // exp(-2 pi i k/nx), k < nx
  void* twy = nullptr;  // exp(-2 pi i k/ny), k < ny
  // discrete adjoint (allocated on first use)
  void* UB = nullptr;    // ubar
  void* N1B = nullptr;   // N1bar
  void* N2B = nullptr;   // 2 beta lambda, then N2bar
  void* AB = nullptr;    // abar
  void* LAM = nullptr;   // adjoint state at the end of the window (in), start (out)
  void* IA[2] = {nullptr, nullptr};           // adjoint row inputs: y-inverted i kx s / W, i l s / W
  void* OA[3] = {nullptr, nullptr, nullptr};  // adjoint row outputs (x-forward): u-, v-, zeta-adjoint terms
  void* CK = nullptr;    // checkpoints (blocked kept band)
  int64_t ck_cap = 0;    // checkpoints allocated
  void* FS = nullptr;    // stored stage fields of one recomputed segment
  int64_t fs_cap = 0;    // steps of stored fields allocated
  // the checkpointed forward run the checkpoints belong to
  const void* ck_ts = nullptr;
  int64_t ck_nsteps = -1, ck_ncheck = 0;
  int64_t ck_id = 0;     // identifies that run (0: none); each forward_ckpt call takes a new one
  bool fs_last = false;  // FS holds the last segment's fields from the forward run
  // kVortChunk captured steps of the forward run, replayed as one graph launch;
  // the graph holds the stepper's packed-coefficient pointer it was captured with
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* graph_packed = nullptr;
  bool graph_filter = false;
};

constexpr int kVortChunk = 10;  // ETDRK4 steps per captured graph (80 kernel nodes)

__device__ __forceinline__ int vfreq(int i, int n) { return i < n / 2 ? i : i - n; }

// compact-row index of kept row iy
__device__ __forceinline__ int v_ky(int iy, const VortGeom& g) { return iy < g.ylo ? iy : iy - (g.yhi - g.ylo + 1); }
__device__ __forceinline__ int v_iy(int ky, const VortGeom& g) { return ky < g.ylo ? ky : ky + (g.yhi - g.ylo + 1); }

struct VortColArgs {
  const void* O0;
  const void* O1;
  void* I0;
  void* I1;
  void* I2;
  void* U;
  void* N1;
  void* N2;
  void* A;
  const double* E;    // exp(dt L)
  const double* Eh;   // exp(dt L / 2)
  const double* Z;    // zeta
  const double* Al;   // alpha
  const double* Be;   // beta
  const double* Ga;   // Gamma
  const void* tw;
  int do_inverse;
  const double* Fl = nullptr;  // spectral filter applied after the update (FilteredETDRK4), or null
};

struct VortRowArgs {
  const void* I0;
  const void* I1;
  const void* I2;
  void* O0;
  void* O1;
  const void* tw;
  void* F;  // optional store of this stage's physical fields (u + i v, zeta), null = off
};

// stored fields of one stage: u + i v (complex, normalised) then zeta (real), each [ny][nx]
template <typename R> inline size_t vfield_bytes(const VortGeom& g) {
  return (size_t)g.nx * g.ny * (sizeof(cplx<R>) + sizeof(R));
}

// host launchers implemented in vort2d_fused.hip
int vort_row(int prec, const VortGeom& g, const VortRowArgs& a, hipStream_t st);
int vort_col(int prec, int stage, const VortGeom& g, const VortColArgs& a, hipStream_t st);
int vort_import(int prec, const VortGeom& g, const void* sol, void* U, hipStream_t st);
int vort_export(int prec, const VortGeom& g, void* sol, const void* U, const double* E, const double* F,
                int64_t nsteps, hipStream_t st);
int vort_ensure_packed(ffhip_op* op, ffhip_etdrk4* ts, const VortGeom& g, hipStream_t st);

}  // namespace ffhip
