// Fast kernels of the FFT plans (mul!/ldiv!, src/domains.jl:2-5) on
// power-of-two axes, built from the register-resident team transforms of
// team_fft.h.  The generic Stockham kernels of fft.hip keep every other size.
//
//   rows (x axis of an rfft plan, nx = 2M with M in 256..1024): one wavefront
//     per row.  r2c runs the M-point complex FFT of the row read as M complex
//     pairs (x[2n], x[2n+1]) and splits it into the M + 1 half-spectrum
//     coefficients; the mirrored coefficient Z[M - k] that the split needs
//     sits in lane (64 - l) mod 64 and comes over with ds_bpermute (__shfl).
//     c2r is the mirror image: the pre-split reads X[k] and X[M - k] from
//     memory and one M-point inverse gives the row's sample pairs.
//     Exchanges inside a transform are wavefront-level, so the four rows of
//     a workgroup never wait for each other.
//   columns (y and z axes, n in 64..2048): C adjacent columns per workgroup,
//     one team of n/P threads per column with its threads interleaved lane by
//     lane (thread t in column t % C), so a wavefront's load or store touches
//     C consecutive complex values of 64/C rows.  Workgroups that share 128 B
//     lines run on the same XCD (same L2).
// Every twiddle is in registers, filled from the tables exp(-2 pi i k/n)
// that the plan already holds (generated in FP64 on the host and rounded
// once to the plan precision).  Arithmetic differs from the generic kernels
// only in the radix order (roundoff-level differences).
#include <cstdlib>

#include "plan.h"
#include "prof.h"
#include "team_fft.h"

namespace ffhip {

// ----------------------------------------------------------------- rows
template <typename R, int M>
__global__ void __launch_bounds__(256) k_fr2c(const R* __restrict__ in, cplx<R>* __restrict__ out,
                                              const cplx<R>* __restrict__ rootsM, const cplx<R>* __restrict__ rtw,
                                              int64_t nrows) {
  extern __shared__ unsigned char fsm[];
  constexpr int Q = M / 64;
  constexpr int SL = TeamSched<M, Q>::LDS;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + w;
  if (row >= nrows) return;  // whole wavefront; no workgroup barrier follows
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(fsm) + w * SL;
  TwRegs<M, Q, R> twr;
  twr_fill_roots<M, Q, 1, R, TwRegs<M, Q, R>::NR>(twr, rootsM, lane);
  const cplx<R>* src = reinterpret_cast<const cplx<R>*>(in) + row * M;  // pairs (x[2n], x[2n+1])
  cplx<R> z[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) z[q] = src[lane + 64 * q];
  wave_fft<M, Q, -1, R>(z, lds, lane, twr);
  // X[k] = (Z[k] + conj Z[M-k]) / 2 + rtw[k] (Z[k] - conj Z[M-k]) / (2i); for
  // k = l + 64 q, M - k = (64 - l) + 64 (Q-1-q) (lane l >= 1) or 64 (Q - q) (lane 0)
  cplx<R>* dst = out + row * (M + 1);
  const int partner = (64 - lane) & 63;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cplx<R> mir;
    mir.x = __shfl(z[Q - 1 - q].x, partner, 64);
    mir.y = __shfl(z[Q - 1 - q].y, partner, 64);
    if (lane == 0) mir = z[q == 0 ? 0 : Q - q];
    const int k = lane + 64 * q;
    const cplx<R> zk = z[q], zc = cconj<R>(mir);
    const cplx<R> xe = cscale<R>(cadd<R>(zk, zc), R(0.5));
    const cplx<R> d = csub<R>(zk, zc);
    const cplx<R> xo = mkc<R>(R(0.5) * d.y, -R(0.5) * d.x);  // (zk - zc) / (2i)
    dst[k] = cadd<R>(xe, cmul<R>(rtw[k], xo));
  }
  if (lane == 0) {  // k = M: zk = Z[0], zc = conj Z[0]
    const cplx<R> zk = z[0], zc = cconj<R>(z[0]);
    const cplx<R> xe = cscale<R>(cadd<R>(zk, zc), R(0.5));
    const cplx<R> d = csub<R>(zk, zc);
    const cplx<R> xo = mkc<R>(R(0.5) * d.y, -R(0.5) * d.x);
    dst[M] = cadd<R>(xe, cmul<R>(rtw[M], xo));
  }
}

template <typename R, int M>
__global__ void __launch_bounds__(256) k_fc2r(const cplx<R>* __restrict__ in, R* __restrict__ out,
                                              const cplx<R>* __restrict__ rootsM, const cplx<R>* __restrict__ rtw,
                                              int64_t nrows, R scale) {
  extern __shared__ unsigned char fsm[];
  constexpr int Q = M / 64;
  constexpr int SL = TeamSched<M, Q>::LDS;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + w;
  if (row >= nrows) return;  // whole wavefront; no workgroup barrier follows
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(fsm) + w * SL;
  TwRegs<M, Q, R> twr;
  twr_fill_roots<M, Q, 1, R, TwRegs<M, Q, R>::NR>(twr, rootsM, lane);
  const cplx<R>* src = in + row * (M + 1);
  cplx<R> z[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = lane + 64 * q;
    cplx<R> xk = src[k], xm = src[M - k];
    if (k == 0) {  // imaginary parts of DC and Nyquist are ignored
      xk.y = 0;
      xm.y = 0;
    }
    const cplx<R> xmc = cconj<R>(xm);
    const cplx<R> xe = cscale<R>(cadd<R>(xk, xmc), R(0.5));
    const cplx<R> xo = cmul<R>(cscale<R>(csub<R>(xk, xmc), R(0.5)), cconj<R>(rtw[k]));
    z[q] = mkc<R>(xe.x - xo.y, xe.y + xo.x);  // xe + i xo
  }
  wave_fft<M, Q, 1, R>(z, lds, lane, twr);
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(out) + row * M;
#pragma unroll
  for (int q = 0; q < Q; ++q) dst[lane + 64 * q] = cscale<R>(z[q], scale);
}

// ----------------------------------------------------------------- columns
#ifndef FF_FCOL_P16
#define FF_FCOL_P16 1024  // columns of at least this length take 16 points per thread, shorter ones 8
#endif
constexpr int fcol_pts(int n) { return n >= FF_FCOL_P16 ? 16 : 8; }
template <typename R, int N>
constexpr int fcol_cols() {  // columns per workgroup: >= 256 threads, >= 32 B per row segment
  constexpr int T = N / fcol_pts(N);
  constexpr int cmin = sizeof(R) == 4 ? 4 : 2;
  return 256 / T > cmin ? 256 / T : cmin;
}
template <typename R, int N>
constexpr int fcol_sub() {  // workgroups whose column segments share one 128 B line
  constexpr int seg = fcol_cols<R, N>() * (int)sizeof(cplx<R>);
  return seg >= 128 ? 1 : 128 / seg;
}

// Block b runs on XCD b % 8.  Each XCD takes a contiguous range of gpx line
// groups (group q on XCD q / gpx), its NSUB workgroups per group side by
// side: neighbouring groups then share an L2 too, which matters when the row
// pitch is not a multiple of 128 B (an rfft array's nx/2 + 1 columns) and a
// group's rows straddle the lines of the next group.
__device__ __forceinline__ void fxcd_split(int b, int nsub, int gpx, int* group, int* sub) {
  const int x = b % 8;
  const int slot = b / 8;
  *sub = slot % nsub;
  *group = x * gpx + slot / nsub;
}

// Transform (slab s, column c) has element i at s * slab_dist + c + i * stride;
// in and out may be the same array (each element is read and written by the
// same thread, reads first).
template <typename R, int N, int DIR>
__global__ void __launch_bounds__((fcol_cols<R, N>() * (N / fcol_pts(N)))) k_fcol(
    const cplx<R>* in, cplx<R>* out, const cplx<R>* __restrict__ roots, int64_t ncols, int64_t stride,
    int64_t slab_dist, R scale, int gpx) {
  extern __shared__ unsigned char fsm[];
  constexpr int P = fcol_pts(N);
  constexpr int T = N / P;
  constexpr int C = fcol_cols<R, N>();
  constexpr int SL = team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), C);
  int grp, sub;
  fxcd_split(blockIdx.x, fcol_sub<R, N>(), gpx, &grp, &sub);
  const int64_t col0 = ((int64_t)grp * fcol_sub<R, N>() + sub) * C;
  if (col0 >= ncols) return;  // whole workgroup, before any barrier
  const int t = threadIdx.x;
  const int c = t % C, tid = t / C;
  const int64_t col = col0 + c;
  const bool live = col < ncols;
  const int64_t base = (int64_t)blockIdx.y * slab_dist + (live ? col : 0);
  TwRegs<N, P, R> twr;
  twr_fill_roots<N, P, 1, R, TwRegs<N, P, R>::NR>(twr, roots, tid);
  cplx<R> v[P];
#pragma unroll
  for (int p = 0; p < P; ++p) v[p] = live ? in[base + (int64_t)(tid + p * T) * stride] : mkc<R>(0, 0);
  team_fft<N, P, DIR, R>(v, reinterpret_cast<cplx<R>*>(fsm) + c * SL, tid, nullptr, twr);
  if (live) {
#pragma unroll
    for (int p = 0; p < P; ++p) out[base + (int64_t)(tid + p * T) * stride] = cscale<R>(v[p], scale);
  }
}

// ----------------------------------------------------------------- host side
static bool fast_on() {  // read at each plan creation
  const char* e = getenv("FFHIP_FFT_FAST");
  return !e || atoi(e) != 0;
}

bool fast_rows_ok(int64_t m) { return fast_on() && (m == 256 || m == 512 || m == 1024); }
bool fast_cols_ok(int64_t n) { return fast_on() && n >= 64 && n <= 2048 && (n & (n - 1)) == 0; }

template <typename K>
static int set_lds(K kern, size_t sh) {
  FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)sh));
  return FFHIP_OK;
}

template <typename R, int M>
static int rows_r2c_m(const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, hipStream_t st) {
  const size_t sh = (size_t)4 * TeamSched<M, M / 64>::LDS * sizeof(cplx<R>);
  static bool attr = false;
  if (!attr) {
    FF_TRY(set_lds(k_fr2c<R, M>, sh));
    attr = true;
  }
  const double bytes = (double)nrows * (2.0 * M * sizeof(R) + (M + 1) * sizeof(cplx<R>));
  prof_launch("fft_rows_r2c", bytes, k_fr2c<R, M>, dim3((unsigned)((nrows + 3) / 4)), dim3(256), sh, st,
              reinterpret_cast<const R*>(in), reinterpret_cast<cplx<R>*>(out),
              reinterpret_cast<const cplx<R>*>(ax.tw), reinterpret_cast<const cplx<R>*>(rtw), nrows);
  FF_LAUNCHED("k_fr2c");
  return FFHIP_OK;
}

template <typename R, int M>
static int rows_c2r_m(const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, double scale,
                      hipStream_t st) {
  const size_t sh = (size_t)4 * TeamSched<M, M / 64>::LDS * sizeof(cplx<R>);
  static bool attr = false;
  if (!attr) {
    FF_TRY(set_lds(k_fc2r<R, M>, sh));
    attr = true;
  }
  const double bytes = (double)nrows * (2.0 * M * sizeof(R) + (M + 1) * sizeof(cplx<R>));
  prof_launch("fft_rows_c2r", bytes, k_fc2r<R, M>, dim3((unsigned)((nrows + 3) / 4)), dim3(256), sh, st,
              reinterpret_cast<const cplx<R>*>(in), reinterpret_cast<R*>(out),
              reinterpret_cast<const cplx<R>*>(ax.tw), reinterpret_cast<const cplx<R>*>(rtw), nrows, (R)scale);
  FF_LAUNCHED("k_fc2r");
  return FFHIP_OK;
}

template <typename R>
static int rows_r2c_t(const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, hipStream_t st) {
  switch (ax.n) {
    case 256: return rows_r2c_m<R, 256>(in, out, nrows, ax, rtw, st);
    case 512: return rows_r2c_m<R, 512>(in, out, nrows, ax, rtw, st);
    case 1024: return rows_r2c_m<R, 1024>(in, out, nrows, ax, rtw, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fast r2c rows: m = %lld", (long long)ax.n);
  }
}

template <typename R>
static int rows_c2r_t(const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, double scale,
                      hipStream_t st) {
  switch (ax.n) {
    case 256: return rows_c2r_m<R, 256>(in, out, nrows, ax, rtw, scale, st);
    case 512: return rows_c2r_m<R, 512>(in, out, nrows, ax, rtw, scale, st);
    case 1024: return rows_c2r_m<R, 1024>(in, out, nrows, ax, rtw, scale, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fast c2r rows: m = %lld", (long long)ax.n);
  }
}

int fast_rows_r2c(int prec, const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw,
                  hipStream_t st) {
  return prec == FFHIP_F32 ? rows_r2c_t<float>(in, out, nrows, ax, rtw, st)
                           : rows_r2c_t<double>(in, out, nrows, ax, rtw, st);
}

int fast_rows_c2r(int prec, const void* in, void* out, int64_t nrows, const Axis& ax, const void* rtw, double scale,
                  hipStream_t st) {
  return prec == FFHIP_F32 ? rows_c2r_t<float>(in, out, nrows, ax, rtw, scale, st)
                           : rows_c2r_t<double>(in, out, nrows, ax, rtw, scale, st);
}

template <typename R, int N, int DIR>
static int cols_n_dir(const void* in, void* out, int64_t ncols, int64_t stride, int64_t nslabs, int64_t slab_dist,
                      const Axis& ax, double scale, hipStream_t st) {
  constexpr int P = fcol_pts(N);
  constexpr int C = fcol_cols<R, N>();
  constexpr int SUB = fcol_sub<R, N>();
  constexpr int SL = team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), C);
  const size_t sh = (size_t)C * SL * sizeof(cplx<R>);
  static bool attr = false;
  if (!attr) {
    FF_TRY(set_lds(k_fcol<R, N, DIR>, sh));
    attr = true;
  }
  FF_CHECK(nslabs <= 65535, FFHIP_ERR_UNSUPPORTED, "fast column pass: %lld slabs", (long long)nslabs);
  const int64_t groups = (ncols + (int64_t)C * SUB - 1) / ((int64_t)C * SUB);
  const int64_t gpx = (groups + 7) / 8;  // line groups per XCD
  const int64_t gx = gpx * 8 * SUB;
  FF_CHECK(gx < (1LL << 31), FFHIP_ERR_UNSUPPORTED, "fast column pass: %lld columns", (long long)ncols);
  const double bytes = 2.0 * ncols * nslabs * N * sizeof(cplx<R>);
  prof_launch("fft_cols_c2c", bytes, k_fcol<R, N, DIR>, dim3((unsigned)gx, (unsigned)nslabs), dim3(C * (N / P)), sh,
              st, reinterpret_cast<const cplx<R>*>(in), reinterpret_cast<cplx<R>*>(out),
              reinterpret_cast<const cplx<R>*>(ax.tw), ncols, stride, slab_dist, (R)scale, (int)gpx);
  FF_LAUNCHED("k_fcol");
  return FFHIP_OK;
}

template <typename R, int N>
static int cols_n(const void* in, void* out, int64_t ncols, int64_t stride, int64_t nslabs, int64_t slab_dist,
                  const Axis& ax, int dir, double scale, hipStream_t st) {
  return dir < 0 ? cols_n_dir<R, N, -1>(in, out, ncols, stride, nslabs, slab_dist, ax, scale, st)
                 : cols_n_dir<R, N, 1>(in, out, ncols, stride, nslabs, slab_dist, ax, scale, st);
}

template <typename R>
static int cols_t(const void* in, void* out, int64_t ncols, int64_t stride, int64_t nslabs, int64_t slab_dist,
                  const Axis& ax, int dir, double scale, hipStream_t st) {
  switch (ax.n) {
    case 64: return cols_n<R, 64>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    case 128: return cols_n<R, 128>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    case 256: return cols_n<R, 256>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    case 512: return cols_n<R, 512>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    case 1024: return cols_n<R, 1024>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    case 2048: return cols_n<R, 2048>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fast column pass: n = %lld", (long long)ax.n);
  }
}

int fast_cols_c2c(int prec, const void* in, void* out, int64_t ncols, int64_t stride, int64_t nslabs,
                  int64_t slab_dist, const Axis& ax, int dir, double scale, hipStream_t st) {
  return prec == FFHIP_F32 ? cols_t<float>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st)
                           : cols_t<double>(in, out, ncols, stride, nslabs, slab_dist, ax, dir, scale, st);
}

}  // namespace ffhip
