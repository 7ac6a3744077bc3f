// Pointwise passes of the reference's time steppers, dealiasing, and the
// compensated Parseval reductions.
//
// ETDRK4substep12!/substep3!/update! (src/timesteppers.jl:501-516) and the RK4
// pieces (:225-264) keep the reference's operation order; arithmetic is done
// in FP64 (coefficients are FP64) and rounded once to the state precision.
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "reduce.h"
#include "stepper.h"

namespace ffhip {

constexpr int kPwThreads = 256;

template <typename R> __device__ __forceinline__ double2 ld2(const cplx<R>* p, int64_t i) {
  const cplx<R> v = p[i];
  return make_double2((double)v.x, (double)v.y);
}
template <typename R> __device__ __forceinline__ void st2(cplx<R>* p, int64_t i, double2 v) {
  p[i] = mkc<R>((R)v.x, (R)v.y);
}
__device__ __forceinline__ double2 d2add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 d2mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 d2scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// coefficient c[i] times z: real (CC = false) or complex (CC = true) coefficient
template <bool CC> __device__ __forceinline__ double2 cmulz(const void* c, int64_t i, double2 z) {
  if (CC) return d2mul(reinterpret_cast<const double2*>(c)[i], z);
  return d2scale(z, reinterpret_cast<const double*>(c)[i]);
}

#define FF_GRID_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// out = E½·sol + ζ·N
template <typename R, bool CC>
__global__ void k_etd_sub12(cplx<R>* out, const cplx<R>* sol, const cplx<R>* N, const void* Eh, const void* Z,
                            int64_t n) {
  FF_GRID_LOOP(i, n) { st2<R>(out, i, d2add(cmulz<CC>(Eh, i, ld2<R>(sol, i)), cmulz<CC>(Z, i, ld2<R>(N, i)))); }
}

// out = E½·sol1 + ζ·(2N3 − N1)
template <typename R, bool CC>
__global__ void k_etd_sub3(cplx<R>* out, const cplx<R>* sol1, const cplx<R>* N1, const cplx<R>* N3, const void* Eh,
                           const void* Z, int64_t n) {
  FF_GRID_LOOP(i, n) {
    // 2N3 - N1 is formed in the state type, as the reference's broadcast does
    const cplx<R> n1 = N1[i], n3 = N3[i];
    const double2 d = make_double2((double)(R(2) * n3.x - n1.x), (double)(R(2) * n3.y - n1.y));
    st2<R>(out, i, d2add(cmulz<CC>(Eh, i, ld2<R>(sol1, i)), cmulz<CC>(Z, i, d)));
  }
}

// sol = E·sol + α·N1 + 2β·(N2 + N3) + Γ·N4
template <typename R, bool CC>
__global__ void k_etd_update(cplx<R>* sol, const cplx<R>* N1, const cplx<R>* N2, const cplx<R>* N3,
                             const cplx<R>* N4, const void* E, const void* A, const void* B, const void* G,
                             int64_t n) {
  FF_GRID_LOOP(i, n) {
    double2 acc = d2add(cmulz<CC>(E, i, ld2<R>(sol, i)), cmulz<CC>(A, i, ld2<R>(N1, i)));
    // N2 + N3 is formed in the state type, as the reference's broadcast does
    const cplx<R> n2 = N2[i], n3 = N3[i];
    const double2 s23 = make_double2((double)(n2.x + n3.x), (double)(n2.y + n3.y));
    acc = d2add(acc, d2scale(cmulz<CC>(B, i, s23), 2.0));
    acc = d2add(acc, cmulz<CC>(G, i, ld2<R>(N4, i)));
    st2<R>(sol, i, acc);
  }
}

template <typename R, bool CC>
__global__ void k_rk4_addlinear(cplx<R>* RHS, const cplx<R>* sol, const void* L, int64_t n) {
  FF_GRID_LOOP(i, n) { st2<R>(RHS, i, d2add(ld2<R>(RHS, i), cmulz<CC>(L, i, ld2<R>(sol, i)))); }
}

template <typename R>
__global__ void k_rk4_substepsol(cplx<R>* out, const cplx<R>* sol, const cplx<R>* RHS, double dt, int64_t n) {
  FF_GRID_LOOP(i, n) { st2<R>(out, i, d2add(ld2<R>(sol, i), d2scale(ld2<R>(RHS, i), dt))); }
}

template <typename R>
__global__ void k_rk4_update(cplx<R>* sol, const cplx<R>* R1, const cplx<R>* R2, const cplx<R>* R3,
                             const cplx<R>* R4, double dt6, int64_t n) {
  FF_GRID_LOOP(i, n) {
    double2 s = d2add(ld2<R>(R1, i), d2scale(ld2<R>(R2, i), 2.0));
    s = d2add(s, d2scale(ld2<R>(R3, i), 2.0));
    s = d2add(s, ld2<R>(R4, i));
    st2<R>(sol, i, d2add(ld2<R>(sol, i), d2scale(s, dt6)));
  }
}

// sol *= filter (FP64 filter, rounded once to the state type)
template <typename R>
__global__ void k_filter(cplx<R>* sol, const double* f, int64_t n) {
  FF_GRID_LOOP(i, n) { st2<R>(sol, i, d2scale(ld2<R>(sol, i), f[i])); }
}

// ForwardEuler (src/timesteppers.jl:112-118): sol += dt (L sol + N);
// filtered (:142-149): sol = filter (sol + dt (N + L sol))
template <typename R, bool CC>
__global__ void k_euler(cplx<R>* sol, const cplx<R>* N, const void* L, const double* f, double dt, int64_t n) {
  FF_GRID_LOOP(i, n) {
    const double2 u = ld2<R>(sol, i);
    double2 r = d2add(u, d2scale(d2add(cmulz<CC>(L, i, u), ld2<R>(N, i)), dt));
    if (f) r = d2scale(r, f[i]);
    st2<R>(sol, i, r);
  }
}

// AB3update! (src/timesteppers.jl:630-638): forward Euler while step < 3, then
// sol += dt (23/12 RHS - 16/12 RHS_-1 + 5/12 RHS_-2); then the filter if any
template <typename R>
__global__ void k_ab3(cplx<R>* sol, const cplx<R>* R0, const cplx<R>* R1, const cplx<R>* R2, const double* f,
                      double dt, int euler, int64_t n) {
  FF_GRID_LOOP(i, n) {
    double2 r;
    if (euler) {
      r = d2scale(ld2<R>(R0, i), dt);
    } else {
      r = d2add(d2scale(ld2<R>(R0, i), 23.0 / 12.0), d2scale(ld2<R>(R1, i), -(16.0 / 12.0)));
      r = d2scale(d2add(r, d2scale(ld2<R>(R2, i), 5.0 / 12.0)), dt);
    }
    cplx<R> u = sol[i];
    u = mkc<R>((R)((double)u.x + r.x), (R)((double)u.y + r.y));
    if (f) u = mkc<R>((R)((double)u.x * f[i]), (R)((double)u.y * f[i]));
    sol[i] = u;
  }
}

// LSRK54 stage (src/timesteppers.jl:389-397): S2 = A S2 + dt RHS; sol += B S2.
// A == 0 (the first stage, after S2 = 0) does not read S2.
template <typename R>
__global__ void k_lsrk54(cplx<R>* sol, cplx<R>* S2, const cplx<R>* RHS, double A, double B, double dt, int64_t n) {
  FF_GRID_LOOP(i, n) {
    const double2 r = d2scale(ld2<R>(RHS, i), dt);
    const double2 s = A == 0.0 ? r : d2add(d2scale(ld2<R>(S2, i), A), r);
    st2<R>(S2, i, s);
    const double2 sr = ld2<R>(S2, i);  // sol += B S2 with S2 in the state type
    st2<R>(sol, i, d2add(ld2<R>(sol, i), d2scale(sr, B)));
  }
}

// zero every entry whose index along some dimension lies in that dimension's
// aliased range [lo, hi] (0-based, inclusive; empty when lo > hi)
template <typename R>
__global__ void k_dealias(cplx<R>* f, int64_t d0, int64_t d1, int64_t d2, int64_t total, int64_t lo0, int64_t hi0,
                          int64_t lo1, int64_t hi1, int64_t lo2, int64_t hi2) {
  FF_GRID_LOOP(idx, total) {
    const int64_t i0 = idx % d0;
    int64_t r = idx / d0;
    const int64_t i1 = r % d1;
    r /= d1;
    const int64_t i2 = r % d2;
    if ((i0 >= lo0 && i0 <= hi0) || (i1 >= lo1 && i1 <= hi1) || (i2 >= lo2 && i2 <= hi2)) f[idx] = mkc<R>(0, 0);
  }
}

static unsigned pw_blocks(int64_t n) { return blocks_for(n, kPwThreads, 2048 * 8); }

#define FF_DISPATCH_RC(prec, cc, KERNEL, ...)                                                      \
  do {                                                                                             \
    if (prec == FFHIP_F32) {                                                                       \
      if (cc) KERNEL(float, true, __VA_ARGS__); else KERNEL(float, false, __VA_ARGS__);            \
    } else {                                                                                       \
      if (cc) KERNEL(double, true, __VA_ARGS__); else KERNEL(double, false, __VA_ARGS__);          \
    }                                                                                              \
  } while (0)

int etd_substep12(const ffhip_etdrk4* ts, void* out, const void* sol, const void* N, hipStream_t st) {
  FF_CHECK(ts && out && sol && N, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
#define L12(R, CC, ...)                                                                                         \
  hipLaunchKernelGGL((k_etd_sub12<R, CC>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (cplx<R>*)out,           \
                     (const cplx<R>*)sol, (const cplx<R>*)N, ts->coef[5], ts->coef[0], n)
  FF_DISPATCH_RC(ts->prec, ts->cplx_coef, L12, 0);
#undef L12
  FF_LAUNCHED("k_etd_sub12");
  return FFHIP_OK;
}

int etd_substep3(const ffhip_etdrk4* ts, void* out, const void* sol1, const void* N1, const void* N3,
                 hipStream_t st) {
  FF_CHECK(ts && out && sol1 && N1 && N3, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
#define L3(R, CC, ...)                                                                                         \
  hipLaunchKernelGGL((k_etd_sub3<R, CC>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (cplx<R>*)out,           \
                     (const cplx<R>*)sol1, (const cplx<R>*)N1, (const cplx<R>*)N3, ts->coef[5], ts->coef[0], n)
  FF_DISPATCH_RC(ts->prec, ts->cplx_coef, L3, 0);
#undef L3
  FF_LAUNCHED("k_etd_sub3");
  return FFHIP_OK;
}

int etd_update(const ffhip_etdrk4* ts, void* sol, const void* N1, const void* N2, const void* N3, const void* N4,
               hipStream_t st) {
  FF_CHECK(ts && sol && N1 && N2 && N3 && N4, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
#define LU(R, CC, ...)                                                                                         \
  hipLaunchKernelGGL((k_etd_update<R, CC>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (cplx<R>*)sol,         \
                     (const cplx<R>*)N1, (const cplx<R>*)N2, (const cplx<R>*)N3, (const cplx<R>*)N4, ts->coef[4],  \
                     ts->coef[1], ts->coef[2], ts->coef[3], n)
  FF_DISPATCH_RC(ts->prec, ts->cplx_coef, LU, 0);
#undef LU
  FF_LAUNCHED("k_etd_update");
  return FFHIP_OK;
}

int rk4_addlinear(const ffhip_rk4* ts, void* RHS, const void* sol, hipStream_t st) {
  FF_CHECK(ts && RHS && sol, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
#define LA(R, CC, ...)                                                                                         \
  hipLaunchKernelGGL((k_rk4_addlinear<R, CC>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (cplx<R>*)RHS,      \
                     (const cplx<R>*)sol, ts->L, n)
  FF_DISPATCH_RC(ts->prec, ts->cplx_coef, LA, 0);
#undef LA
  FF_LAUNCHED("k_rk4_addlinear");
  return FFHIP_OK;
}

int rk4_substepsol(const ffhip_rk4* ts, void* newsol, const void* sol, const void* RHS, double dt,
                   hipStream_t st) {
  FF_CHECK(ts && newsol && sol && RHS, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
  if (ts->prec == FFHIP_F32)
    hipLaunchKernelGGL((k_rk4_substepsol<float>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (float2*)newsol,
                       (const float2*)sol, (const float2*)RHS, dt, n);
  else
    hipLaunchKernelGGL((k_rk4_substepsol<double>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (double2*)newsol,
                       (const double2*)sol, (const double2*)RHS, dt, n);
  FF_LAUNCHED("k_rk4_substepsol");
  return FFHIP_OK;
}

int rk4_update(const ffhip_rk4* ts, void* sol, const void* R1, const void* R2, const void* R3, const void* R4,
               double dt, hipStream_t st) {
  FF_CHECK(ts && sol && R1 && R2 && R3 && R4, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
  if (ts->prec == FFHIP_F32)
    hipLaunchKernelGGL((k_rk4_update<float>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (float2*)sol,
                       (const float2*)R1, (const float2*)R2, (const float2*)R3, (const float2*)R4, dt / 6, n);
  else
    hipLaunchKernelGGL((k_rk4_update<double>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (double2*)sol,
                       (const double2*)R1, (const double2*)R2, (const double2*)R3, (const double2*)R4, dt / 6, n);
  FF_LAUNCHED("k_rk4_update");
  return FFHIP_OK;
}

int dealias_launch(void* fh, int prec, int ndim, const int64_t* dims, int64_t batch, const int64_t* lo,
                   const int64_t* hi, hipStream_t st) {
  FF_CHECK(fh && dims && lo && hi, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(ndim >= 1 && ndim <= 3, FFHIP_ERR_ARG, "ndim must be 1..3");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  FF_CHECK(batch >= 1, FFHIP_ERR_ARG, "batch must be >= 1");
  int64_t d[3] = {1, 1, 1}, l[3] = {1, 1, 1}, h[3] = {0, 0, 0};
  for (int i = 0; i < ndim; ++i) {
    d[i] = dims[i];
    l[i] = lo[i] - 1;  // 1-based -> 0-based
    h[i] = hi[i] - 1;
  }
  const int64_t total = d[0] * d[1] * d[2] * batch;
  if (total == 0) return FFHIP_OK;
  if (prec == FFHIP_F32)
    hipLaunchKernelGGL((k_dealias<float>), dim3(pw_blocks(total)), dim3(kPwThreads), 0, st, (float2*)fh, d[0], d[1],
                       d[2], total, l[0], h[0], l[1], h[1], l[2], h[2]);
  else
    hipLaunchKernelGGL((k_dealias<double>), dim3(pw_blocks(total)), dim3(kPwThreads), 0, st, (double2*)fh, d[0],
                       d[1], d[2], total, l[0], h[0], l[1], h[1], l[2], h[2]);
  FF_LAUNCHED("k_dealias");
  return FFHIP_OK;
}

// ----------------------------------------------------------------- reductions
// MODE 0: sum w*|u|^2; MODE 1: sum w*Re(u).  w = 2 except at x indices 0 and nx/2
// (only when rfft != 0).
template <typename R, int MODE>
__global__ void __launch_bounds__(kRedThreads) k_parseval(const cplx<R>* __restrict__ u, int64_t total, int64_t d0,
                                                          int rfft, int64_t nyq, double2* __restrict__ partial) {
  double2 acc = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const cplx<R> v = u[i];
    double2 c;
    if (MODE == 0) {
      c = dd_add(dd_sq((double)v.x), dd_sq((double)v.y));
    } else {
      c = make_double2((double)v.x, 0.0);
    }
    if (rfft) {
      const int64_t ix = i % d0;
      if (ix != 0 && ix != nyq) c = make_double2(2 * c.x, 2 * c.y);
    }
    acc = dd_add(acc, c);
  }
  acc = block_dd_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(kRedThreads) k_dd_final(const double2* __restrict__ partial, int n,
                                                          double2* __restrict__ out) {
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = dd_add(acc, partial[i]);
  acc = block_dd_sum(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

static std::mutex g_red_mu;
static double2* g_red_buf = nullptr;  // kRedBlocks partials + 1 result

int red_buffer(double2** p) {
  std::lock_guard<std::mutex> lk(g_red_mu);
  if (!g_red_buf) {
    void* q = nullptr;
    FF_TRY(dev_alloc(&q, (kRedBlocks + 1) * sizeof(double2)));
    g_red_buf = reinterpret_cast<double2*>(q);
  }
  *p = g_red_buf;
  return FFHIP_OK;
}

int parseval_launch(int mode, const void* uh, int prec, int ndim, const int64_t* dims, int rfft, int64_t nx,
                    double norm, double* out, hipStream_t st) {
  FF_CHECK(uh && dims && out, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(ndim >= 1 && ndim <= 4, FFHIP_ERR_ARG, "ndim must be 1..4");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  const int64_t total = prod(dims, ndim);
  if (rfft) FF_CHECK(dims[0] == nx / 2 + 1, FFHIP_ERR_ARG, "rfft layout needs dims[0] == nx/2+1");
  double2* buf = nullptr;
  FF_TRY(red_buffer(&buf));
  int blocks = (int)((total + kRedThreads - 1) / kRedThreads);
  if (blocks > kRedBlocks) blocks = kRedBlocks;
  if (blocks < 1) blocks = 1;
  const int64_t nyq = rfft ? dims[0] - 1 : -1;
#define LP(R, M)                                                                                                  \
  hipLaunchKernelGGL((k_parseval<R, M>), dim3(blocks), dim3(kRedThreads), 0, st, (const cplx<R>*)uh, total, dims[0], \
                     rfft, nyq, buf)
  if (prec == FFHIP_F32) {
    if (mode == 0) LP(float, 0); else LP(float, 1);
  } else {
    if (mode == 0) LP(double, 0); else LP(double, 1);
  }
#undef LP
  FF_LAUNCHED("k_parseval");
  hipLaunchKernelGGL(k_dd_final, dim3(1), dim3(kRedThreads), 0, st, buf, blocks, buf + kRedBlocks);
  FF_LAUNCHED("k_dd_final");
  double2 h;
  FF_HIP(hipMemcpyAsync(&h, buf + kRedBlocks, sizeof h, hipMemcpyDeviceToHost, st));
  FF_HIP(hipStreamSynchronize(st));
  *out = (h.x + h.y) * norm;
  if (!std::isfinite(*out)) return fail(FFHIP_ERR_NONFINITE, "non-finite value in reduction");
  return FFHIP_OK;
}

// ----------------------------------------------------------------- allocation-free reductions
// ffhip_parseval_noalloc keeps its partial sums in this fixed array instead of
// a library allocation: kRedBlocks double-double partials (16 KiB), one copy
// per GPU that the library runs on, held until the process ends.  In host code
// the name is a host-side stand-in, not a device address, so it is never
// passed to a kernel: red_scratch_setup() looks up each device's address once
// (hipGetSymbolAddress, that device current) and every launch takes that
// pointer.  Each launch writes partials 0..blocks-1 before the final kernel
// reads them, so nothing needs clearing between calls; calls that overlap on
// one GPU (concurrent streams or threads) race on the array.
static __device__ __align__(256) double2 scratch[kRedBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double2* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

int red_scratch_setup(int device) {
  FF_CHECK(device >= 0 && device < kMaxDevices, FFHIP_ERR_ARG, "device %d beyond the scratch table (%d)", device,
           kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return FFHIP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(FFHIP_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
  g_scratch[device] = reinterpret_cast<double2*>(p);
  return FFHIP_OK;
}

// (acc.x + acc.y) * norm of the partials, written to device memory
__global__ void __launch_bounds__(kRedThreads) k_dd_final_scaled(const double2* __restrict__ partial, int n,
                                                                 double norm, double* __restrict__ out) {
#pragma clang fp contract(off)
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = dd_add(acc, partial[i]);
  acc = block_dd_sum(acc);
  if (threadIdx.x == 0) out[0] = (acc.x + acc.y) * norm;
}

// ----------------------------------------------------------------- steppers: create/destroy
static int upload_f64(void** d, const double* h, int64_t count) {
  FF_TRY(dev_alloc(d, (size_t)count * sizeof(double)));
  FF_HIP(hipMemcpy(*d, h, (size_t)count * sizeof(double), hipMemcpyHostToDevice));
  return FFHIP_OK;
}

static int upload_coef(void** d, const double* re, const double* im, int64_t count) {
  if (!im) return upload_f64(d, re, count);
  std::vector<double> z(2 * count);
  for (int64_t i = 0; i < count; ++i) {
    z[2 * i] = re[i];
    z[2 * i + 1] = im[i];
  }
  return upload_f64(d, z.data(), 2 * count);
}

int apply_filter(int prec, int64_t count, const void* filter, void* sol, hipStream_t st) {
  if (!filter) return FFHIP_OK;
  if (prec == FFHIP_F32)
    hipLaunchKernelGGL((k_filter<float>), dim3(pw_blocks(count)), dim3(kPwThreads), 0, st, (float2*)sol,
                       (const double*)filter, count);
  else
    hipLaunchKernelGGL((k_filter<double>), dim3(pw_blocks(count)), dim3(kPwThreads), 0, st, (double2*)sol,
                       (const double*)filter, count);
  FF_LAUNCHED("k_filter");
  return FFHIP_OK;
}

int euler_update(const ffhip_rk4* ts, void* sol, const void* N, double dt, hipStream_t st) {
  FF_CHECK(ts && sol && N, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
#define LE(R, CC, ...)                                                                                     \
  hipLaunchKernelGGL((k_euler<R, CC>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (cplx<R>*)sol,          \
                     (const cplx<R>*)N, ts->L, (const double*)ts->filter, dt, n)
  FF_DISPATCH_RC(ts->prec, ts->cplx_coef, LE, 0);
#undef LE
  FF_LAUNCHED("k_euler");
  return FFHIP_OK;
}

int ab3_update(const ffhip_rk4* ts, void* sol, const void* RHS, const void* RHSm1, const void* RHSm2, double dt,
               int64_t step, hipStream_t st) {
  FF_CHECK(ts && sol && RHS && RHSm1 && RHSm2, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
  const int euler = step < 3;
  if (ts->prec == FFHIP_F32)
    hipLaunchKernelGGL((k_ab3<float>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (float2*)sol,
                       (const float2*)RHS, (const float2*)RHSm1, (const float2*)RHSm2, (const double*)ts->filter,
                       dt, euler, n);
  else
    hipLaunchKernelGGL((k_ab3<double>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (double2*)sol,
                       (const double2*)RHS, (const double2*)RHSm1, (const double2*)RHSm2,
                       (const double*)ts->filter, dt, euler, n);
  FF_LAUNCHED("k_ab3");
  return FFHIP_OK;
}

int lsrk54_stage(const ffhip_rk4* ts, void* sol, void* S2, const void* RHS, double A, double B, double dt,
                 hipStream_t st) {
  FF_CHECK(ts && sol && S2 && RHS, FFHIP_ERR_ARG, "null argument");
  const int64_t n = ts->count;
  if (ts->prec == FFHIP_F32)
    hipLaunchKernelGGL((k_lsrk54<float>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (float2*)sol, (float2*)S2,
                       (const float2*)RHS, A, B, dt, n);
  else
    hipLaunchKernelGGL((k_lsrk54<double>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, st, (double2*)sol,
                       (double2*)S2, (const double2*)RHS, A, B, dt, n);
  FF_LAUNCHED("k_lsrk54");
  return FFHIP_OK;
}

// scratch arrays are allocated on first use (callers that bring their own
// arrays, and the fused operators, never touch them)
int etd_scratch(ffhip_etdrk4* t) {
  for (int s = 0; s < 6; ++s)
    if (!t->scratch[s]) FF_TRY(dev_alloc(&t->scratch[s], (size_t)t->count * cplx_size(t->prec)));
  return FFHIP_OK;
}

int rk4_scratch(ffhip_rk4* t) {
  for (int s = 0; s < 5; ++s)
    if (!t->scratch[s]) FF_TRY(dev_alloc(&t->scratch[s], (size_t)t->count * cplx_size(t->prec)));
  return FFHIP_OK;
}

}  // namespace ffhip

using namespace ffhip;

static void etd_free(ffhip_etdrk4* t) {
  if (!t) return;
  for (auto& p : t->coef) dev_free(p);
  for (auto& p : t->scratch) dev_free(p);
  dev_free(t->packed);
  dev_free(t->filter);
  delete[] t->L_host_re;
  delete[] t->L_host_im;
  delete t;
}

static void rk4_free(ffhip_rk4* t) {
  if (!t) return;
  dev_free(t->L);
  dev_free(t->filter);
  for (auto& p : t->scratch) dev_free(p);
  delete t;
}

static int check_dims(int ndim, const int64_t* dims) {
  FF_CHECK(ndim >= 1 && ndim <= 3, FFHIP_ERR_ARG, "ndim must be 1..3 (got %d)", ndim);
  FF_CHECK(dims, FFHIP_ERR_ARG, "null dims");
  for (int i = 0; i < ndim; ++i) FF_CHECK(dims[i] >= 1, FFHIP_ERR_ARG, "dims must be positive");
  return FFHIP_OK;
}

extern "C" {

int ffhip_dealias(void* fh, int prec, int ndim, const int64_t* dims, int64_t batch, const int64_t* lo,
                  const int64_t* hi, void* stream) {
  clear_error();
  return dealias_launch(fh, prec, ndim, dims, batch, lo, hi, as_stream(stream));
}

int ffhip_parsevalsum2(const void* uh, int prec, int ndim, const int64_t* dims, int rfft, int64_t nx, double norm,
                       double* out, void* stream) {
  clear_error();
  return parseval_launch(0, uh, prec, ndim, dims, rfft, nx, norm, out, as_stream(stream));
}

int ffhip_parsevalsum(const void* uh, int prec, int ndim, const int64_t* dims, int rfft, int64_t nx, double norm,
                      double* out, void* stream) {
  clear_error();
  return parseval_launch(1, uh, prec, ndim, dims, rfft, nx, norm, out, as_stream(stream));
}

int ffhip_parseval_noalloc(int mode, const void* uh, int prec, int ndim, const int64_t* dims, int rfft, int64_t nx,
                           double norm, double* out_dev, void* stream) {
  clear_error();
  FF_CHECK(mode == 0 || mode == 1, FFHIP_ERR_ARG, "mode must be 0 (parsevalsum2) or 1 (parsevalsum)");
  FF_CHECK(uh && dims && out_dev, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(ndim >= 1 && ndim <= 4, FFHIP_ERR_ARG, "ndim must be 1..4");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  for (int i = 0; i < ndim; ++i) FF_CHECK(dims[i] >= 1, FFHIP_ERR_ARG, "dims[%d] = %lld", i, (long long)dims[i]);
  const int64_t total = prod(dims, ndim);
  if (rfft) FF_CHECK(dims[0] == nx / 2 + 1, FFHIP_ERR_ARG, "rfft layout needs dims[0] == nx/2+1");
  int dev = 0;
  FF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev]) FF_TRY(red_scratch_setup(dev));  // ffhip_init not run here
  double2* part = g_scratch[dev];
  // grid-stride partial sums: one per workgroup, at most the kRedBlocks the scratch array holds
  const int64_t want = (total + kRedThreads - 1) / kRedThreads;
  const int blocks = (int)(want < kRedBlocks ? want : kRedBlocks);
  FF_CHECK(blocks >= 1 && blocks <= kRedBlocks, FFHIP_ERR_ARG, "%d partial sums exceed the scratch array (%d)", blocks,
           kRedBlocks);
  hipStream_t st = as_stream(stream);
  const int64_t nyq = rfft ? dims[0] - 1 : -1;
#define LP(R, M)                                                                                                  \
  hipLaunchKernelGGL((k_parseval<R, M>), dim3(blocks), dim3(kRedThreads), 0, st, (const cplx<R>*)uh, total, dims[0], \
                     rfft, nyq, part)
  if (prec == FFHIP_F32) {
    if (mode == 0) LP(float, 0); else LP(float, 1);
  } else {
    if (mode == 0) LP(double, 0); else LP(double, 1);
  }
#undef LP
  FF_LAUNCHED("k_parseval");
  hipLaunchKernelGGL(k_dd_final_scaled, dim3(1), dim3(kRedThreads), 0, st, part, blocks, norm, out_dev);
  FF_LAUNCHED("k_dd_final_scaled");
  return FFHIP_OK;
}

int ffhip_etdrk4_create(ffhip_etdrk4** out, const double* Lre, const double* Lim, int ndim, const int64_t* dims,
                        double dt, int prec) {
  clear_error();
  FF_CHECK(out, FFHIP_ERR_ARG, "null output pointer");
  *out = nullptr;
  FF_TRY(check_dims(ndim, dims));
  FF_CHECK(Lre, FFHIP_ERR_ARG, "null L");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  ffhip_etdrk4* t = new (std::nothrow) ffhip_etdrk4();
  FF_CHECK(t, FFHIP_ERR_OOM, "host allocation failed");
  t->ndim = ndim;
  for (int i = 0; i < ndim; ++i) t->dims[i] = dims[i];
  t->count = prod(dims, ndim);
  t->prec = prec;
  t->cplx_coef = Lim != nullptr;
  // dt = fltype(equation.T)(dt)  (src/timesteppers.jl:457)
  t->dt = prec == FFHIP_F32 ? (double)(float)dt : dt;
  const int64_t n = t->count;
  std::vector<double> zr(n), ar(n), br(n), gr(n), zi, ai, bi, gi;
  if (Lim) {
    zi.resize(n);
    ai.resize(n);
    bi.resize(n);
    gi.resize(n);
  }
  int rc = ffhip_etdcoeffs(t->dt, Lre, Lim, n, 32, 1.0, zr.data(), ar.data(), br.data(), gr.data(),
                           Lim ? zi.data() : nullptr, Lim ? ai.data() : nullptr, Lim ? bi.data() : nullptr,
                           Lim ? gi.data() : nullptr);
  // getexpLs (:673-678): exp(dt L), exp(dt L / 2)
  std::vector<double> er(n), ehr(n), ei, ehi;
  if (Lim) {
    ei.resize(n);
    ehi.resize(n);
  }
  for (int64_t i = 0; i < n; ++i) {
    if (Lim) {
      std::complex<double> hl(t->dt * Lre[i], t->dt * Lim[i]);
      std::complex<double> e = std::exp(hl), eh = std::exp(hl / 2.0);
      er[i] = e.real();
      ei[i] = e.imag();
      ehr[i] = eh.real();
      ehi[i] = eh.imag();
    } else {
      er[i] = std::exp(t->dt * Lre[i]);
      ehr[i] = std::exp(t->dt * Lre[i] / 2);
    }
  }
  const double* re[6] = {zr.data(), ar.data(), br.data(), gr.data(), er.data(), ehr.data()};
  const double* im[6] = {Lim ? zi.data() : nullptr, Lim ? ai.data() : nullptr, Lim ? bi.data() : nullptr,
                         Lim ? gi.data() : nullptr, Lim ? ei.data() : nullptr, Lim ? ehi.data() : nullptr};
  for (int c = 0; c < 6 && rc == FFHIP_OK; ++c) rc = upload_coef(&t->coef[c], re[c], im[c], n);
  if (rc == FFHIP_OK) {
    t->L_host_re = new (std::nothrow) double[n];
    if (!t->L_host_re) rc = fail(FFHIP_ERR_OOM, "host allocation failed");
    else memcpy(t->L_host_re, Lre, n * sizeof(double));
    if (rc == FFHIP_OK && Lim) {
      t->L_host_im = new (std::nothrow) double[n];
      if (!t->L_host_im) rc = fail(FFHIP_ERR_OOM, "host allocation failed");
      else memcpy(t->L_host_im, Lim, n * sizeof(double));
    }
  }
  if (rc != FFHIP_OK) {
    etd_free(t);
    return rc;
  }
  *out = t;
  return FFHIP_OK;
}

int ffhip_etdrk4_destroy(ffhip_etdrk4* ts) {
  clear_error();
  etd_free(ts);
  return FFHIP_OK;
}

int ffhip_etdrk4_scratch(ffhip_etdrk4* ts, int which, void** ptr) {
  clear_error();
  FF_CHECK(ts && ptr, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(which >= 0 && which < 6, FFHIP_ERR_ARG, "scratch index %d out of range", which);
  FF_TRY(etd_scratch(ts));
  *ptr = ts->scratch[which];
  return FFHIP_OK;
}

int ffhip_etdrk4_coeff(ffhip_etdrk4* ts, int which, void** ptr, int* is_complex) {
  clear_error();
  FF_CHECK(ts && ptr, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(which >= 0 && which < 6, FFHIP_ERR_ARG, "coefficient index %d out of range", which);
  *ptr = ts->coef[which];
  if (is_complex) *is_complex = ts->cplx_coef;
  return FFHIP_OK;
}

int ffhip_etdrk4_substep12(ffhip_etdrk4* ts, void* out, const void* sol, const void* N, void* stream) {
  clear_error();
  return etd_substep12(ts, out, sol, N, as_stream(stream));
}

int ffhip_etdrk4_substep3(ffhip_etdrk4* ts, void* out, const void* sol1, const void* N1, const void* N3,
                          void* stream) {
  clear_error();
  return etd_substep3(ts, out, sol1, N1, N3, as_stream(stream));
}

int ffhip_etdrk4_update(ffhip_etdrk4* ts, void* sol, const void* N1, const void* N2, const void* N3, const void* N4,
                        void* stream) {
  clear_error();
  return etd_update(ts, sol, N1, N2, N3, N4, as_stream(stream));
}

int ffhip_rk4_create(ffhip_rk4** out, const double* Lre, const double* Lim, int ndim, const int64_t* dims,
                     int prec) {
  clear_error();
  FF_CHECK(out, FFHIP_ERR_ARG, "null output pointer");
  *out = nullptr;
  FF_TRY(check_dims(ndim, dims));
  FF_CHECK(Lre, FFHIP_ERR_ARG, "null L");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  ffhip_rk4* t = new (std::nothrow) ffhip_rk4();
  FF_CHECK(t, FFHIP_ERR_OOM, "host allocation failed");
  t->ndim = ndim;
  for (int i = 0; i < ndim; ++i) t->dims[i] = dims[i];
  t->count = prod(dims, ndim);
  t->prec = prec;
  t->cplx_coef = Lim != nullptr;
  int rc = upload_coef(&t->L, Lre, Lim, t->count);
  if (rc != FFHIP_OK) {
    rk4_free(t);
    return rc;
  }
  *out = t;
  return FFHIP_OK;
}

int ffhip_rk4_destroy(ffhip_rk4* ts) {
  clear_error();
  rk4_free(ts);
  return FFHIP_OK;
}

int ffhip_rk4_scratch(ffhip_rk4* ts, int which, void** ptr) {
  clear_error();
  FF_CHECK(ts && ptr, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(which >= 0 && which < 5, FFHIP_ERR_ARG, "scratch index %d out of range", which);
  FF_TRY(rk4_scratch(ts));
  *ptr = ts->scratch[which];
  return FFHIP_OK;
}

int ffhip_rk4_addlinearterm(ffhip_rk4* ts, void* RHS, const void* sol, void* stream) {
  clear_error();
  return rk4_addlinear(ts, RHS, sol, as_stream(stream));
}

int ffhip_rk4_substepsol(ffhip_rk4* ts, void* newsol, const void* sol, const void* RHS, double dt, void* stream) {
  clear_error();
  return rk4_substepsol(ts, newsol, sol, RHS, dt, as_stream(stream));
}

int ffhip_rk4_update(ffhip_rk4* ts, void* sol, const void* R1, const void* R2, const void* R3, const void* R4,
                     double dt, void* stream) {
  clear_error();
  return rk4_update(ts, sol, R1, R2, R3, R4, dt, as_stream(stream));
}

static int set_filter(void** slot, const double* filter, int64_t count) {
  dev_free(*slot);
  *slot = nullptr;
  if (!filter) return FFHIP_OK;
  for (int64_t i = 0; i < count; ++i)
    FF_CHECK(std::isfinite(filter[i]), FFHIP_ERR_ARG, "filter entry %lld is not finite", (long long)i);
  return upload_f64(slot, filter, count);
}

int ffhip_etdrk4_set_filter(ffhip_etdrk4* ts, const double* filter) {
  clear_error();
  FF_CHECK(ts, FFHIP_ERR_ARG, "null stepper");
  FF_TRY(set_filter(&ts->filter, filter, ts->count));
  ts->packed_key = 0;  // fused paths repack their coefficient copies
  return FFHIP_OK;
}

int ffhip_rk4_set_filter(ffhip_rk4* ts, const double* filter) {
  clear_error();
  FF_CHECK(ts, FFHIP_ERR_ARG, "null stepper");
  return set_filter(&ts->filter, filter, ts->count);
}

int ffhip_etdrk4_apply_filter(ffhip_etdrk4* ts, void* sol, void* stream) {
  clear_error();
  FF_CHECK(ts && sol, FFHIP_ERR_ARG, "null argument");
  return apply_filter(ts->prec, ts->count, ts->filter, sol, as_stream(stream));
}

int ffhip_rk4_apply_filter(ffhip_rk4* ts, void* sol, void* stream) {
  clear_error();
  FF_CHECK(ts && sol, FFHIP_ERR_ARG, "null argument");
  return apply_filter(ts->prec, ts->count, ts->filter, sol, as_stream(stream));
}

int ffhip_euler_update(ffhip_rk4* ts, void* sol, const void* N, double dt, void* stream) {
  clear_error();
  return euler_update(ts, sol, N, dt, as_stream(stream));
}

int ffhip_ab3_update(ffhip_rk4* ts, void* sol, const void* RHS, const void* RHSm1, const void* RHSm2, double dt,
                     int64_t step, void* stream) {
  clear_error();
  return ab3_update(ts, sol, RHS, RHSm1, RHSm2, dt, step, as_stream(stream));
}

int ffhip_lsrk54_stage(ffhip_rk4* ts, void* sol, void* S2, const void* RHS, double A, double B, double dt,
                       void* stream) {
  clear_error();
  return lsrk54_stage(ts, sol, S2, RHS, A, B, dt, as_stream(stream));
}

}  // extern "C"
