// Shared helpers for libffhip: error reporting, complex arithmetic, precision traits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "ffhip.h"

namespace ffhip {

// ----------------------------------------------------------------- errors
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define FF_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::ffhip::fail(FFHIP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define FF_CHECK(cond, code, ...)                 \
  do {                                            \
    if (!(cond)) return ::ffhip::fail(code, __VA_ARGS__); \
  } while (0)

#define FF_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != FFHIP_OK) return rc_; \
  } while (0)

// Launch check: reports the error of the most recent launch.
#define FF_LAUNCHED(name)                                                          \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess)                                                          \
      return ::ffhip::fail(FFHIP_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e_)); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device allocation that reports OOM as FFHIP_ERR_OOM.
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
int max_lds_per_block();

// ----------------------------------------------------------------- complex
template <typename R> struct ctype;
template <> struct ctype<float> { using type = float2; };
template <> struct ctype<double> { using type = double2; };
template <typename R> using cplx = typename ctype<R>::type;

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mkc(R x, R y) {
  cplx<R> c;
  c.x = x;
  c.y = y;
  return c;
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return mkc<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return mkc<R>(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cconj(cplx<R> a) { return mkc<R>(a.x, -a.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return mkc<R>(a.x * s, a.y * s); }
// a * (S i), S = +1 or -1
template <int S, typename R>
__host__ __device__ __forceinline__ cplx<R> cmul_si(cplx<R> a) { return mkc<R>(-S * a.y, S * a.x); }
// twiddle for direction DIR (-1 forward uses the table value, +1 its conjugate)
template <int DIR, typename R>
__device__ __forceinline__ cplx<R> twid(cplx<R> w) { return DIR < 0 ? w : cconj<R>(w); }

template <typename R> constexpr int prec_code();
template <> constexpr int prec_code<float>() { return FFHIP_F32; }
template <> constexpr int prec_code<double>() { return FFHIP_F64; }

inline size_t real_size(int prec) { return prec == FFHIP_F32 ? 4 : 8; }
inline size_t cplx_size(int prec) { return prec == FFHIP_F32 ? 8 : 16; }
inline bool valid_prec(int prec) { return prec == FFHIP_F32 || prec == FFHIP_F64; }

inline int64_t prod(const int64_t* d, int n) {
  int64_t p = 1;
  for (int i = 0; i < n; ++i) p *= d[i];
  return p;
}

inline unsigned blocks_for(int64_t n, int threads, int64_t cap = 65535LL * 64) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace ffhip
