// FFT plans: r2c/c2r and c2c transforms over 1-3 dimensions, batched.
//
// mul!(out, plan, in)  (src/domains.jl:2-5)  -> ffhip_plan_forward
// ldiv!(out, plan, in) (AbstractFFTs scaled inverse) -> ffhip_plan_inverse
//
// Execution: the x axis is transformed per row (contiguous), the y and z
// axes per column tile (C adjacent columns loaded as C-wide contiguous runs).
// A transform whose ping-pong tile fits the LDS budget runs in one kernel per
// axis (Stockham in LDS); longer ones run one global-memory pass per radix.
// The r2c x transform packs the n reals as n/2 complex values, runs an n/2
// complex FFT and splits the result with the exp(-2 pi i k/n) twiddles; c2r
// mirrors it and ignores the imaginary parts of the DC and Nyquist entries.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "fft_device.h"
#include "plan.h"
#include "prof.h"

namespace ffhip {

// ----------------------------------------------------------------- host helpers
static const long double kPiL = 3.141592653589793238462643383279502884L;

// (cos, sin)(2 pi k / n) with exact values at multiples of n/8 symmetry points.
void unit_root(int64_t k, int64_t n, double* c, double* s) {
  k %= n;
  if (k < 0) k += n;
  // quadrant q and remainder: 4k = q n + rem
  __int128 k4 = (__int128)4 * k;
  int q = (int)(k4 / n);
  int64_t rem = (int64_t)(k4 - (__int128)q * n);  // angle within quadrant = (pi/2) rem/n
  long double cc, ss;
  if (rem == 0) {
    cc = 1.0L;
    ss = 0.0L;
  } else if (2 * rem == n) {
    cc = ss = 0.70710678118654752440084436210484903928L;
  } else if (2 * rem < n) {
    long double a = kPiL / 2 * (long double)rem / (long double)n;
    cc = cosl(a);
    ss = sinl(a);
  } else {
    long double a = kPiL / 2 * (long double)(n - rem) / (long double)n;
    cc = sinl(a);
    ss = cosl(a);
  }
  long double C, S;
  switch (q & 3) {
    case 0: C = cc; S = ss; break;
    case 1: C = -ss; S = cc; break;
    case 2: C = -cc; S = -ss; break;
    default: C = ss; S = -cc; break;
  }
  *c = (double)C;
  *s = (double)S;
}

std::vector<int> factor_radices(int64_t n) {
  std::vector<int> r;
  int64_t m = n;
  while (m % 8 == 0) { r.push_back(8); m /= 8; }
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 5 == 0) { r.push_back(5); m /= 5; }
  for (int64_t p = 7; m > 1; p += 2) {
    while (m % p == 0) { r.push_back((int)p); m /= p; }
    if (p * p > m && m > 1) { r.push_back((int)m); m = 1; }
  }
  return r;
}

// Table exp(-2 pi i k / n) for k in [0, count), uploaded in precision prec.
int upload_roots(int64_t n, int64_t count, int prec, void** dptr) {
  std::vector<double> h(2 * count);
  for (int64_t k = 0; k < count; ++k) {
    double c, s;
    unit_root(k, n, &c, &s);
    h[2 * k] = c;
    h[2 * k + 1] = -s;
  }
  size_t bytes = (size_t)count * cplx_size(prec);
  FF_TRY(dev_alloc(dptr, bytes));
  if (prec == FFHIP_F64) {
    FF_HIP(hipMemcpy(*dptr, h.data(), bytes, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(2 * count);
    for (int64_t i = 0; i < 2 * count; ++i) f[i] = (float)h[i];
    FF_HIP(hipMemcpy(*dptr, f.data(), bytes, hipMemcpyHostToDevice));
  }
  return FFHIP_OK;
}

// ----------------------------------------------------------------- LDS kernels
constexpr int kThreads = 256;

// c2c along contiguous rows (x axis of the complex plan).
template <int DIR, typename R>
__global__ void __launch_bounds__(kThreads) k_rows_c2c(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                      Radices rd, const cplx<R>* __restrict__ tw, int64_t nrows,
                                                      int nb, R scale) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  const int n = rd.n;
  cplx<R>* b = a + nb * n;
  const int64_t row0 = (int64_t)blockIdx.x * nb;
  const int nloc = (int)((nrows - row0) < nb ? (nrows - row0) : nb);
  for (int i = threadIdx.x; i < nb * n; i += blockDim.x) {
    const int t = i / n;
    a[i] = t < nloc ? in[(row0 + t) * n + (i - t * n)] : mkc<R>(0, 0);
  }
  __syncthreads();
  cplx<R>* res = stockham_lds<DIR, R>(a, b, rd, tw, nb, n, 1, false);
  for (int i = threadIdx.x; i < nloc * n; i += blockDim.x) out[row0 * n + i] = cscale<R>(res[i], scale);
}

// r2c along contiguous rows: real (n = 2m) -> complex (m + 1).
template <typename R>
__global__ void __launch_bounds__(kThreads) k_rows_r2c(const R* __restrict__ in, cplx<R>* __restrict__ out,
                                                      Radices rd, const cplx<R>* __restrict__ tw,
                                                      const cplx<R>* __restrict__ rtw, int64_t nrows, int nb) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  const int m = rd.n;
  cplx<R>* b = a + nb * m;
  const int64_t row0 = (int64_t)blockIdx.x * nb;
  const int nloc = (int)((nrows - row0) < nb ? (nrows - row0) : nb);
  const cplx<R>* inc = reinterpret_cast<const cplx<R>*>(in);  // pairs (x[2j], x[2j+1])
  for (int i = threadIdx.x; i < nb * m; i += blockDim.x) {
    const int t = i / m;
    a[i] = t < nloc ? inc[(row0 + t) * m + (i - t * m)] : mkc<R>(0, 0);
  }
  __syncthreads();
  cplx<R>* Z = stockham_lds<-1, R>(a, b, rd, tw, nb, m, 1, false);
  const int nk = m + 1;
  for (int i = threadIdx.x; i < nloc * nk; i += blockDim.x) {
    const int t = i / nk;
    const int k = i - t * nk;
    const cplx<R> zk = Z[t * m + (k == m ? 0 : k)];
    const cplx<R> zc = cconj<R>(Z[t * m + (k == 0 ? 0 : m - k)]);
    const cplx<R> xe = cscale<R>(cadd<R>(zk, zc), R(0.5));
    const cplx<R> d = csub<R>(zk, zc);
    const cplx<R> xo = mkc<R>(R(0.5) * d.y, -R(0.5) * d.x);  // (zk - zc) / (2i)
    out[(row0 + t) * nk + k] = cadd<R>(xe, cmul<R>(rtw[k], xo));
  }
}

// c2r along contiguous rows: complex (m + 1) -> real (2m), scaled.
template <typename R>
__global__ void __launch_bounds__(kThreads) k_rows_c2r(const cplx<R>* __restrict__ in, R* __restrict__ out,
                                                      Radices rd, const cplx<R>* __restrict__ tw,
                                                      const cplx<R>* __restrict__ rtw, int64_t nrows, int nb,
                                                      R scale) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  const int m = rd.n;
  cplx<R>* b = a + nb * m;
  const int nk = m + 1;
  const int64_t row0 = (int64_t)blockIdx.x * nb;
  const int nloc = (int)((nrows - row0) < nb ? (nrows - row0) : nb);
  for (int i = threadIdx.x; i < nb * m; i += blockDim.x) {
    const int t = i / m;
    const int k = i - t * m;
    cplx<R> z = mkc<R>(0, 0);
    if (t < nloc) {
      const cplx<R>* row = in + (row0 + t) * nk;
      cplx<R> xk = row[k];
      cplx<R> xm = row[m - k];
      if (k == 0) {  // imaginary parts of DC and Nyquist are ignored
        xk.y = 0;
        xm.y = 0;
      }
      const cplx<R> xmc = cconj<R>(xm);
      const cplx<R> xe = cscale<R>(cadd<R>(xk, xmc), R(0.5));
      const cplx<R> xo = cmul<R>(cscale<R>(csub<R>(xk, xmc), R(0.5)), cconj<R>(rtw[k]));
      z = mkc<R>(xe.x - xo.y, xe.y + xo.x);  // xe + i xo
    }
    a[i] = z;
  }
  __syncthreads();
  cplx<R>* res = stockham_lds<1, R>(a, b, rd, tw, nb, m, 1, false);
  cplx<R>* outc = reinterpret_cast<cplx<R>*>(out);
  for (int i = threadIdx.x; i < nloc * m; i += blockDim.x) outc[row0 * m + i] = cscale<R>(res[i], scale);
}

// c2c along a strided axis: transform (slab s, column c) has element i at
// s*slab_dist + c + i*stride.  A block takes C adjacent columns of one slab.
template <int DIR, typename R>
__global__ void __launch_bounds__(kThreads) k_cols_c2c(const cplx<R>* in, cplx<R>* out,  // may alias (in place)
                                                      Radices rd, const cplx<R>* __restrict__ tw, int64_t ncols,
                                                      int64_t stride, int64_t slab_dist, int C, R scale) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  const int n = rd.n;
  cplx<R>* b = a + C * n;
  const int64_t col0 = (int64_t)blockIdx.x * C;
  const int64_t base = (int64_t)blockIdx.y * slab_dist + col0;
  const int cl = (int)((ncols - col0) < C ? (ncols - col0) : C);
  for (int i = threadIdx.x; i < C * n; i += blockDim.x) {
    const int c = i % C;
    const int e = i / C;
    a[i] = c < cl ? in[base + c + e * stride] : mkc<R>(0, 0);
  }
  __syncthreads();
  cplx<R>* res = stockham_lds<DIR, R>(a, b, rd, tw, C, 1, C, true);
  for (int i = threadIdx.x; i < C * n; i += blockDim.x) {
    const int c = i % C;
    const int e = i / C;
    if (c < cl) out[base + c + e * stride] = cscale<R>(res[i], scale);
  }
}

// ----------------------------------------------------------------- global-memory path
// One radix pass over all transforms.  Transform t = (slab, col): element i
// of src at slab*s.slab + col*s.col + i*s.elem (same form for dst).
struct Layout {
  int64_t ncols, col, slab, elem;
};

__device__ __forceinline__ int64_t lay_addr(const Layout& L, int64_t t, int64_t i) {
  const int64_t sl = t / L.ncols;
  const int64_t c = t - sl * L.ncols;
  return sl * L.slab + c * L.col + i * L.elem;
}

template <int RAD, int DIR, typename R>
__global__ void __launch_bounds__(kThreads) k_pass_global(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst,
                                                         Layout ls, Layout ld, int64_t ntrans, int64_t n, int p,
                                                         int64_t Ns, const cplx<R>* __restrict__ tw, R scale,
                                                         bool tfast) {
  const int64_t m = n / p;
  const int64_t total = ntrans * m;
  const int64_t tstep = n / (Ns * p);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t t, j;
    if (tfast) {
      t = idx % ntrans;
      j = idx / ntrans;
    } else {
      t = idx / m;
      j = idx - t * m;
    }
    const int64_t k = j % Ns;
    const int64_t o = (j - k) * p + k;
    if constexpr (RAD > 0) {
      cplx<R> v[RAD];
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        v[r] = src[lay_addr(ls, t, j + r * m)];
        if (Ns > 1 && r > 0) v[r] = cmul<R>(v[r], twid<DIR, R>(tw[r * k * tstep]));
      }
      Bfly<RAD, DIR, R>::run(v);
#pragma unroll
      for (int r = 0; r < RAD; ++r) dst[lay_addr(ld, t, o + r * Ns)] = cscale<R>(v[r], scale);
    } else {
      const int64_t wstep = n / p;
      for (int q = 0; q < p; ++q) {
        cplx<R> acc = mkc<R>(0, 0);
        for (int r = 0; r < p; ++r) {
          cplx<R> x = src[lay_addr(ls, t, j + r * m)];
          if (Ns > 1 && r > 0) x = cmul<R>(x, twid<DIR, R>(tw[r * k * tstep]));
          const int e = (r * q) % p;
          acc = cadd<R>(acc, e == 0 ? x : cmul<R>(x, twid<DIR, R>(tw[e * wstep])));
        }
        dst[lay_addr(ld, t, o + q * Ns)] = cscale<R>(acc, scale);
      }
    }
  }
}

template <typename R>
__global__ void k_gather(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst, Layout ls, Layout ld,
                         int64_t ntrans, int64_t n) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < ntrans * n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = idx / n, i = idx - t * n;
    dst[lay_addr(ld, t, i)] = src[lay_addr(ls, t, i)];
  }
}

// r2c split / c2r merge for the global path (rows of m complex <-> m+1).
template <typename R>
__global__ void k_r2c_split(const cplx<R>* __restrict__ Z, cplx<R>* __restrict__ out,
                            const cplx<R>* __restrict__ rtw, int64_t nrows, int64_t m) {
  const int64_t nk = m + 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows * nk;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / nk, k = i - t * nk;
    const cplx<R> zk = Z[t * m + (k == m ? 0 : k)];
    const cplx<R> zc = cconj<R>(Z[t * m + (k == 0 ? 0 : m - k)]);
    const cplx<R> xe = cscale<R>(cadd<R>(zk, zc), R(0.5));
    const cplx<R> d = csub<R>(zk, zc);
    const cplx<R> xo = mkc<R>(R(0.5) * d.y, -R(0.5) * d.x);
    out[i] = cadd<R>(xe, cmul<R>(rtw[k], xo));
  }
}

template <typename R>
__global__ void k_c2r_merge(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ Z,
                            const cplx<R>* __restrict__ rtw, int64_t nrows, int64_t m) {
  const int64_t nk = m + 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows * m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / m, k = i - t * m;
    cplx<R> xk = in[t * nk + k];
    cplx<R> xm = in[t * nk + (m - k)];
    if (k == 0) {
      xk.y = 0;
      xm.y = 0;
    }
    const cplx<R> xmc = cconj<R>(xm);
    const cplx<R> xe = cscale<R>(cadd<R>(xk, xmc), R(0.5));
    const cplx<R> xo = cmul<R>(cscale<R>(csub<R>(xk, xmc), R(0.5)), cconj<R>(rtw[k]));
    Z[i] = mkc<R>(xe.x - xo.y, xe.y + xo.x);
  }
}

// ----------------------------------------------------------------- dispatch helpers
static int lds_budget() {
  int m = max_lds_per_block();
  return m < 65536 ? m : 65536;
}

template <typename R>
static int global_axis(const cplx<R>* src, cplx<R>* dst, Layout ls, Layout ld, int64_t ntrans, const Axis& ax,
                       int dir, R scale, cplx<R>* w0, cplx<R>* w1, hipStream_t st) {
  // passes: src -> w0 -> w1 -> ... -> dst (the last pass writes dst)
  const int64_t n = ax.n;
  const int nst = (int)ax.rad.size();
  if (nst == 0) {  // n == 1: copy with scale through a radix-less pass
    return fail(FFHIP_ERR_UNSUPPORTED, "length-1 axis on the global path");
  }
  Layout lw{1, 0, n, 1};  // contiguous [t][n]
  const cplx<R>* cur = src;
  if (nst == 1 && (const void*)src == (const void*)dst) {
    // a single pass cannot run in place: stage the input in w1 first
    prof_launch("fft_gather", 2.0 * ntrans * n * sizeof(cplx<R>), k_gather<R>,
                dim3(blocks_for(ntrans * n, kThreads, 1 << 20)), dim3(kThreads), 0, st, src, w1, ls, lw, ntrans, n);
    FF_LAUNCHED("k_gather");
    cur = w1;
    ls = lw;
  }
  Layout lc = ls;
  int64_t Ns = 1;
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(ax.tw);
  for (int s = 0; s < nst; ++s) {
    const int p = ax.rad[s];
    const bool last = s == nst - 1;
    cplx<R>* out = last ? dst : (s % 2 == 0 ? w0 : w1);
    Layout lo = last ? ld : lw;
    const R sc = last ? scale : R(1);
    const int64_t total = ntrans * (n / p);
    const unsigned blocks = blocks_for(total, kThreads, 1 << 20);
    const bool tfast = lc.elem != 1;
    const double pbytes = 2.0 * ntrans * n * sizeof(cplx<R>);
#define FF_GPASS(RAD)                                                                                   \
  if (dir < 0)                                                                                          \
    prof_launch("fft_pass_global", pbytes, k_pass_global<RAD, -1, R>, dim3(blocks), dim3(kThreads), 0, st, \
                cur, out, lc, lo, ntrans, n, p, Ns, tw, sc, tfast);                                     \
  else                                                                                                  \
    prof_launch("fft_pass_global", pbytes, k_pass_global<RAD, 1, R>, dim3(blocks), dim3(kThreads), 0, st,  \
                cur, out, lc, lo, ntrans, n, p, Ns, tw, sc, tfast);
    switch (p) {
      case 2: FF_GPASS(2); break;
      case 3: FF_GPASS(3); break;
      case 4: FF_GPASS(4); break;
      case 5: FF_GPASS(5); break;
      case 8: FF_GPASS(8); break;
      default: FF_GPASS(0); break;
    }
#undef FF_GPASS
    FF_LAUNCHED("k_pass_global");
    cur = out;
    lc = lo;
    Ns *= p;
  }
  return FFHIP_OK;
}

static Radices to_radices(const Axis& ax) {
  Radices r;
  r.n = (int)ax.n;
  r.nst = (int)ax.rad.size();
  for (int i = 0; i < r.nst; ++i) r.r[i] = ax.rad[i];
  return r;
}

// c2c of `ntrans` contiguous rows of length ax.n
template <typename R>
static int run_rows_c2c(const cplx<R>* in, cplx<R>* out, int64_t nrows, const Axis& ax, int dir, R scale,
                        ffhip_plan* p, hipStream_t st) {
  const int n = (int)ax.n;
  if (ax.lds) {
    int nb = ax.nb;
    size_t sh = 2 * (size_t)nb * n * sizeof(cplx<R>);
    unsigned blocks = (unsigned)((nrows + nb - 1) / nb);
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(ax.tw);
    const double bytes = 2.0 * nrows * n * sizeof(cplx<R>);
    if (dir < 0)
      prof_launch("fft_rows_c2c", bytes, k_rows_c2c<-1, R>, dim3(blocks), dim3(kThreads), sh, st, in, out,
                  to_radices(ax), tw, nrows, nb, scale);
    else
      prof_launch("fft_rows_c2c", bytes, k_rows_c2c<1, R>, dim3(blocks), dim3(kThreads), sh, st, in, out,
                  to_radices(ax), tw, nrows, nb, scale);
    FF_LAUNCHED("k_rows_c2c");
    return FFHIP_OK;
  }
  Layout L{1, 0, ax.n, 1};
  return global_axis<R>(in, out, L, L, nrows, ax, dir, scale, reinterpret_cast<cplx<R>*>(p->g0),
                        reinterpret_cast<cplx<R>*>(p->g1), st);
}

template <typename R>
static int run_cols_c2c(const cplx<R>* in, cplx<R>* out, int64_t ncols, int64_t stride, int64_t nslabs,
                        int64_t slab_dist, const Axis& ax, int dir, R scale, ffhip_plan* p, hipStream_t st) {
  if (ax.fast) return fast_cols_c2c(p->prec, in, out, ncols, stride, nslabs, slab_dist, ax, dir, (double)scale, st);
  const int n = (int)ax.n;
  if (ax.lds) {
    const int C = (int)(ax.nb < ncols ? ax.nb : ncols);
    size_t sh = 2 * (size_t)C * n * sizeof(cplx<R>);
    dim3 grid((unsigned)((ncols + C - 1) / C), (unsigned)nslabs);
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(ax.tw);
    const double bytes = 2.0 * ncols * nslabs * n * sizeof(cplx<R>);
    if (dir < 0)
      prof_launch("fft_cols_c2c", bytes, k_cols_c2c<-1, R>, grid, dim3(kThreads), sh, st, in, out, to_radices(ax),
                  tw, ncols, stride, slab_dist, C, scale);
    else
      prof_launch("fft_cols_c2c", bytes, k_cols_c2c<1, R>, grid, dim3(kThreads), sh, st, in, out, to_radices(ax),
                  tw, ncols, stride, slab_dist, C, scale);
    FF_LAUNCHED("k_cols_c2c");
    return FFHIP_OK;
  }
  Layout L{ncols, 1, slab_dist, stride};
  return global_axis<R>(in, out, L, L, ncols * nslabs, ax, dir, scale, reinterpret_cast<cplx<R>*>(p->g0),
                        reinterpret_cast<cplx<R>*>(p->g1), st);
}

template <typename R>
static int run_rows_r2c(const R* in, cplx<R>* out, int64_t nrows, ffhip_plan* p, hipStream_t st) {
  const Axis& ax = p->ax[0];
  if (ax.fast) return fast_rows_r2c(p->prec, in, out, nrows, ax, p->rtw, st);
  const int m = (int)ax.n;
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(ax.tw);
  const cplx<R>* rtw = reinterpret_cast<const cplx<R>*>(p->rtw);
  if (ax.lds) {
    const int nb = ax.nb;
    size_t sh = 2 * (size_t)nb * m * sizeof(cplx<R>);
    unsigned blocks = (unsigned)((nrows + nb - 1) / nb);
    prof_launch("fft_rows_r2c", (double)nrows * (2.0 * m * sizeof(R) + (m + 1) * sizeof(cplx<R>)), k_rows_r2c<R>,
                dim3(blocks), dim3(kThreads), sh, st, in, out, to_radices(ax), tw, rtw, nrows, nb);
    FF_LAUNCHED("k_rows_r2c");
    return FFHIP_OK;
  }
  cplx<R>* Z = reinterpret_cast<cplx<R>*>(p->g2);
  Layout L{1, 0, ax.n, 1};
  FF_TRY(global_axis<R>(reinterpret_cast<const cplx<R>*>(in), Z, L, L, nrows, ax, -1, R(1),
                        reinterpret_cast<cplx<R>*>(p->g0), reinterpret_cast<cplx<R>*>(p->g1), st));
  prof_launch("fft_r2c_split", (double)nrows * (2 * m + 1) * sizeof(cplx<R>), k_r2c_split<R>,
              dim3(blocks_for(nrows * (m + 1), kThreads, 1 << 20)), dim3(kThreads), 0, st, Z, out, rtw, nrows,
              (int64_t)m);
  FF_LAUNCHED("k_r2c_split");
  return FFHIP_OK;
}

template <typename R>
static int run_rows_c2r(const cplx<R>* in, R* out, int64_t nrows, R scale, ffhip_plan* p, hipStream_t st) {
  const Axis& ax = p->ax[0];
  if (ax.fast) return fast_rows_c2r(p->prec, in, out, nrows, ax, p->rtw, (double)scale, st);
  const int m = (int)ax.n;
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(ax.tw);
  const cplx<R>* rtw = reinterpret_cast<const cplx<R>*>(p->rtw);
  if (ax.lds) {
    const int nb = ax.nb;
    size_t sh = 2 * (size_t)nb * m * sizeof(cplx<R>);
    unsigned blocks = (unsigned)((nrows + nb - 1) / nb);
    prof_launch("fft_rows_c2r", (double)nrows * (2.0 * m * sizeof(R) + (m + 1) * sizeof(cplx<R>)), k_rows_c2r<R>,
                dim3(blocks), dim3(kThreads), sh, st, in, out, to_radices(ax), tw, rtw, nrows, nb, scale);
    FF_LAUNCHED("k_rows_c2r");
    return FFHIP_OK;
  }
  cplx<R>* Z = reinterpret_cast<cplx<R>*>(p->g2);
  prof_launch("fft_c2r_merge", (double)nrows * (2 * m + 1) * sizeof(cplx<R>), k_c2r_merge<R>,
              dim3(blocks_for(nrows * m, kThreads, 1 << 20)), dim3(kThreads), 0, st, in, Z, rtw, nrows, (int64_t)m);
  FF_LAUNCHED("k_c2r_merge");
  Layout L{1, 0, ax.n, 1};
  return global_axis<R>(Z, reinterpret_cast<cplx<R>*>(out), L, L, nrows, ax, 1, scale,
                        reinterpret_cast<cplx<R>*>(p->g0), reinterpret_cast<cplx<R>*>(p->g1), st);
}

// ----------------------------------------------------------------- plan execution
template <typename R>
static int exec_forward(ffhip_plan* p, const void* vin, void* vout, hipStream_t st) {
  const int64_t* n = p->n;
  const int nd = p->ndim;
  const int64_t nx0 = p->kind == 0 ? n[0] / 2 + 1 : n[0];  // complex x extent
  const int64_t rest = (nd > 1 ? n[1] : 1) * (nd > 2 ? n[2] : 1) * p->batch;
  cplx<R>* out = reinterpret_cast<cplx<R>*>(vout);
  if (p->kind == 0) {
    FF_TRY(run_rows_r2c<R>(reinterpret_cast<const R*>(vin), out, rest, p, st));
  } else {
    FF_TRY(run_rows_c2c<R>(reinterpret_cast<const cplx<R>*>(vin), out, rest, p->ax[0], -1, R(1), p, st));
  }
  if (nd >= 2)
    FF_TRY(run_cols_c2c<R>(out, out, nx0, nx0, (nd > 2 ? n[2] : 1) * p->batch, nx0 * n[1], p->ax[1], -1, R(1), p,
                           st));
  if (nd >= 3)
    FF_TRY(run_cols_c2c<R>(out, out, nx0 * n[1], nx0 * n[1], p->batch, nx0 * n[1] * n[2], p->ax[2], -1, R(1), p,
                           st));
  return FFHIP_OK;
}

// extra: factor on top of the 1/prod(n) normalisation (prod(n) gives the unnormalised inverse)
template <typename R>
static int exec_inverse(ffhip_plan* p, const void* vin, void* vout, hipStream_t st, double extra = 1.0) {
  const int64_t* n = p->n;
  const int nd = p->ndim;
  const int64_t nx0 = p->kind == 0 ? n[0] / 2 + 1 : n[0];
  const int64_t rest = (nd > 1 ? n[1] : 1) * (nd > 2 ? n[2] : 1) * p->batch;
  double total = 1;
  for (int d = 0; d < nd; ++d) total *= (double)n[d];
  const cplx<R>* cur = reinterpret_cast<const cplx<R>*>(vin);
  cplx<R>* ws = reinterpret_cast<cplx<R>*>(p->ws);
  if (p->kind == 1) {
    // complex plan: z, y into ws (input preserved), x last with the scale
    cplx<R>* out = reinterpret_cast<cplx<R>*>(vout);
    if (nd >= 3) {
      FF_TRY(run_cols_c2c<R>(cur, ws, nx0 * n[1], nx0 * n[1], p->batch, nx0 * n[1] * n[2], p->ax[2], 1, R(1), p,
                             st));
      cur = ws;
    }
    if (nd >= 2) {
      FF_TRY(run_cols_c2c<R>(cur, ws, nx0, nx0, (nd > 2 ? n[2] : 1) * p->batch, nx0 * n[1], p->ax[1], 1, R(1), p,
                             st));
      cur = ws;
    }
    return run_rows_c2c<R>(cur, out, rest, p->ax[0], 1, R(extra / total), p, st);
  }
  if (nd >= 3) {
    FF_TRY(run_cols_c2c<R>(cur, ws, nx0 * n[1], nx0 * n[1], p->batch, nx0 * n[1] * n[2], p->ax[2], 1, R(1), p,
                           st));
    cur = ws;
  }
  if (nd >= 2) {
    FF_TRY(run_cols_c2c<R>(cur, ws, nx0, nx0, (nd > 2 ? n[2] : 1) * p->batch, nx0 * n[1], p->ax[1], 1, R(1), p,
                           st));
    cur = ws;
  }
  // x: the split into even/odd halves already carries a factor 1/2, so the
  // normalised inverse along x is (1/m) * (unnormalised length-m inverse);
  // the other axes contribute 1/prod(n[1:]).
  const double m = (double)(n[0] / 2);
  const double other = total / (double)n[0];
  return run_rows_c2r<R>(cur, reinterpret_cast<R*>(vout), rest, R(extra / (m * other)), p, st);
}

// out = in * (edge on the kx = 0 and kx = nx/2 columns, mid elsewhere), rfft layout [rest][nx0]
template <typename R>
__global__ void k_rfft_weight(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int64_t count,
                              int64_t nx0, R edge, R mid) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kx = i % nx0;
    const R s = (kx == 0 || kx == nx0 - 1) ? edge : mid;
    const cplx<R> v = in[i];
    out[i] = mkc<R>(v.x * s, v.y * s);
  }
}

static int64_t spec_entries(const ffhip_plan* p) {
  int64_t c = p->kind == 0 ? p->n[0] / 2 + 1 : p->n[0];
  for (int d = 1; d < p->ndim; ++d) c *= p->n[d];
  return c * p->batch;
}

static double plan_total(const ffhip_plan* p) {
  double t = 1;
  for (int d = 0; d < p->ndim; ++d) t *= (double)p->n[d];
  return t;
}

// ldiv!^T(phi) = W * mul!(phi) / prod(n)  (W = 1 at kx = 0, nx/2, else 2; c2c: W = 1)
template <typename R>
static int exec_inverse_adjoint(ffhip_plan* p, const void* in, void* out, hipStream_t st) {
  FF_TRY(exec_forward<R>(p, in, out, st));
  const double t = plan_total(p);
  const int64_t cnt = spec_entries(p);
  const R edge = (R)(1.0 / t), mid = (R)((p->kind == 0 ? 2.0 : 1.0) / t);
  hipLaunchKernelGGL((k_rfft_weight<R>), dim3(blocks_for(cnt, 256, 8192)), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(out), reinterpret_cast<cplx<R>*>(out), cnt,
                     p->kind == 0 ? p->n[0] / 2 + 1 : p->n[0], edge, mid);
  FF_LAUNCHED("k_rfft_weight");
  return FFHIP_OK;
}

// mul!^T(Y) = prod(n) * ldiv!(Y / W)
template <typename R>
static int exec_forward_adjoint(ffhip_plan* p, const void* in, void* out, hipStream_t st) {
  const double t = plan_total(p);
  if (p->kind == 1) return exec_inverse<R>(p, in, out, st, t);
  const int64_t cnt = spec_entries(p);
  if (!p->adj) FF_TRY(dev_alloc(&p->adj, (size_t)cnt * sizeof(cplx<R>)));
  hipLaunchKernelGGL((k_rfft_weight<R>), dim3(blocks_for(cnt, 256, 8192)), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(in), reinterpret_cast<cplx<R>*>(p->adj), cnt,
                     p->n[0] / 2 + 1, R(1), R(0.5));
  FF_LAUNCHED("k_rfft_weight");
  return exec_inverse<R>(p, p->adj, out, st, t);
}

// ----------------------------------------------------------------- plan creation
static int setup_axis(Axis& ax, int64_t n, int prec, int64_t ntrans_hint, bool force_global) {
  ax.n = n;
  ax.rad = factor_radices(n);
  FF_CHECK((int)ax.rad.size() <= kMaxStages, FFHIP_ERR_UNSUPPORTED, "too many radix stages for n=%lld",
           (long long)n);
  FF_CHECK(n <= (1LL << 30), FFHIP_ERR_UNSUPPORTED, "axis length %lld too large", (long long)n);
  FF_TRY(upload_roots(n, n, prec, &ax.tw));
  const size_t csz = cplx_size(prec);
  const int budget = lds_budget();
  const size_t per = 2 * (size_t)n * csz;  // ping-pong tile of one transform
  ax.lds = !force_global && n >= 1 && per <= (size_t)budget;
  if (ax.lds) {
    int64_t cap = (int64_t)(budget / per);
    int64_t want = (2048 + n - 1) / n;  // ~2048 elements per block
    if (want < 8 && ntrans_hint >= 8 && cap >= 8) want = 8;
    int64_t nb = want < cap ? want : cap;
    if (nb < 1) nb = 1;
    if (nb > 64) nb = 64;
    ax.nb = (int)nb;
  }
  return FFHIP_OK;
}

static void destroy_plan(ffhip_plan* p) {
  if (!p) return;
  for (auto& a : p->ax) dev_free(a.tw);
  dev_free(p->rtw);
  dev_free(p->ws);
  dev_free(p->g0);
  dev_free(p->g1);
  dev_free(p->g2);
  dev_free(p->adj);
  delete p;
}

static int create_plan(ffhip_plan** out, int kind, int ndim, const int64_t* n, int64_t batch, int prec) {
  FF_CHECK(out, FFHIP_ERR_ARG, "plan output pointer is NULL");
  *out = nullptr;
  FF_CHECK(ndim >= 1 && ndim <= 3, FFHIP_ERR_ARG, "ndim must be 1, 2 or 3 (got %d)", ndim);
  FF_CHECK(n, FFHIP_ERR_ARG, "size array is NULL");
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "precision must be 32 or 64 (got %d)", prec);
  FF_CHECK(batch >= 1, FFHIP_ERR_ARG, "batch must be >= 1");
  for (int d = 0; d < ndim; ++d) FF_CHECK(n[d] >= 1, FFHIP_ERR_ARG, "sizes must be positive");
  if (kind == 0) FF_CHECK(n[0] % 2 == 0, FFHIP_ERR_DOMAIN, "nx must be even for a real transform");
  ffhip_plan* p = new (std::nothrow) ffhip_plan();
  FF_CHECK(p, FFHIP_ERR_OOM, "host allocation failed");
  p->kind = kind;
  p->ndim = ndim;
  p->batch = batch;
  p->prec = prec;
  for (int d = 0; d < 3; ++d) p->n[d] = d < ndim ? n[d] : 1;
  const bool force_global = getenv("FFHIP_FORCE_GLOBAL_FFT") && atoi(getenv("FFHIP_FORCE_GLOBAL_FFT")) != 0;
  const int64_t nx0 = kind == 0 ? n[0] / 2 + 1 : n[0];
  const int64_t rows = (ndim > 1 ? n[1] : 1) * (ndim > 2 ? n[2] : 1) * batch;
  int rc = FFHIP_OK;
  const int64_t xlen = kind == 0 ? n[0] / 2 : n[0];
  rc = setup_axis(p->ax[0], xlen, prec, rows, force_global);
  p->ax[0].fast = !force_global && kind == 0 && fast_rows_ok(xlen);
  if (rc == FFHIP_OK && kind == 0) rc = upload_roots(n[0], n[0] / 2 + 1, prec, &p->rtw);
  for (int d = 1; d < ndim && rc == FFHIP_OK; ++d) {
    rc = setup_axis(p->ax[d], n[d], prec, nx0, force_global);
    p->ax[d].fast = !force_global && fast_cols_ok(n[d]);
  }
  const int64_t spec = nx0 * (ndim > 1 ? n[1] : 1) * (ndim > 2 ? n[2] : 1) * batch;
  const size_t sbytes = (size_t)spec * cplx_size(prec);
  if (rc == FFHIP_OK && ndim > 1) rc = dev_alloc(&p->ws, sbytes);
  bool any_global = false;
  for (int d = 0; d < ndim; ++d) any_global |= !p->ax[d].lds;
  if (rc == FFHIP_OK && any_global) rc = dev_alloc(&p->g0, sbytes);
  if (rc == FFHIP_OK && any_global) rc = dev_alloc(&p->g1, sbytes);
  if (rc == FFHIP_OK && any_global && kind == 0) rc = dev_alloc(&p->g2, sbytes);
  if (rc != FFHIP_OK) {
    destroy_plan(p);
    return rc;
  }
  (void)rows;
  *out = p;
  return FFHIP_OK;
}

int plan_inverse_scaled(ffhip_plan* p, const void* in, void* out, double extra, hipStream_t st) {
  return p->prec == FFHIP_F32 ? exec_inverse<float>(p, in, out, st, extra)
                              : exec_inverse<double>(p, in, out, st, extra);
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_plan_rfft(ffhip_plan** plan, int ndim, const int64_t* n, int64_t batch, int prec) {
  clear_error();
  return create_plan(plan, 0, ndim, n, batch, prec);
}

int ffhip_plan_fft(ffhip_plan** plan, int ndim, const int64_t* n, int64_t batch, int prec) {
  clear_error();
  return create_plan(plan, 1, ndim, n, batch, prec);
}

int ffhip_plan_destroy(ffhip_plan* plan) {
  clear_error();
  destroy_plan(plan);
  return FFHIP_OK;
}

int ffhip_plan_forward(ffhip_plan* p, const void* in, void* out, void* stream) {
  clear_error();
  FF_CHECK(p && in && out, FFHIP_ERR_ARG, "null plan or buffer");
  FF_CHECK(in != out, FFHIP_ERR_ARG, "in-place transforms are not supported");
  return p->prec == FFHIP_F32 ? exec_forward<float>(p, in, out, as_stream(stream))
                              : exec_forward<double>(p, in, out, as_stream(stream));
}

int ffhip_plan_inverse(ffhip_plan* p, const void* in, void* out, void* stream) {
  clear_error();
  FF_CHECK(p && in && out, FFHIP_ERR_ARG, "null plan or buffer");
  FF_CHECK(in != out, FFHIP_ERR_ARG, "in-place transforms are not supported");
  return p->prec == FFHIP_F32 ? exec_inverse<float>(p, in, out, as_stream(stream))
                              : exec_inverse<double>(p, in, out, as_stream(stream));
}

int ffhip_plan_forward_adjoint(ffhip_plan* p, const void* in, void* out, void* stream) {
  clear_error();
  FF_CHECK(p && in && out, FFHIP_ERR_ARG, "null plan or buffer");
  FF_CHECK(in != out, FFHIP_ERR_ARG, "in-place transforms are not supported");
  return p->prec == FFHIP_F32 ? exec_forward_adjoint<float>(p, in, out, as_stream(stream))
                              : exec_forward_adjoint<double>(p, in, out, as_stream(stream));
}

int ffhip_plan_inverse_adjoint(ffhip_plan* p, const void* in, void* out, void* stream) {
  clear_error();
  FF_CHECK(p && in && out, FFHIP_ERR_ARG, "null plan or buffer");
  FF_CHECK(in != out, FFHIP_ERR_ARG, "in-place transforms are not supported");
  return p->prec == FFHIP_F32 ? exec_inverse_adjoint<float>(p, in, out, as_stream(stream))
                              : exec_inverse_adjoint<double>(p, in, out, as_stream(stream));
}

int ffhip_plan_info(const ffhip_plan* p, int* ndim, int64_t* n, int64_t* batch, int* prec, int* kind) {
  clear_error();
  FF_CHECK(p, FFHIP_ERR_ARG, "null plan");
  if (ndim) *ndim = p->ndim;
  if (n)
    for (int d = 0; d < p->ndim; ++d) n[d] = p->n[d];
  if (batch) *batch = p->batch;
  if (prec) *prec = p->prec;
  if (kind) *kind = p->kind;
  return FFHIP_OK;
}

}  // extern "C"
