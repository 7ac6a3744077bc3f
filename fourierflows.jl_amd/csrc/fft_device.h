// Device-side FFT building blocks: radix-2/3/4/5/8 butterflies, a generic
// prime-radix DFT stage, and a Stockham auto-sort driver over LDS tiles.
//
// Conventions: DIR = -1 is the forward transform (kernel exp(-2 pi i jk/n),
// FFTW's sign), DIR = +1 the unnormalised backward one.  Twiddles come from a
// table tw[k] = exp(-2 pi i k/n), k in [0, n), generated on the host in
// extended precision and rounded once to the working precision.
#pragma once

#include "common.h"

namespace ffhip {

constexpr int kMaxStages = 24;

// Radix schedule of one complex FFT length, passed by value to kernels.
struct Radices {
  int n;
  int nst;
  int r[kMaxStages];
};

// ----------------------------------------------------------------- butterflies
template <int RAD, int DIR, typename R> struct Bfly;

template <int DIR, typename R> struct Bfly<2, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    cplx<R> a = v[0], b = v[1];
    v[0] = cadd<R>(a, b);
    v[1] = csub<R>(a, b);
  }
};

template <int DIR, typename R> struct Bfly<4, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    cplx<R> t0 = cadd<R>(v[0], v[2]);
    cplx<R> t1 = csub<R>(v[0], v[2]);
    cplx<R> t2 = cadd<R>(v[1], v[3]);
    cplx<R> t3 = cmul_si<DIR, R>(csub<R>(v[1], v[3]));
    v[0] = cadd<R>(t0, t2);
    v[2] = csub<R>(t0, t2);
    v[1] = cadd<R>(t1, t3);
    v[3] = csub<R>(t1, t3);
  }
};

template <int DIR, typename R> struct Bfly<8, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    const R h = R(0.70710678118654752440084436210484904);
    cplx<R> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<R> o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<4, DIR, R>::run(e);
    Bfly<4, DIR, R>::run(o);
    // o1 *= (1 + DIR i)/sqrt2, o2 *= DIR i, o3 *= (-1 + DIR i)/sqrt2
    o[1] = mkc<R>(h * (o[1].x - DIR * o[1].y), h * (o[1].y + DIR * o[1].x));
    o[2] = cmul_si<DIR, R>(o[2]);
    o[3] = mkc<R>(h * (-o[3].x - DIR * o[3].y), h * (-o[3].y + DIR * o[3].x));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd<R>(e[k], o[k]);
      v[k + 4] = csub<R>(e[k], o[k]);
    }
  }
};

template <int DIR, typename R> struct Bfly<3, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    const R s60 = R(0.86602540378443864676372317075293618);
    cplx<R> s = cadd<R>(v[1], v[2]);
    cplx<R> d = csub<R>(v[1], v[2]);
    cplx<R> y0 = cadd<R>(v[0], s);
    cplx<R> t = csub<R>(v[0], cscale<R>(s, R(0.5)));
    cplx<R> u = cscale<R>(cmul_si<DIR, R>(d), s60);
    v[0] = y0;
    v[1] = cadd<R>(t, u);
    v[2] = csub<R>(t, u);
  }
};

template <int DIR, typename R> struct Bfly<5, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    const R c1 = R(0.30901699437494742410229341718281906);   // cos(2pi/5)
    const R c2 = R(-0.80901699437494742410229341718281906);  // cos(4pi/5)
    const R s1 = R(0.95105651629515357211643933337938214);   // sin(2pi/5)
    const R s2 = R(0.58778525229247312916870595463907277);   // sin(4pi/5)
    cplx<R> a14 = cadd<R>(v[1], v[4]), d14 = csub<R>(v[1], v[4]);
    cplx<R> a23 = cadd<R>(v[2], v[3]), d23 = csub<R>(v[2], v[3]);
    cplx<R> y0 = cadd<R>(v[0], cadd<R>(a14, a23));
    cplx<R> t1 = cadd<R>(v[0], cadd<R>(cscale<R>(a14, c1), cscale<R>(a23, c2)));
    cplx<R> t2 = cadd<R>(v[0], cadd<R>(cscale<R>(a14, c2), cscale<R>(a23, c1)));
    cplx<R> u1 = cmul_si<DIR, R>(cadd<R>(cscale<R>(d14, s1), cscale<R>(d23, s2)));
    cplx<R> u2 = cmul_si<DIR, R>(csub<R>(cscale<R>(d14, s2), cscale<R>(d23, s1)));
    v[0] = y0;
    v[1] = cadd<R>(t1, u1);
    v[4] = csub<R>(t1, u1);
    v[2] = cadd<R>(t2, u2);
    v[3] = csub<R>(t2, u2);
  }
};

// ----------------------------------------------------------------- Stockham pass
// Tile of `nb` transforms of length n in LDS; element (t, i) lives at
// base[t*sb + i*si].  One radix-RAD pass with span Ns (product of the radices
// already applied): butterfly j reads i = j + r*n/RAD and writes
// i = (j - j%Ns)*RAD + j%Ns + r*Ns.  `tfast` picks which index varies fastest
// across threads (transform index for interleaved tiles).
template <int RAD, int DIR, typename R>
__device__ __forceinline__ void stockham_pass(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst,
                                              const cplx<R>* __restrict__ tw, int n, int Ns, int nb,
                                              int sb, int si, bool tfast) {
  const int m = n / RAD;
  const int total = nb * m;
  const int tstep = n / (Ns * RAD);
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int t, j;
    if (tfast) {
      t = idx % nb;
      j = idx / nb;
    } else {
      t = idx / m;
      j = idx - t * m;
    }
    const cplx<R>* s = src + t * sb;
    cplx<R>* d = dst + t * sb;
    cplx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[r] = s[(j + r * m) * si];
    const int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[r] = cmul<R>(v[r], twid<DIR, R>(tw[r * k * tstep]));
    }
    Bfly<RAD, DIR, R>::run(v);
    const int o = (j - k) * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) d[(o + r * Ns) * si] = v[r];
  }
}

// Generic prime radix p (direct O(p^2) DFT, inputs read from LDS).
template <int DIR, typename R>
__device__ __forceinline__ void stockham_pass_generic(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst,
                                                      const cplx<R>* __restrict__ tw, int n, int p, int Ns,
                                                      int nb, int sb, int si, bool tfast) {
  const int m = n / p;
  const int total = nb * m;
  const int tstep = n / (Ns * p);
  const int wstep = n / p;  // W_p = tw[wstep]
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int t, j;
    if (tfast) {
      t = idx % nb;
      j = idx / nb;
    } else {
      t = idx / m;
      j = idx - t * m;
    }
    const cplx<R>* s = src + t * sb;
    cplx<R>* d = dst + t * sb;
    const int k = j % Ns;
    const int o = (j - k) * p + k;
    for (int q = 0; q < p; ++q) {
      cplx<R> acc = mkc<R>(0, 0);
      for (int r = 0; r < p; ++r) {
        cplx<R> x = s[(j + r * m) * si];
        if (Ns > 1 && r > 0) x = cmul<R>(x, twid<DIR, R>(tw[r * k * tstep]));
        const int e = (r * q) % p;
        acc = cadd<R>(acc, e == 0 ? x : cmul<R>(x, twid<DIR, R>(tw[e * wstep])));
      }
      d[(o + q * Ns) * si] = acc;
    }
  }
}

// Full transform of an LDS tile: ping-pongs between a and b; returns the
// buffer that holds the result.  Ends with a barrier.
template <int DIR, typename R>
__device__ cplx<R>* stockham_lds(cplx<R>* a, cplx<R>* b, const Radices& rd, const cplx<R>* __restrict__ tw,
                                 int nb, int sb, int si, bool tfast) {
  int Ns = 1;
  const int n = rd.n;
  for (int s = 0; s < rd.nst; ++s) {
    const int p = rd.r[s];
    switch (p) {
      case 2: stockham_pass<2, DIR, R>(a, b, tw, n, Ns, nb, sb, si, tfast); break;
      case 3: stockham_pass<3, DIR, R>(a, b, tw, n, Ns, nb, sb, si, tfast); break;
      case 4: stockham_pass<4, DIR, R>(a, b, tw, n, Ns, nb, sb, si, tfast); break;
      case 5: stockham_pass<5, DIR, R>(a, b, tw, n, Ns, nb, sb, si, tfast); break;
      case 8: stockham_pass<8, DIR, R>(a, b, tw, n, Ns, nb, sb, si, tfast); break;
      default: stockham_pass_generic<DIR, R>(a, b, tw, n, p, Ns, nb, sb, si, tfast); break;
    }
    __syncthreads();
    cplx<R>* t = a;
    a = b;
    b = t;
    Ns *= p;
  }
  return a;
}

}  // namespace ffhip
