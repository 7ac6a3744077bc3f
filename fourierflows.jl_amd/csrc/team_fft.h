// Register-resident Stockham FFT of a power-of-two length N by a "team" of
// T = N/P threads, each holding P points.
//
// On entry thread `tid` holds x[tid + p*T] in v[p]; on exit it holds
// X[tid + p*T].  Passes use radix P (and one smaller radix for the
// remainder); the first pass reads registers, the last writes registers, the
// passes in between exchange through a per-team LDS buffer of N + N/16
// complex values (one pad slot per 16 keeps strided writes off a single
// bank).  The caller's block runs all teams through the same sequence, so
// __syncthreads() is the team barrier.
//
// Twiddles come from an LDS table in pass order (team_tw_load): pass S's
// factors for r = 1..RK-1 and k = 0..NS-1 sit at tw_off(S) + (r-1) NS + k,
// so consecutive lanes read consecutive slots.  Reading them from global
// memory instead put an L2 round trip into every pass and, since stores and
// loads share one counter, made each pass wait for the stores before it.
// Kernels that run several transforms of one size can keep a thread's
// factors in registers instead (TwRegs), and two independent transforms can
// run side by side through one set of barriers (team_fft2).
#pragma once

#include "fft_device.h"

namespace ffhip {

template <int DIR, typename R> struct Bfly<16, DIR, R> {
  static __device__ __forceinline__ void run(cplx<R>* v) {
    // 16 = 4 x 4: DFT4 over n1 for each n2, twiddle W16^(n2 k1), DFT4 over n2
    cplx<R> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cplx<R> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Bfly<4, DIR, R>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    const R c1 = R(0.92387953251128675612818318939678829);  // cos(pi/8)
    const R s1 = R(0.38268343236508977172845998403039887);  // sin(pi/8)
    const R h = R(0.70710678118654752440084436210484904);
    // W16^m = exp(DIR*2*pi*i*m/16) for m = n2*k1 in {1,2,3,4,6,9}
    auto rot = [&](cplx<R> z, R c, R s) { return mkc<R>(z.x * c - DIR * z.y * s, z.y * c + DIR * z.x * s); };
    a[1][1] = rot(a[1][1], c1, s1);
    a[1][2] = rot(a[1][2], h, h);
    a[1][3] = rot(a[1][3], s1, c1);
    a[2][1] = rot(a[2][1], h, h);
    a[2][2] = cmul_si<DIR, R>(a[2][2]);
    a[2][3] = rot(a[2][3], -h, h);
    a[3][1] = rot(a[3][1], s1, c1);
    a[3][2] = rot(a[3][2], -h, h);
    a[3][3] = rot(a[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cplx<R> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Bfly<4, DIR, R>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

constexpr int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }
constexpr int cfull(int N, int P) { return (N % P == 0 && N >= P) ? 1 + cfull(N / P, P) : 0; }

#ifndef FF_LDS_PAD_SHIFT
#define FF_LDS_PAD_SHIFT 4  // one pad slot per 2^shift slots of a team's LDS buffer
#endif
template <int N, int P> struct TeamSched {
  static_assert((N & (N - 1)) == 0 && (P & (P - 1)) == 0, "powers of two only");
  static constexpr int T = N / P;
  static constexpr int NFULL = cfull(N, P);
  static constexpr int REM = N / cpow(P, NFULL);
  static constexpr int NPASS = NFULL + (REM > 1 ? 1 : 0);
  static constexpr int LDS = N + (N >> FF_LDS_PAD_SHIFT);  // complex slots per team
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> FF_LDS_PAD_SHIFT); }

#ifndef FF_LDS_BANKS
#define FF_LDS_BANKS 32  // 4-byte LDS banks the slice rotation targets (0: no rotation)
#endif

// Complex slots between the LDS slices of k teams whose threads are
// interleaved lane by lane (thread t belongs to team t % k).  Lanes of one
// wavefront then touch the same slot index in k different slices; a stride
// of slots + pad rotates slice c by c * BANKS / k banks, so those accesses
// land on different banks instead of serialising on one.
constexpr int team_slice(int slots, int slot_bytes, int k) {
  if (k <= 1 || FF_LDS_BANKS == 0) return slots;
  const int dw = slot_bytes / 4;
  const int target = (FF_LDS_BANKS / k) % FF_LDS_BANKS;
  for (int s = slots; s < slots + FF_LDS_BANKS; ++s)
    if ((s * dw) % FF_LDS_BANKS == target) return s;
  return slots;
}

// LDS twiddle table: slots before pass S's factors, and the table size
template <int N, int P, int S>
constexpr int tw_off() {
  if constexpr (S <= 1) {
    return 0;
  } else {
    constexpr int RK = S - 1 < TeamSched<N, P>::NFULL ? P : TeamSched<N, P>::REM;
    return tw_off<N, P, S - 1>() + (RK - 1) * cpow(P, S - 1);
  }
}
template <int N, int P>
constexpr int tw_slots() {
  return tw_off<N, P, TeamSched<N, P>::NPASS>();
}
// the same count for sizes known at run time (power-of-two n, points p)
constexpr int tw_slots_rt(int n, int p) {
  int total = 0, ns = 1, rest = n;
  while (rest % p == 0 && rest >= p) {
    if (ns > 1) total += (p - 1) * ns;
    ns *= p;
    rest /= p;
  }
  if (rest > 1 && ns > 1) total += (rest - 1) * ns;
  return total;
}

static_assert(tw_slots<2048, 16>() == tw_slots_rt(2048, 16) && tw_slots<1024, 16>() == tw_slots_rt(1024, 16) &&
                  tw_slots<512, 16>() == tw_slots_rt(512, 16) && tw_slots<256, 16>() == tw_slots_rt(256, 16) &&
                  tw_slots<1024, 8>() == tw_slots_rt(1024, 8) && tw_slots<512, 8>() == tw_slots_rt(512, 8) &&
                  tw_slots<256, 8>() == tw_slots_rt(256, 8) && tw_slots<128, 8>() == tw_slots_rt(128, 8) &&
                  tw_slots<64, 8>() == tw_slots_rt(64, 8),
              "run-time twiddle table size must match the compile-time one");

// RS: roots holds the RS*N roots of unity (root j of N is entry RS*j)
template <int N, int P, int S, typename R, int RS = 1>
__device__ __forceinline__ void tw_fill_pass(cplx<R>* ltw, const cplx<R>* roots, int t, int nt) {
  using TS = TeamSched<N, P>;
  if constexpr (S < TS::NPASS) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    constexpr int STEP = RS * (N / (NS * RK));
    constexpr int OFF = tw_off<N, P, S>();
    for (int e = t; e < (RK - 1) * NS; e += nt) {
      const int r = e / NS + 1, k = e - (r - 1) * NS;
      ltw[OFF + e] = roots[r * k * STEP];
    }
    tw_fill_pass<N, P, S + 1, R, RS>(ltw, roots, t, nt);
  }
}

// Fill the pass-ordered table from the N roots of unity (all threads of the
// block take part; the block barrier at the end publishes it).  Returns ltw.
// RS > 1 reads every RS-th entry of a table of RS*N roots.
template <int N, int P, typename R, int RS = 1>
__device__ __forceinline__ const cplx<R>* team_tw_load(cplx<R>* ltw, const void* roots) {
  tw_fill_pass<N, P, 1, R, RS>(ltw, reinterpret_cast<const cplx<R>*>(roots), (int)threadIdx.x, (int)blockDim.x);
  __syncthreads();
  return ltw;
}

// Per-thread twiddle cache: a thread's factors of every pass in registers
// (pass S, butterfly b, r at twr_off(S) + b (RK-1) + r-1), for kernels that
// run several transforms of one size and have the registers to spare.
template <int N, int P, int S>
constexpr int twr_off() {
  if constexpr (S <= 1) {
    return 0;
  } else {
    constexpr int RK = S - 1 < TeamSched<N, P>::NFULL ? P : TeamSched<N, P>::REM;
    return twr_off<N, P, S - 1>() + (P / RK) * (RK - 1);
  }
}
// NREG: passes 1..NREG keep their factors in registers, later passes read LDS
template <int N, int P, typename R, int NREG = TeamSched<N, P>::NPASS - 1>
struct TwRegs {
  static constexpr int NR = NREG;
  cplx<R> w[twr_off<N, P, NREG + 1>()];
};

template <int N, int P, int S, typename R, int NREG>
__device__ __forceinline__ void twr_fill(TwRegs<N, P, R, NREG>& t, const cplx<R>* ltw, int tid) {
  using TS = TeamSched<N, P>;
  if constexpr (S < TS::NPASS && S <= NREG) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    constexpr int NB = P / RK;
    constexpr int T = N / P;
    constexpr int OFF = tw_off<N, P, S>();
    constexpr int RO = twr_off<N, P, S>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = (tid + b * T) & (NS - 1);
#pragma unroll
      for (int r = 1; r < RK; ++r) t.w[RO + b * (RK - 1) + r - 1] = ltw[OFF + (r - 1) * NS + k];
    }
    twr_fill<N, P, S + 1, R, NREG>(t, ltw, tid);
  }
}
template <int N, int P, typename R, int NREG>
__device__ __forceinline__ void twr_fill(TwRegs<N, P, R, NREG>& t, const cplx<R>* ltw, int tid) {
  twr_fill<N, P, 1, R, NREG>(t, ltw, tid);
}

// The same registers filled straight from the N roots of unity in global
// memory (for kernels that keep no LDS table).
template <int N, int P, int S, typename R, int NREG>
__device__ __forceinline__ void twr_fill_roots(TwRegs<N, P, R, NREG>& t, const cplx<R>* roots, int tid) {
  using TS = TeamSched<N, P>;
  if constexpr (S < TS::NPASS && S <= NREG) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    constexpr int NB = P / RK;
    constexpr int T = N / P;
    constexpr int STEP = N / (NS * RK);
    constexpr int RO = twr_off<N, P, S>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = (tid + b * T) & (NS - 1);
#pragma unroll
      for (int r = 1; r < RK; ++r) t.w[RO + b * (RK - 1) + r - 1] = roots[r * k * STEP];
    }
    twr_fill_roots<N, P, S + 1, R, NREG>(t, roots, tid);
  }
}

// Barrier of a team: the whole workgroup, or (WAVE) one wavefront that is the
// whole team.  A wavefront's LDS operations execute in issue order, so the
// wavefront-scope fences and the wave barrier only stop the compiler from
// moving LDS reads and writes of different lanes across the exchange.
template <bool WAVE>
__device__ __forceinline__ void team_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One pass in two halves: tp_load gathers a thread's butterfly inputs (from
// registers on the first pass, else from LDS), tp_exec applies twiddles and
// butterflies and writes the outputs (to registers on the last pass, else to
// LDS).  team_pass puts the barriers between them; team_pass2 runs two
// independent transforms through one set of barriers.
template <int N, int P, int RK, bool FIRST, typename R>
__device__ __forceinline__ void tp_load(const cplx<R> (&v)[P], cplx<R> (&x)[P], const cplx<R>* __restrict__ lds,
                                        int tid) {
  constexpr int T = N / P;
  constexpr int NB = P / RK;  // butterflies per thread
  constexpr int M = N / RK;
  if constexpr (FIRST) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < RK; ++r) x[b * RK + r] = v[b + r * NB];
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < RK; ++r) x[b * RK + r] = lds[lpad(tid + b * T + r * M)];
  }
}

template <int N, int P, int RK, int NS, int TWO, int TRO, bool REG, bool LAST, int DIR, typename R>
__device__ __forceinline__ void tp_exec(cplx<R> (&v)[P], cplx<R> (&x)[P], cplx<R>* __restrict__ lds, int tid,
                                        const cplx<R>* __restrict__ tw, const cplx<R>* twr) {
  constexpr int T = N / P;
  constexpr int NB = P / RK;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = tid + b * T;
    const int k = j & (NS - 1);
    cplx<R>* y = x + b * RK;
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < RK; ++r) {
        const cplx<R> w = REG ? twr[TRO + b * (RK - 1) + r - 1] : tw[TWO + (r - 1) * NS + k];
        y[r] = cmul<R>(y[r], twid<DIR, R>(w));
      }
    }
    Bfly<RK, DIR, R>::run(y);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < RK; ++r) v[b + r * NB] = y[r];
    } else {
      const int o = (j - k) * RK + k;
#pragma unroll
      for (int r = 0; r < RK; ++r) lds[lpad(o + r * NS)] = y[r];
    }
  }
}

template <int N, int P, int RK, int NS, int TWO, int TRO, bool REG, bool FIRST, bool LAST, int DIR, typename R,
          bool WAVE = false>
__device__ __forceinline__ void team_pass(cplx<R> (&v)[P], cplx<R>* __restrict__ lds, int tid,
                                          const cplx<R>* __restrict__ tw, const cplx<R>* twr) {
  cplx<R> x[P];
  tp_load<N, P, RK, FIRST, R>(v, x, lds, tid);
  if constexpr (!FIRST) team_sync<WAVE>();
  tp_exec<N, P, RK, NS, TWO, TRO, REG, LAST, DIR, R>(v, x, lds, tid, tw, twr);
  if constexpr (!LAST) team_sync<WAVE>();
}

template <int N, int P, int RK, int NS, int TWO, int TRO, bool REG, bool FIRST, bool LAST, int DIR, typename R,
          bool WAVE = false>
__device__ __forceinline__ void team_pass2(cplx<R> (&v1)[P], cplx<R> (&v2)[P], cplx<R>* __restrict__ lds1,
                                           cplx<R>* __restrict__ lds2, int tid, const cplx<R>* __restrict__ tw,
                                           const cplx<R>* twr) {
  cplx<R> x1[P], x2[P];
  tp_load<N, P, RK, FIRST, R>(v1, x1, lds1, tid);
  tp_load<N, P, RK, FIRST, R>(v2, x2, lds2, tid);
  if constexpr (!FIRST) team_sync<WAVE>();
  tp_exec<N, P, RK, NS, TWO, TRO, REG, LAST, DIR, R>(v1, x1, lds1, tid, tw, twr);
  tp_exec<N, P, RK, NS, TWO, TRO, REG, LAST, DIR, R>(v2, x2, lds2, tid, tw, twr);
  if constexpr (!LAST) team_sync<WAVE>();
}

template <int N, int P, int S, int DIR, int NREG, typename R, bool WAVE = false>
__device__ __forceinline__ void team_passes(cplx<R> (&v)[P], cplx<R>* lds, int tid, const cplx<R>* tw,
                                            const cplx<R>* twr) {
  using TS = TeamSched<N, P>;
  if constexpr (S < TS::NPASS) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    team_pass<N, P, RK, NS, tw_off<N, P, S>(), twr_off<N, P, S>(), (S <= NREG), S == 0, S == TS::NPASS - 1, DIR,
              R, WAVE>(v, lds, tid, tw, twr);
    team_passes<N, P, S + 1, DIR, NREG, R, WAVE>(v, lds, tid, tw, twr);
  }
}

template <int N, int P, int S, int DIR, int NREG, typename R, bool WAVE = false>
__device__ __forceinline__ void team_passes2(cplx<R> (&v1)[P], cplx<R> (&v2)[P], cplx<R>* lds1, cplx<R>* lds2,
                                             int tid, const cplx<R>* tw, const cplx<R>* twr) {
  using TS = TeamSched<N, P>;
  if constexpr (S < TS::NPASS) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    team_pass2<N, P, RK, NS, tw_off<N, P, S>(), twr_off<N, P, S>(), (S <= NREG), S == 0, S == TS::NPASS - 1, DIR,
               R, WAVE>(v1, v2, lds1, lds2, tid, tw, twr);
    team_passes2<N, P, S + 1, DIR, NREG, R, WAVE>(v1, v2, lds1, lds2, tid, tw, twr);
  }
}

// Full transform.  Requires NPASS >= 2 (N >= 2P) so that the first pass
// writes LDS and the last pass reads it.  Ends with LDS free for reuse.
// tw is the table team_tw_load filled.
template <int N, int P, int DIR, typename R>
__device__ __forceinline__ void team_fft(cplx<R> (&v)[P], cplx<R>* lds, int tid, const cplx<R>* tw) {
  static_assert(TeamSched<N, P>::NPASS >= 2, "team_fft needs at least two passes");
  team_passes<N, P, 0, DIR, 0, R>(v, lds, tid, tw, nullptr);
}

// Two independent transforms of one direction side by side (separate LDS
// buffers, shared barriers: half the barriers of two team_fft calls and two
// streams of independent work between them), twiddles as below.
template <int N, int P, int DIR, typename R, int NREG>
__device__ __forceinline__ void team_fft2(cplx<R> (&v1)[P], cplx<R> (&v2)[P], cplx<R>* lds1, cplx<R>* lds2, int tid,
                                          const cplx<R>* tw, const TwRegs<N, P, R, NREG>& t) {
  static_assert(TeamSched<N, P>::NPASS >= 2, "team_fft needs at least two passes");
  team_passes2<N, P, 0, DIR, NREG, R>(v1, v2, lds1, lds2, tid, tw, t.w);
}

// Two independent transforms side by side with every twiddle from the LDS table.
template <int N, int P, int DIR, typename R>
__device__ __forceinline__ void team_fft2(cplx<R> (&v1)[P], cplx<R> (&v2)[P], cplx<R>* lds1, cplx<R>* lds2, int tid,
                                          const cplx<R>* tw) {
  static_assert(TeamSched<N, P>::NPASS >= 2, "team_fft needs at least two passes");
  team_passes2<N, P, 0, DIR, 0, R>(v1, v2, lds1, lds2, tid, tw, nullptr);
}

// The same transform with the first t.NR passes' twiddles from registers
// (twr_fill) and the rest from the LDS table tw.
template <int N, int P, int DIR, typename R, int NREG>
__device__ __forceinline__ void team_fft(cplx<R> (&v)[P], cplx<R>* lds, int tid, const cplx<R>* tw,
                                         const TwRegs<N, P, R, NREG>& t) {
  static_assert(TeamSched<N, P>::NPASS >= 2, "team_fft needs at least two passes");
  team_passes<N, P, 0, DIR, NREG, R>(v, lds, tid, tw, t.w);
}

// One-wavefront teams (T = N/P = 64 lanes, the team is the wavefront): the
// same transforms with wavefront-level exchange barriers, so teams never
// wait for each other.  Every twiddle comes from registers (t.NR covers all
// passes that have factors), so no LDS table is needed.
template <int N, int P, int DIR, typename R, int NREG>
__device__ __forceinline__ void wave_fft(cplx<R> (&v)[P], cplx<R>* lds, int lane, const TwRegs<N, P, R, NREG>& t) {
  static_assert(N / P == 64, "a wavefront team has 64 lanes");
  static_assert(NREG >= TeamSched<N, P>::NPASS - 1, "wave_fft keeps every twiddle in registers");
  team_passes<N, P, 0, DIR, NREG, R, true>(v, lds, lane, nullptr, t.w);
}
template <int N, int P, int DIR, typename R, int NREG>
__device__ __forceinline__ void wave_fft2(cplx<R> (&v1)[P], cplx<R> (&v2)[P], cplx<R>* lds1, cplx<R>* lds2, int lane,
                                          const TwRegs<N, P, R, NREG>& t) {
  static_assert(N / P == 64, "a wavefront team has 64 lanes");
  static_assert(NREG >= TeamSched<N, P>::NPASS - 1, "wave_fft keeps every twiddle in registers");
  team_passes2<N, P, 0, DIR, NREG, R, true>(v1, v2, lds1, lds2, lane, nullptr, t.w);
}

// Registers of a half-length transform (M = N/2 points) filled from the N
// roots of unity exp(-2 pi i k/N): the M-point factor of index e is root 2e.
template <int M, int P, int S, typename R, int NREG>
__device__ __forceinline__ void twr_fill_roots_half(TwRegs<M, P, R, NREG>& t, const cplx<R>* roots, int tid) {
  using TS = TeamSched<M, P>;
  if constexpr (S < TS::NPASS && S <= NREG) {
    constexpr int RK = S < TS::NFULL ? P : TS::REM;
    constexpr int NS = cpow(P, S);
    constexpr int NB = P / RK;
    constexpr int T = M / P;
    constexpr int STEP = M / (NS * RK);
    constexpr int RO = twr_off<M, P, S>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = (tid + b * T) & (NS - 1);
#pragma unroll
      for (int r = 1; r < RK; ++r) t.w[RO + b * (RK - 1) + r - 1] = roots[2 * r * k * STEP];
    }
    twr_fill_roots_half<M, P, S + 1, R, NREG>(t, roots, tid);
  }
}

}  // namespace ffhip
