// Fused RK4 step of the 3D passive tracer in the steady ABC flow on
// power-of-two grids (64..512 per axis).
//
// One RK4 stage is four kernels over the kept (dealiased) band only:
//   y-inverse   : inverse y-FFT of the z-inverted stage state        (per kx, z)
//   x pass      : c2r in x, products u c, v c, w c with the analytic
//                 velocity, forward x-FFTs of (u c + i v c) and w c   (per y, z row)
//   y-forward   : forward y-FFTs of the three products, kept y only (per kx, z)
//   z pass      : forward z-FFTs of the three products -> N = -i (kx U + l V + m W)
//                 at the kept modes -> the RK4 stage of
//                 src/timesteppers.jl:237-274 (addlinearterm!, substepsol!,
//                 RK4update!, FP64 arithmetic, the state type at every store
//                 as there) -> inverse z-FFT of the next stage state (per kx, y)
// The reference's calcN! would run one c2r and three r2c 3D transforms plus
// six broadcasts per stage over the full arrays.
//
// Layouts (complex, blocks of 4 kx as in the 2D passes):
//   state, stage state, RHS1..3 : [TX][KY][KZ][4]
//   y-pass input/output A1, A2  : [nz][TX][ny][4]   (A1's aliased y rows stay zero)
//   x-pass output B (3 fields)  : [nz][TX][ny][4]
//   z-pass input C (3 fields)   : [TX][KY][nz][4]
#include <cstdlib>
#include <new>

#include "vort2d.h"

namespace ffhip {

#ifndef FF_TAX_TWREG
#define FF_TAX_TWREG 1  // adjoint x pass: twiddles in registers across its three transforms
#endif

struct TracerGeom {
  int nx, ny, nz, nkr;
  int KX, KXp, TX;
  int ylo, yhi, KY;  // aliased y rows [ylo, yhi] (0-based), kept count
  int zlo, zhi, KZ;
  double dk, dl, dm;
  double norm;       // 1 / (nx ny nz): the three unnormalised inverses
  double A, B, C;    // ABC amplitudes
};

struct TracerState {
  TracerGeom g;
  void* S = nullptr;   // state
  void* S1 = nullptr;  // stage state
  void* R[3] = {nullptr, nullptr, nullptr};  // RHS1..3
  void* A1 = nullptr;
  void* A2 = nullptr;
  void* Bf[3] = {nullptr, nullptr, nullptr};
  void* Cf[3] = {nullptr, nullptr, nullptr};
  void* Lp = nullptr;  // L packed to the state blocking (FP64)
  const ffhip_rk4* lp_owner = nullptr;
  void* twx = nullptr;
  void* twy = nullptr;
  void* twz = nullptr;
  const double* tab = nullptr;  // sin x, cos x | sin y, cos y | sin z, cos z (FP64, owned by the operator)
};

struct TracerArgs {
  void* S;
  void* S1;
  void* R0;
  void* R1;
  void* R2;
  void* A1;
  void* A2;
  void* B0;
  void* B1;
  void* B2;
  void* C0;
  void* C1;
  void* C2;
  const double* L;
  const double* tab;
  const void* twx;
  const void* twy;
  const void* twz;
  double h2, h6, h;  // dt/2 (rounded as the reference rounds it), dt/6, dt
  int do_inverse;
};

__device__ __forceinline__ int t_kept_index(int i, int lo, int hi) { return i < lo ? i : i - (hi - lo + 1); }
__device__ __forceinline__ int t_full_index(int k, int lo, int hi) { return k < lo ? k : k + (hi - lo + 1); }

// ----------------------------------------------------------------- z pass
// STAGE 0: inverse z-FFT of the state only.  STAGE 1..4: RK4 stages.
template <typename R, int NZ, int P, int STAGE>
__global__ void __launch_bounds__(kVCB*(NZ / P)) k_tz(TracerGeom g, TracerArgs a) {
  constexpr int T = NZ / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char tsm[];
  // one 4-column tile at one kept y per workgroup; the four y rows that share
  // the 128 B lines of the [nz][TX][ny][4] output run on one XCD (one L2)
  int q, sub;
  xcd_split(blockIdx.x, 4, &q, &sub);
  const int nq = g.ny / 4;
  if (q >= g.TX * nq) return;
  const int tx = q / nq, iy = (q % nq) * 4 + sub;
  if (iy >= g.ylo && iy <= g.yhi) return;  // aliased row: no work
  const int ky = t_kept_index(iy, g.ylo, g.yhi);
  const int t = threadIdx.x;
  const int c = t % CB;
  const int tid = t / CB;
  constexpr int SL = team_slice(TeamSched<NZ, P>::LDS, sizeof(cplx<R>), CB);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(tsm) + c * SL;
  const cplx<R>* tw = team_tw_load<NZ, P, R>(reinterpret_cast<cplx<R>*>(tsm) + CB * SL, a.twz);
  const int ix = tx * CB + c;
  const bool colkept = ix < g.KX;
  const R kx = (R)(g.dk * (double)ix);
  const R l = (R)(g.dl * (double)vfreq(iy, g.ny));
  const int64_t cbase = ((int64_t)tx * g.KY + ky) * NZ * CB + c;      // C[f][tx][ky][z][c]
  const int64_t sbase = ((int64_t)tx * g.KY + ky) * g.KZ * CB + c;    // state [tx][ky][kz][c]

  cplx<R> nn[P];
  if constexpr (STAGE > 0) {
    // N = -i (kx U + l V + m W), accumulated field by field in the state type
    const cplx<R>* Cs[3] = {reinterpret_cast<const cplx<R>*>(a.C0), reinterpret_cast<const cplx<R>*>(a.C1),
                            reinterpret_cast<const cplx<R>*>(a.C2)};
#pragma unroll
    for (int p = 0; p < P; ++p) nn[p] = mkc<R>(0, 0);
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
      cplx<R> v[P];
#pragma unroll
      for (int p = 0; p < P; ++p) v[p] = Cs[f][cbase + (int64_t)(tid + p * T) * CB];
      team_fft<NZ, P, -1, R>(v, lds, tid, tw);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int iz = tid + p * T;
        const R q = f == 0 ? kx : (f == 1 ? l : (R)(g.dm * (double)vfreq(iz, NZ)));
        // running sum s = kx U + l V + m W in the order of k_tracer_assemble
        if (f == 0) {
          nn[p] = mkc<R>(q * v[p].x, q * v[p].y);
        } else {
          nn[p] = mkc<R>(nn[p].x + q * v[p].x, nn[p].y + q * v[p].y);
        }
      }
    }
  }

  cplx<R>* S = reinterpret_cast<cplx<R>*>(a.S);
  cplx<R>* S1 = reinterpret_cast<cplx<R>*>(a.S1);
  cplx<R>* R0 = reinterpret_cast<cplx<R>*>(a.R0);
  cplx<R>* R1 = reinterpret_cast<cplx<R>*>(a.R1);
  cplx<R>* R2 = reinterpret_cast<cplx<R>*>(a.R2);
  cplx<R> nw[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int iz = tid + p * T;
    nw[p] = mkc<R>(0, 0);
    const bool kept = colkept && (iz < g.zlo || iz > g.zhi);
    if (!kept) continue;
    const int64_t si = sbase + (int64_t)t_kept_index(iz, g.zlo, g.zhi) * CB;
    if constexpr (STAGE == 0) {
      nw[p] = S[si];
    } else {
      // calcN! output in the state type (N = -i s), then addlinearterm!: RHS = N + L x
      const cplx<R> N = mkc<R>(nn[p].y, -nn[p].x);
      const cplx<R> x = STAGE == 1 ? S[si] : S1[si];
      const double L = a.L[si];
      const cplx<R> rhs = mkc<R>((R)((double)N.x + L * (double)x.x), (R)((double)N.y + L * (double)x.y));
      const cplx<R> u = STAGE == 1 ? x : S[si];
      if constexpr (STAGE < 4) {
        // substepsol!: sol + h RHS (h = dt/2, dt/2, dt)
        const double h = STAGE == 3 ? a.h : a.h2;
        (STAGE == 1 ? R0 : (STAGE == 2 ? R1 : R2))[si] = rhs;
        nw[p] = mkc<R>((R)((double)u.x + h * (double)rhs.x), (R)((double)u.y + h * (double)rhs.y));
        S1[si] = nw[p];
      } else {
        // RK4update!: sol += dt/6 (RHS1 + 2 RHS2 + 2 RHS3 + RHS4)
        const cplx<R> r1 = R0[si], r2 = R1[si], r3 = R2[si];
        double sx = (double)r1.x + 2.0 * (double)r2.x, sy = (double)r1.y + 2.0 * (double)r2.y;
        sx = sx + 2.0 * (double)r3.x;
        sy = sy + 2.0 * (double)r3.y;
        sx = sx + (double)rhs.x;
        sy = sy + (double)rhs.y;
        nw[p] = mkc<R>((R)((double)u.x + a.h6 * sx), (R)((double)u.y + a.h6 * sy));
        S[si] = nw[p];
      }
    }
  }
  if (!a.do_inverse) return;  // uniform over the block
  team_fft<NZ, P, 1, R>(nw, lds, tid, tw);
  cplx<R>* A1 = reinterpret_cast<cplx<R>*>(a.A1);
  const int64_t plane = (int64_t)g.TX * g.ny * CB;  // one z plane of A1
  const int64_t abase = ((int64_t)tx * g.ny + iy) * CB + c;
#pragma unroll
  for (int p = 0; p < P; ++p) A1[(int64_t)(tid + p * T) * plane + abase] = nw[p];
}

// ----------------------------------------------------------------- y passes
// NF fields, one y-column each: src rows read at the kept y only when
// src_kept (the others are zero), dst written in full ([nz][TX][ny][4]) or
// at the kept rows only into [TX][KY][nz][4] when dst_kept.  src == dst is
// allowed (each column is read whole before it is written).
struct TracerYArgs {
  const void* src[3];
  void* dst[3];
  const void* tw;
  int src_kept, dst_kept;
};

template <typename R, int NY, int P, int DIR, int NF>
__global__ void __launch_bounds__(kVCB*(NY / P)) k_ty(TracerGeom g, TracerYArgs a) {
  constexpr int T = NY / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char tsm[];
  // one 4-column tile at one z per workgroup; the four z planes that share the
  // 128 B lines of a [TX][KY][nz][4] output run on one XCD (one L2)
  int q, sub;
  xcd_split(blockIdx.x, 4, &q, &sub);
  const int nq = g.nz / 4;
  if (q >= g.TX * nq) return;
  const int tx = q % g.TX, iz = (q / g.TX) * 4 + sub;
  const int t = threadIdx.x;
  const int c = t % CB;
  const int tid = t / CB;
  constexpr int SL = team_slice(TeamSched<NY, P>::LDS, sizeof(cplx<R>), CB);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(tsm) + c * SL;
  const cplx<R>* tw = team_tw_load<NY, P, R>(reinterpret_cast<cplx<R>*>(tsm) + CB * SL, a.tw);
  const int64_t base = ((int64_t)iz * g.TX + tx) * NY * CB + c;  // [z][tx][y][c]
#pragma unroll 1
  for (int f = 0; f < NF; ++f) {
    const cplx<R>* src = reinterpret_cast<const cplx<R>*>(a.src[f]);
    cplx<R>* dst = reinterpret_cast<cplx<R>*>(a.dst[f]);
    cplx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      const bool zero = a.src_kept && iy >= g.ylo && iy <= g.yhi;
      v[p] = zero ? mkc<R>(0, 0) : src[base + (int64_t)iy * CB];
    }
    team_fft<NY, P, DIR, R>(v, lds, tid, tw);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      if (!a.dst_kept) {
        dst[base + (int64_t)iy * CB] = v[p];
      } else if (iy < g.ylo || iy > g.yhi) {
        const int ky = t_kept_index(iy, g.ylo, g.yhi);
        dst[(((int64_t)tx * g.KY + ky) * g.nz + iz) * CB + c] = v[p];
      }
    }
  }
}

// ----------------------------------------------------------------- x pass
template <typename R, int NX, int P>
__global__ void __launch_bounds__(4 * (NX / P)) k_tx(TracerGeom g, TracerArgs a) {
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char tsm[];
  const int grp = (int)blockIdx.x;  // four rows (one 128 B line group) at one z per workgroup
  const int iz = grp / (g.ny / 4);
  const int r = (int)(threadIdx.x % 4);
  const int y = (grp % (g.ny / 4)) * 4 + r;
  const int tid = threadIdx.x / 4;
  constexpr int SL = team_slice(TeamSched<NX, P>::LDS, sizeof(cplx<R>), 4);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(tsm) + r * SL;
  const cplx<R>* tw = team_tw_load<NX, P, R>(reinterpret_cast<cplx<R>*>(tsm) + 4 * SL, a.twx);
  const int KX = g.KX;
  const int64_t ny = g.ny;
  const int64_t zoff = (int64_t)iz * g.TX * ny * CB;
  auto xaddr = [&](int k) -> int64_t { return zoff + ((int64_t)(k / CB) * ny + y) * CB + (k % CB); };
  const cplx<R>* A2 = reinterpret_cast<const cplx<R>*>(a.A2);

  // c (real) as one complex inverse of its Hermitian-extended half spectrum
  cplx<R> w[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {  // branch-free: all loads issue before the first wait
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;  // direct / mirrored (conjugate) source
    const cplx<R> c = A2[xaddr(lo ? k : (hi ? NX - k : 0))];
    const R sx = (lo || hi) ? c.x : R(0);
    const R sy = lo ? (k == 0 ? R(0) : c.y) : (hi ? -c.y : R(0));  // c2r ignores Im at kx = 0
    w[p] = mkc<R>(sx, sy);
  }
  team_fft<NX, P, 1, R>(w, lds, tid, tw);

  // u = A sin z + C cos y (constant along the row), v = B sin x + A cos z, w = C sin y + B cos x
  const double* sx = a.tab;
  const double* cx = a.tab + g.nx;
  const double* sy = a.tab + 2 * g.nx;
  const double* cy = sy + g.ny;
  const double* sz = a.tab + 2 * g.nx + 2 * g.ny;
  const double* cz = sz + g.nz;
  const R u = (R)(g.A * sz[iz] + g.C * cy[y]);
  const double acz = g.A * cz[iz], csy = g.C * sy[y];
  const R nrm = (R)g.norm;
  cplx<R> z[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int ix = tid + p * T;
    const R cval = w[p].x * nrm;
    const R v = (R)(g.B * sx[ix] + acz);
    const R ww = (R)(csy + g.B * cx[ix]);
    z[p] = mkc<R>(u * cval, v * cval);  // u c + i v c
    w[p] = mkc<R>(ww * cval, 0);        // w c
  }
  team_fft<NX, P, -1, R>(z, lds, tid, tw);
  // split (u c + i v c) into the half spectra of u c and v c
#pragma unroll
  for (int p = 0; p < P; ++p) lds[lpad(tid + p * T)] = z[p];
  __syncthreads();
  cplx<R>* B0 = reinterpret_cast<cplx<R>*>(a.B0);
  cplx<R>* B1 = reinterpret_cast<cplx<R>*>(a.B1);
  cplx<R>* B2 = reinterpret_cast<cplx<R>*>(a.B2);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    cplx<R> Pk = mkc<R>(0, 0), Qk = mkc<R>(0, 0);
    if (k < KX) {
      const cplx<R> wk = z[p];
      const cplx<R> wc = cconj<R>(lds[lpad((NX - k) & (NX - 1))]);
      Pk = mkc<R>(R(0.5) * (wk.x + wc.x), R(0.5) * (wk.y + wc.y));
      Qk = mkc<R>(R(0.5) * (wk.y - wc.y), -R(0.5) * (wk.x - wc.x));
    }
    B0[xaddr(k)] = Pk;
    B1[xaddr(k)] = Qk;
  }
  __syncthreads();
  team_fft<NX, P, -1, R>(w, lds, tid, tw);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    B2[xaddr(k)] = k < KX ? w[p] : mkc<R>(0, 0);
  }
}

// ----------------------------------------------------------------- adjoint passes
// Transpose of the RK4 step (the tracer is linear in c, so no forward states
// are needed), with the arithmetic of the library's generic tracer adjoint:
//   (N + L)^T x = P (W/Nt) mul!( u p + v q + w r ) + L x,
//   p, q, r = unnormalised c2r of i kx x / W, i l x / W, i m x / W (x masked),
// W = 1 on the kx = 0 column, 2 elsewhere.  Passes per evaluation:
//   z pass   : forward z-FFT of the previous evaluation's product spectrum ->
//              its output o -> the reverse RK4 combination -> the next input
//              kb -> inverse z-FFTs of its three fields
//   y-inverse: the three fields, in place
//   x pass   : c2r of p + i q (packed) and of r, the product u p + v q + w r,
//              its forward x-FFT
//   y-forward: the product, kept rows only
// Reverse step (lam the adjoint state, h6 = dt/6, h3 = 2 dt/6, h2 = dt/2):
//   kb = h6 lam;  o = f(kb); ub = lam + o;  kb = h3 lam + h o
//   o = f(kb); ub += o; kb = h3 lam + h2 o; o = f(kb); ub += o; kb = h6 lam + h2 o
//   o = f(kb); lam = ub + o
struct TracerAdjArgs {
  void* LAM;  // [TX][KY][KZ][4]
  void* UB;
  void* KB;
  const void* F;  // forward z input of the product spectrum, [TX][KY][nz][4]
  void* Z0;       // inverse z outputs of the three fields, [nz][TX][ny][4]
  void* Z1;
  void* Z2;
  const double* L;
  const void* twz;
  double h, h2, h3, h6;
  int cont;  // STAGE 4: 1 = start the next (earlier) step, 0 = the run ends here
};

template <typename R, int NZ, int P, int STAGE>
__global__ void __launch_bounds__(kVCB*(NZ / P)) k_taz(TracerGeom g, TracerAdjArgs a) {
  constexpr int T = NZ / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char tsm[];
  int q, sub;  // work mapping as in k_tz
  xcd_split(blockIdx.x, 4, &q, &sub);
  const int nq = g.ny / 4;
  if (q >= g.TX * nq) return;
  const int tx = q / nq, iy = (q % nq) * 4 + sub;
  if (iy >= g.ylo && iy <= g.yhi) return;
  const int ky = t_kept_index(iy, g.ylo, g.yhi);
  const int t = threadIdx.x;
  const int c = t % CB;
  const int tid = t / CB;
  constexpr int SL = team_slice(TeamSched<NZ, P>::LDS, sizeof(cplx<R>), CB);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(tsm) + c * SL;
  const cplx<R>* tw = team_tw_load<NZ, P, R>(reinterpret_cast<cplx<R>*>(tsm) + CB * SL, a.twz);
  const int ix = tx * CB + c;
  const bool colkept = ix < g.KX;
  const int64_t fbase = ((int64_t)tx * g.KY + ky) * NZ * CB + c;
  const int64_t sbase = ((int64_t)tx * g.KY + ky) * g.KZ * CB + c;

  cplx<R> fz[P];
  if constexpr (STAGE > 0) {
    const cplx<R>* F = reinterpret_cast<const cplx<R>*>(a.F);
#pragma unroll
    for (int p = 0; p < P; ++p) fz[p] = F[fbase + (int64_t)(tid + p * T) * CB];
    team_fft<NZ, P, -1, R>(fz, lds, tid, tw);
  }
  cplx<R>* LAM = reinterpret_cast<cplx<R>*>(a.LAM);
  cplx<R>* UB = reinterpret_cast<cplx<R>*>(a.UB);
  cplx<R>* KB = reinterpret_cast<cplx<R>*>(a.KB);
  const double wn = (ix == 0 ? 1.0 : 2.0) * g.norm;  // W / Nt
  cplx<R> kb[P];
  bool next = true;
  if constexpr (STAGE == 4) next = a.cont != 0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int iz = tid + p * T;
    kb[p] = mkc<R>(0, 0);
    const bool kept = colkept && (iz < g.zlo || iz > g.zhi);
    if (!kept) continue;
    const int64_t si = sbase + (int64_t)t_kept_index(iz, g.zlo, g.zhi) * CB;
    const cplx<R> lam = LAM[si];
    if constexpr (STAGE == 0) {
      kb[p] = mkc<R>((R)(a.h6 * (double)lam.x), (R)(a.h6 * (double)lam.y));
    } else {
      const cplx<R> x = KB[si];
      const double l = a.L[si];
      const cplx<R> o = mkc<R>((R)(l * (double)x.x + wn * (double)fz[p].x), (R)(l * (double)x.y + wn * (double)fz[p].y));
      if constexpr (STAGE < 4) {
        const cplx<R> base = STAGE == 1 ? lam : UB[si];
        UB[si] = mkc<R>(base.x + o.x, base.y + o.y);
        const double cl = STAGE == 3 ? a.h6 : a.h3, cu = STAGE == 1 ? a.h : a.h2;
        kb[p] = mkc<R>((R)(cl * (double)lam.x + cu * (double)o.x), (R)(cl * (double)lam.y + cu * (double)o.y));
      } else {
        const cplx<R> ub = UB[si];
        const cplx<R> nl = mkc<R>(ub.x + o.x, ub.y + o.y);
        LAM[si] = nl;
        if (next) kb[p] = mkc<R>((R)(a.h6 * (double)nl.x), (R)(a.h6 * (double)nl.y));
      }
    }
    if (next) KB[si] = kb[p];
  }
  if (!next) return;  // uniform over the block

  // inverse z-FFTs of i kx kb / W, i l kb / W, i m kb / W
  const double iw = ix == 0 ? 1.0 : 0.5;
  const R qx = (R)(g.dk * (double)ix * iw), qy = (R)(g.dl * (double)vfreq(iy, g.ny) * iw);
  void* dsts[3] = {a.Z0, a.Z1, a.Z2};
  const int64_t plane = (int64_t)g.TX * g.ny * CB;
  const int64_t abase = ((int64_t)tx * g.ny + iy) * CB + c;
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    cplx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iz = tid + p * T;
      const R q = f == 0 ? qx : (f == 1 ? qy : (R)(g.dm * (double)vfreq(iz, NZ) * iw));
      v[p] = mkc<R>(-q * kb[p].y, q * kb[p].x);
    }
    team_fft<NZ, P, 1, R>(v, lds, tid, tw);
    cplx<R>* d = reinterpret_cast<cplx<R>*>(dsts[f]);
#pragma unroll
    for (int p = 0; p < P; ++p) d[(int64_t)(tid + p * T) * plane + abase] = v[p];
  }
}

// adjoint x pass: src fields (p, q, r half spectra after the y-inverse) -> product spectrum in dst
template <typename R, int NX, int P>
__global__ void __launch_bounds__(4 * (NX / P)) k_tax(TracerGeom g, const void* s0, const void* s1, const void* s2,
                                                      void* dst, const double* tab, const void* twx) {
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char tsm[];
  const int grp = (int)blockIdx.x;
  const int iz = grp / (g.ny / 4);
  const int r = (int)(threadIdx.x % 4);
  const int y = (grp % (g.ny / 4)) * 4 + r;
  const int tid = threadIdx.x / 4;
  constexpr int SL = team_slice(TeamSched<NX, P>::LDS, sizeof(cplx<R>), 4);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(tsm) + r * SL;
  const cplx<R>* tw = team_tw_load<NX, P, R>(reinterpret_cast<cplx<R>*>(tsm) + 4 * SL, twx);
  const int KX = g.KX;
  const int64_t ny = g.ny;
  const int64_t zoff = (int64_t)iz * g.TX * ny * CB;
  auto xaddr = [&](int k) -> int64_t { return zoff + ((int64_t)(k / CB) * ny + y) * CB + (k % CB); };
  const cplx<R>* Ph = reinterpret_cast<const cplx<R>*>(s0);
  const cplx<R>* Qh = reinterpret_cast<const cplx<R>*>(s1);
  const cplx<R>* Rh = reinterpret_cast<const cplx<R>*>(s2);
  cplx<R> z[P], w[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {  // branch-free: all loads issue before the first wait
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;  // direct / mirrored (conjugate) source
    const int64_t q = xaddr(lo ? k : (hi ? NX - k : 0));
    const cplx<R> A = Ph[q], B = Qh[q], C = Rh[q];
    const R ay = k == 0 ? R(0) : A.y, by = k == 0 ? R(0) : B.y, cy = k == 0 ? R(0) : C.y;  // Im ignored at kx = 0
    z[p] = mkc<R>(lo ? A.x - by : (hi ? A.x + B.y : R(0)), lo ? ay + B.x : (hi ? B.x - A.y : R(0)));
    w[p] = mkc<R>((lo || hi) ? C.x : R(0), lo ? cy : (hi ? -C.y : R(0)));
  }
#if FF_TAX_TWREG
  TwRegs<NX, P, R> twr;  // the three transforms share the thread's twiddles
  twr_fill(twr, tw, tid);
  team_fft<NX, P, 1, R>(z, lds, tid, tw, twr);  // p + i q
  team_fft<NX, P, 1, R>(w, lds, tid, tw, twr);  // r
#else
  team_fft<NX, P, 1, R>(z, lds, tid, tw);  // p + i q
  team_fft<NX, P, 1, R>(w, lds, tid, tw);  // r
#endif
  const double* sx = tab;
  const double* cx = tab + g.nx;
  const double* sy = tab + 2 * g.nx;
  const double* cy = sy + g.ny;
  const double* sz = tab + 2 * g.nx + 2 * g.ny;
  const double* cz = sz + g.nz;
  const R u = (R)(g.A * sz[iz] + g.C * cy[y]);
  const double acz = g.A * cz[iz], csy = g.C * sy[y];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int ix = tid + p * T;
    const R v = (R)(g.B * sx[ix] + acz);
    const R ww = (R)(csy + g.B * cx[ix]);
    z[p] = mkc<R>(u * z[p].x + v * z[p].y + ww * w[p].x, 0);  // u p + v q + w r
  }
#if FF_TAX_TWREG
  team_fft<NX, P, -1, R>(z, lds, tid, tw, twr);
#else
  team_fft<NX, P, -1, R>(z, lds, tid, tw);
#endif
  cplx<R>* D = reinterpret_cast<cplx<R>*>(dst);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    D[xaddr(k)] = k < KX ? z[p] : mkc<R>(0, 0);
  }
}

// aliased modes of the adjoint state: the reverse RK4 combination with f(x) = L x
template <typename R>
__global__ void k_taexport(cplx<R>* __restrict__ lam, const cplx<R>* __restrict__ S, const double* __restrict__ L,
                           double h2, double h, double h3, double h6, int64_t nsteps, TracerGeom g) {
  const int64_t total = (int64_t)g.nz * g.ny * g.nkr;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(idx % g.nkr);
    const int64_t r = idx / g.nkr;
    const int iy = (int)(r % g.ny), iz = (int)(r / g.ny);
    const bool kept = ix < g.KX && (iy < g.ylo || iy > g.yhi) && (iz < g.zlo || iz > g.zhi);
    if (kept) {
      const int tx = ix / kVCB, c = ix % kVCB;
      const int ky = t_kept_index(iy, g.ylo, g.yhi), kz = t_kept_index(iz, g.zlo, g.zhi);
      lam[idx] = S[(((int64_t)tx * g.KY + ky) * g.KZ + kz) * kVCB + c];
      continue;
    }
    cplx<R> u = lam[idx];
    if (u.x == R(0) && u.y == R(0)) continue;
    const double l = L[idx];
    auto f = [&](cplx<R> x) { return mkc<R>((R)(l * (double)x.x), (R)(l * (double)x.y)); };
    auto lin = [&](cplx<R> a, double ca, cplx<R> b, double cb) {
      return mkc<R>((R)(ca * (double)a.x + cb * (double)b.x), (R)(ca * (double)a.y + cb * (double)b.y));
    };
    for (int64_t s = 0; s < nsteps; ++s) {
      cplx<R> kb = mkc<R>((R)(h6 * (double)u.x), (R)(h6 * (double)u.y));
      cplx<R> o = f(kb);
      cplx<R> ub = mkc<R>(u.x + o.x, u.y + o.y);
      kb = lin(u, h3, o, h);
      o = f(kb);
      ub = mkc<R>(ub.x + o.x, ub.y + o.y);
      kb = lin(u, h3, o, h2);
      o = f(kb);
      ub = mkc<R>(ub.x + o.x, ub.y + o.y);
      kb = lin(u, h6, o, h2);
      o = f(kb);
      u = mkc<R>(ub.x + o.x, ub.y + o.y);
    }
    lam[idx] = u;
  }
}

// ----------------------------------------------------------------- layout conversion
template <typename R>
__global__ void k_timport(const cplx<R>* __restrict__ sol, cplx<R>* __restrict__ S, TracerGeom g) {
  const int64_t total = (int64_t)g.TX * g.KY * g.KZ * kVCB;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % kVCB);
    int64_t r = idx / kVCB;
    const int kz = (int)(r % g.KZ);
    r /= g.KZ;
    const int ky = (int)(r % g.KY);
    const int tx = (int)(r / g.KY);
    const int ix = tx * kVCB + c;
    const int iy = t_full_index(ky, g.ylo, g.yhi), iz = t_full_index(kz, g.zlo, g.zhi);
    S[idx] = ix < g.KX ? sol[((int64_t)iz * g.ny + iy) * g.nkr + ix] : mkc<R>(0, 0);
  }
}

static __global__ void k_tpackL(const double* __restrict__ L, double* __restrict__ Lp, TracerGeom g) {
  const int64_t total = (int64_t)g.TX * g.KY * g.KZ * kVCB;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % kVCB);
    int64_t r = idx / kVCB;
    const int kz = (int)(r % g.KZ);
    r /= g.KZ;
    const int ky = (int)(r % g.KY);
    const int tx = (int)(r / g.KY);
    const int ix = tx * kVCB + c;
    const int iy = t_full_index(ky, g.ylo, g.yhi), iz = t_full_index(kz, g.zlo, g.zhi);
    Lp[idx] = ix < g.KX ? L[((int64_t)iz * g.ny + iy) * g.nkr + ix] : 0.0;
  }
}

// Scatter the kept band back; modes outside it (N = 0 there) take the RK4
// steps of du/dt = L u with the reference's substep roundings.
template <typename R>
__global__ void k_texport(cplx<R>* __restrict__ sol, const cplx<R>* __restrict__ S, const double* __restrict__ L,
                          double h2, double h, double h6, int64_t nsteps, TracerGeom g) {
  const int64_t total = (int64_t)g.nz * g.ny * g.nkr;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(idx % g.nkr);
    const int64_t r = idx / g.nkr;
    const int iy = (int)(r % g.ny), iz = (int)(r / g.ny);
    const bool kept = ix < g.KX && (iy < g.ylo || iy > g.yhi) && (iz < g.zlo || iz > g.zhi);
    if (kept) {
      const int tx = ix / kVCB, c = ix % kVCB;
      const int ky = t_kept_index(iy, g.ylo, g.yhi), kz = t_kept_index(iz, g.zlo, g.zhi);
      sol[idx] = S[(((int64_t)tx * g.KY + ky) * g.KZ + kz) * kVCB + c];
    } else {
      cplx<R> u = sol[idx];
      if (u.x == R(0) && u.y == R(0)) continue;
      const double l = L[idx];
      for (int64_t s = 0; s < nsteps; ++s) {
        auto rhs = [&](cplx<R> x) { return mkc<R>((R)(l * (double)x.x), (R)(l * (double)x.y)); };
        auto sub = [&](cplx<R> b, cplx<R> k, double hh) {
          return mkc<R>((R)((double)b.x + hh * (double)k.x), (R)((double)b.y + hh * (double)k.y));
        };
        const cplx<R> r1 = rhs(u);
        const cplx<R> r2 = rhs(sub(u, r1, h2));
        const cplx<R> r3 = rhs(sub(u, r2, h2));
        const cplx<R> r4 = rhs(sub(u, r3, h));
        double sx = (double)r1.x + 2.0 * (double)r2.x, sy = (double)r1.y + 2.0 * (double)r2.y;
        sx = sx + 2.0 * (double)r3.x;
        sy = sy + 2.0 * (double)r3.y;
        sx = sx + (double)r4.x;
        sy = sy + (double)r4.y;
        u = mkc<R>((R)((double)u.x + h6 * sx), (R)((double)u.y + h6 * sy));
      }
      sol[idx] = u;
    }
  }
}

// ----------------------------------------------------------------- host side
static bool tpow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename R, int N>
static size_t t_lds(int teams) {
  return ((size_t)teams * team_slice(TeamSched<N, vpts_col(N)>::LDS, sizeof(cplx<R>), teams) +
          tw_slots<N, vpts_col(N)>()) * sizeof(cplx<R>);
}

// workgroups for `groups` groups of four XCD-sharing workgroups (xcd_split)
static int t_grid4(int groups) { return (groups + kXCD - 1) / kXCD * kXCD * 4; }

template <typename K>
static int t_attr(K kern, size_t bytes) {
  FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
  return FFHIP_OK;
}

template <typename R, int N, int STAGE>
static int launch_tz(const TracerGeom& g, const TracerArgs& a, hipStream_t st) {
  constexpr int P = vpts_col(N);
  auto kern = k_tz<R, N, P, STAGE>;
  const size_t sh = t_lds<R, N>(kVCB);
  static bool set = false;
  if (!set) {
    FF_TRY(t_attr(kern, sh));
    set = true;
  }
  static const char* names[5] = {"tracer_z_stage0", "tracer_z_stage1", "tracer_z_stage2", "tracer_z_stage3",
                                 "tracer_z_stage4"};
  const double cs = sizeof(cplx<R>);
  const double nk = (double)g.KX * g.KY;
  const double bytes = (STAGE > 0 ? 3 : 0) * nk * g.nz * cs + (a.do_inverse ? nk * g.nz * cs : 0) +
                       nk * g.KZ * (STAGE == 0 ? cs : (STAGE < 4 ? 3 * cs + 8 : 6 * cs + 8));
  prof_launch(names[STAGE], bytes, kern, dim3(t_grid4(g.TX * (g.ny / 4))), dim3(kVCB * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_tz");
  return FFHIP_OK;
}

template <typename R, int N, int DIR, int NF>
static int launch_ty(const TracerGeom& g, const TracerYArgs& a, const char* name, hipStream_t st) {
  constexpr int P = vpts_col(N);
  auto kern = k_ty<R, N, P, DIR, NF>;
  const size_t sh = t_lds<R, N>(kVCB);
  static bool set = false;
  if (!set) {
    FF_TRY(t_attr(kern, sh));
    set = true;
  }
  const double cs = sizeof(cplx<R>);
  const double nk = (double)g.KX * g.nz;
  const double bytes = NF * nk * ((a.src_kept ? g.KY : g.ny) + (a.dst_kept ? g.KY : g.ny)) * cs;
  prof_launch(name, bytes, kern, dim3(t_grid4(g.TX * (g.nz / 4))), dim3(kVCB * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_ty");
  return FFHIP_OK;
}

template <typename R, int N>
static int launch_tx(const TracerGeom& g, const TracerArgs& a, hipStream_t st) {
  constexpr int P = vpts_row(N);
  auto kern = k_tx<R, N, P>;
  const size_t sh =
      ((size_t)4 * team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), 4) + tw_slots<N, P>()) * sizeof(cplx<R>);
  static bool set = false;
  if (!set) {
    FF_TRY(t_attr(kern, sh));
    set = true;
  }
  prof_launch("tracer_x", 4.0 * g.KX * (double)g.ny * g.nz * sizeof(cplx<R>), kern,
              dim3(g.nz * (g.ny / 4)), dim3(4 * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_tx");
  return FFHIP_OK;
}

template <typename R, int STAGE>
static int disp_tz(const TracerGeom& g, const TracerArgs& a, hipStream_t st) {
  switch (g.nz) {
    case 64: return launch_tz<R, 64, STAGE>(g, a, st);
    case 128: return launch_tz<R, 128, STAGE>(g, a, st);
    case 256: return launch_tz<R, 256, STAGE>(g, a, st);
    default: return launch_tz<R, 512, STAGE>(g, a, st);
  }
}

template <typename R>
static int disp_tz_stage(int stage, const TracerGeom& g, const TracerArgs& a, hipStream_t st) {
  switch (stage) {
    case 0: return disp_tz<R, 0>(g, a, st);
    case 1: return disp_tz<R, 1>(g, a, st);
    case 2: return disp_tz<R, 2>(g, a, st);
    case 3: return disp_tz<R, 3>(g, a, st);
    default: return disp_tz<R, 4>(g, a, st);
  }
}

template <typename R, int DIR, int NF>
static int disp_ty(const TracerGeom& g, const TracerYArgs& a, const char* name, hipStream_t st) {
  switch (g.ny) {
    case 64: return launch_ty<R, 64, DIR, NF>(g, a, name, st);
    case 128: return launch_ty<R, 128, DIR, NF>(g, a, name, st);
    case 256: return launch_ty<R, 256, DIR, NF>(g, a, name, st);
    default: return launch_ty<R, 512, DIR, NF>(g, a, name, st);
  }
}

static TracerYArgs yargs(const void* s0, const void* s1, const void* s2, void* d0, void* d1, void* d2,
                         const void* tw, int src_kept, int dst_kept) {
  TracerYArgs y;
  y.src[0] = s0;
  y.src[1] = s1;
  y.src[2] = s2;
  y.dst[0] = d0;
  y.dst[1] = d1;
  y.dst[2] = d2;
  y.tw = tw;
  y.src_kept = src_kept;
  y.dst_kept = dst_kept;
  return y;
}

template <typename R>
static int disp_tx(const TracerGeom& g, const TracerArgs& a, hipStream_t st) {
  switch (g.nx) {
    case 64: return launch_tx<R, 64>(g, a, st);
    case 128: return launch_tx<R, 128>(g, a, st);
    case 256: return launch_tx<R, 256>(g, a, st);
    default: return launch_tx<R, 512>(g, a, st);
  }
}

template <typename R, int N, int STAGE>
static int launch_taz(const TracerGeom& g, const TracerAdjArgs& a, hipStream_t st) {
  constexpr int P = vpts_col(N);
  auto kern = k_taz<R, N, P, STAGE>;
  const size_t sh = t_lds<R, N>(kVCB);
  static bool set = false;
  if (!set) {
    FF_TRY(t_attr(kern, sh));
    set = true;
  }
  static const char* names[5] = {"tracer_adj_z_stage0", "tracer_adj_z_stage1", "tracer_adj_z_stage2",
                                 "tracer_adj_z_stage3", "tracer_adj_z_stage4"};
  const double cs = sizeof(cplx<R>);
  const double nk = (double)g.KX * g.KY;
  const double bytes = (STAGE > 0 ? nk * g.nz * cs : 0) + 3 * nk * g.nz * cs + nk * g.KZ * (4 * cs + 8);
  prof_launch(names[STAGE], bytes, kern, dim3(t_grid4(g.TX * (g.ny / 4))), dim3(kVCB * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_taz");
  return FFHIP_OK;
}

template <typename R, int STAGE>
static int disp_taz(const TracerGeom& g, const TracerAdjArgs& a, hipStream_t st) {
  switch (g.nz) {
    case 64: return launch_taz<R, 64, STAGE>(g, a, st);
    case 128: return launch_taz<R, 128, STAGE>(g, a, st);
    case 256: return launch_taz<R, 256, STAGE>(g, a, st);
    default: return launch_taz<R, 512, STAGE>(g, a, st);
  }
}

template <typename R>
static int disp_taz_stage(int stage, const TracerGeom& g, const TracerAdjArgs& a, hipStream_t st) {
  switch (stage) {
    case 0: return disp_taz<R, 0>(g, a, st);
    case 1: return disp_taz<R, 1>(g, a, st);
    case 2: return disp_taz<R, 2>(g, a, st);
    case 3: return disp_taz<R, 3>(g, a, st);
    default: return disp_taz<R, 4>(g, a, st);
  }
}

template <typename R, int N>
static int launch_tax(const TracerGeom& g, TracerState* s, hipStream_t st) {
  constexpr int P = vpts_row(N);
  auto kern = k_tax<R, N, P>;
  const size_t sh =
      ((size_t)4 * team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), 4) + tw_slots<N, P>()) * sizeof(cplx<R>);
  static bool set = false;
  if (!set) {
    FF_TRY(t_attr(kern, sh));
    set = true;
  }
  prof_launch("tracer_adj_x", 4.0 * g.KX * (double)g.ny * g.nz * sizeof(cplx<R>), kern,
              dim3(g.nz * (g.ny / 4)), dim3(4 * (N / P)), sh, st, g, (const void*)s->Bf[0],
              (const void*)s->Bf[1], (const void*)s->Bf[2], s->A2, s->tab, (const void*)s->twx);
  FF_LAUNCHED("k_tax");
  return FFHIP_OK;
}

template <typename R>
static int disp_tax(const TracerGeom& g, TracerState* s, hipStream_t st) {
  switch (g.nx) {
    case 64: return launch_tax<R, 64>(g, s, st);
    case 128: return launch_tax<R, 128>(g, s, st);
    case 256: return launch_tax<R, 256>(g, s, st);
    default: return launch_tax<R, 512>(g, s, st);
  }
}

static void tstate_free(TracerState* s) {
  if (!s) return;
  void* ps[] = {s->S, s->S1, s->R[0], s->R[1], s->R[2], s->A1, s->A2, s->Bf[0], s->Bf[1], s->Bf[2],
                s->Cf[0], s->Cf[1], s->Cf[2], s->Lp, s->twx, s->twy, s->twz};
  for (void* p : ps) dev_free(p);
  delete s;
}

int tracer_fused_create(ffhip_op* op) {
  op->fused3 = nullptr;
  if (op->kind != FFHIP_OP_TRACER3D) return FFHIP_OK;
  const char* dis = getenv("FFHIP_DISABLE_FUSED");
  if (dis && atoi(dis) != 0) return FFHIP_OK;
  for (int d = 0; d < 3; ++d)
    if (!tpow2(op->n[d]) || op->n[d] < 64 || op->n[d] > 512) return FFHIP_OK;
  if (!(op->af > 0)) return FFHIP_OK;
  TracerGeom g;
  g.nx = (int)op->n[0];
  g.ny = (int)op->n[1];
  g.nz = (int)op->n[2];
  g.nkr = (int)op->nkr;
  g.KX = (int)op->lo[0] - 1;
  g.KXp = (g.KX + kVCB - 1) / kVCB * kVCB;
  g.TX = g.KXp / kVCB;
  g.ylo = (int)op->lo[1] - 1;
  g.yhi = (int)op->hi[1] - 1;
  g.KY = g.ny - (g.yhi - g.ylo + 1);
  g.zlo = (int)op->lo[2] - 1;
  g.zhi = (int)op->hi[2] - 1;
  g.KZ = g.nz - (g.zhi - g.zlo + 1);
  const double twopi = 6.283185307179586476925286766559;
  g.dk = twopi / op->box[0];
  g.dl = twopi / op->box[1];
  g.dm = twopi / op->box[2];
  g.norm = 1.0 / ((double)g.nx * (double)g.ny * (double)g.nz);
  g.A = op->params[0];
  g.B = op->params[1];
  g.C = op->params[2];
  if (g.KX < 1 || g.KX > g.nx / 2 || g.ylo < 1 || g.yhi < g.ylo || g.ylo > g.ny / 2 || g.yhi < g.ny / 2 ||
      g.zlo < 1 || g.zhi < g.zlo || g.zlo > g.nz / 2 || g.zhi < g.nz / 2)
    return FFHIP_OK;
  TracerState* s = new (std::nothrow) TracerState();
  FF_CHECK(s, FFHIP_ERR_OOM, "host allocation failed");
  s->g = g;
  s->tab = reinterpret_cast<const double*>(op->tables);
  const size_t cs = cplx_size(op->prec);
  const size_t sb = (size_t)g.TX * g.KY * g.KZ * kVCB * cs;
  const size_t ab = (size_t)g.nz * g.TX * g.ny * kVCB * cs;
  const size_t cb = (size_t)g.TX * g.KY * g.nz * kVCB * cs;
  int rc = dev_alloc(&s->S, sb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->S1, sb);
  for (int i = 0; i < 3 && rc == FFHIP_OK; ++i) rc = dev_alloc(&s->R[i], sb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->A1, ab);  // zero-filled: its aliased y rows stay zero
  if (rc == FFHIP_OK) rc = dev_alloc(&s->A2, ab);
  for (int i = 0; i < 3 && rc == FFHIP_OK; ++i) rc = dev_alloc(&s->Bf[i], ab);
  for (int i = 0; i < 3 && rc == FFHIP_OK; ++i) rc = dev_alloc(&s->Cf[i], cb);
  if (rc == FFHIP_OK) rc = dev_alloc(&s->Lp, (size_t)g.TX * g.KY * g.KZ * kVCB * sizeof(double));
  if (rc == FFHIP_OK) rc = upload_roots(g.nx, g.nx, op->prec, &s->twx);
  if (rc == FFHIP_OK) rc = upload_roots(g.ny, g.ny, op->prec, &s->twy);
  if (rc == FFHIP_OK) rc = upload_roots(g.nz, g.nz, op->prec, &s->twz);
  if (rc != FFHIP_OK) {
    tstate_free(s);
    return rc;
  }
  op->fused3 = s;
  return FFHIP_OK;
}

void tracer_fused_destroy(ffhip_op* op) {
  tstate_free(reinterpret_cast<TracerState*>(op->fused3));
  op->fused3 = nullptr;
}

template <typename R>
static int trun(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, hipStream_t st) {
  TracerState* s = reinterpret_cast<TracerState*>(op->fused3);
  const TracerGeom& g = s->g;
  const int64_t n = (int64_t)g.TX * g.KY * g.KZ * kVCB;
  if (s->lp_owner != ts) {
    hipLaunchKernelGGL(k_tpackL, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const double*>(ts->L), reinterpret_cast<double*>(s->Lp), g);
    FF_LAUNCHED("k_tpackL");
    s->lp_owner = ts;
  }
  hipLaunchKernelGGL((k_timport<R>), dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(sol), reinterpret_cast<cplx<R>*>(s->S), g);
  FF_LAUNCHED("k_timport");
  TracerArgs a;
  a.S = s->S;
  a.S1 = s->S1;
  a.R0 = s->R[0];
  a.R1 = s->R[1];
  a.R2 = s->R[2];
  a.A1 = s->A1;
  a.A2 = s->A2;
  a.B0 = s->Bf[0];
  a.B1 = s->Bf[1];
  a.B2 = s->Bf[2];
  a.C0 = s->Cf[0];
  a.C1 = s->Cf[1];
  a.C2 = s->Cf[2];
  a.L = reinterpret_cast<const double*>(s->Lp);
  a.tab = s->tab;
  a.twx = s->twx;
  a.twy = s->twy;
  a.twz = s->twz;
  // step sizes as rk4_run_generic takes them (dt already rounded to the state type)
  a.h = dt;
  a.h2 = op->prec == FFHIP_F32 ? (double)((float)dt / 2.0f) : dt / 2;
  a.h6 = dt / 6;
  a.do_inverse = 1;
  const TracerYArgs yinv = yargs(s->A1, nullptr, nullptr, s->A2, nullptr, nullptr, s->twy, 1, 0);
  const TracerYArgs yfwd = yargs(s->Bf[0], s->Bf[1], s->Bf[2], s->Cf[0], s->Cf[1], s->Cf[2], s->twy, 0, 1);
  FF_TRY(disp_tz_stage<R>(0, g, a, st));
  for (int64_t k = 0; k < nsteps; ++k) {
    for (int stage = 1; stage <= 4; ++stage) {
      FF_TRY((disp_ty<R, 1, 1>(g, yinv, "tracer_y_inverse", st)));
      FF_TRY(disp_tx<R>(g, a, st));
      FF_TRY((disp_ty<R, -1, 3>(g, yfwd, "tracer_y_forward", st)));
      a.do_inverse = (stage < 4 || k + 1 < nsteps) ? 1 : 0;
      FF_TRY(disp_tz_stage<R>(stage, g, a, st));
    }
  }
  const int64_t total = (int64_t)g.nz * g.ny * g.nkr;
  hipLaunchKernelGGL((k_texport<R>), dim3(blocks_for(total, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<cplx<R>*>(sol), reinterpret_cast<const cplx<R>*>(s->S),
                     reinterpret_cast<const double*>(ts->L), a.h2, a.h, a.h6, nsteps, g);
  FF_LAUNCHED("k_texport");
  return FFHIP_OK;
}

// lam <- (d sol_T / d sol_0)^T lam over nsteps fused RK4 steps
template <typename R>
static int tadj(ffhip_op* op, ffhip_rk4* ts, void* lam, double dt, int64_t nsteps, hipStream_t st) {
  TracerState* s = reinterpret_cast<TracerState*>(op->fused3);
  const TracerGeom& g = s->g;
  const int64_t n = (int64_t)g.TX * g.KY * g.KZ * kVCB;
  if (s->lp_owner != ts) {
    hipLaunchKernelGGL(k_tpackL, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const double*>(ts->L), reinterpret_cast<double*>(s->Lp), g);
    FF_LAUNCHED("k_tpackL");
    s->lp_owner = ts;
  }
  hipLaunchKernelGGL((k_timport<R>), dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(lam), reinterpret_cast<cplx<R>*>(s->S), g);
  FF_LAUNCHED("k_timport");
  TracerAdjArgs a;
  a.LAM = s->S;
  a.UB = s->S1;
  a.KB = s->R[0];
  a.F = s->Cf[0];
  a.Z0 = s->Bf[0];
  a.Z1 = s->Bf[1];
  a.Z2 = s->Bf[2];
  a.L = reinterpret_cast<const double*>(s->Lp);
  a.twz = s->twz;
  a.h = dt;
  a.h2 = op->prec == FFHIP_F32 ? (double)((float)dt / 2.0f) : dt / 2;
  a.h6 = dt / 6;
  a.h3 = 2 * (dt / 6);
  a.cont = 1;
  // y-inverse of the three fields in place; y-forward of the product into F
  const TracerYArgs yinv = yargs(s->Bf[0], s->Bf[1], s->Bf[2], s->Bf[0], s->Bf[1], s->Bf[2], s->twy, 1, 0);
  const TracerYArgs yfwd = yargs(s->A2, nullptr, nullptr, s->Cf[0], nullptr, nullptr, s->twy, 0, 1);
  FF_TRY(disp_taz_stage<R>(0, g, a, st));
  for (int64_t k = 0; k < nsteps; ++k) {
    for (int stage = 1; stage <= 4; ++stage) {
      FF_TRY((disp_ty<R, 1, 3>(g, yinv, "tracer_adj_y_inverse", st)));
      FF_TRY(disp_tax<R>(g, s, st));
      FF_TRY((disp_ty<R, -1, 1>(g, yfwd, "tracer_adj_y_forward", st)));
      a.cont = (stage < 4 || k + 1 < nsteps) ? 1 : 0;
      FF_TRY(disp_taz_stage<R>(stage, g, a, st));
    }
  }
  const int64_t total = (int64_t)g.nz * g.ny * g.nkr;
  hipLaunchKernelGGL((k_taexport<R>), dim3(blocks_for(total, 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<cplx<R>*>(lam), reinterpret_cast<const cplx<R>*>(s->S),
                     reinterpret_cast<const double*>(ts->L), a.h2, a.h, a.h3, a.h6, nsteps, g);
  FF_LAUNCHED("k_taexport");
  return FFHIP_OK;
}

int tracer_fused_rk4_adjoint(ffhip_op* op, ffhip_rk4* ts, void* lam, double dt, int64_t nsteps, hipStream_t st) {
  FF_CHECK(op->fused3, FFHIP_ERR_UNSUPPORTED, "no fused tracer path for this operator");
  return op->prec == FFHIP_F32 ? tadj<float>(op, ts, lam, dt, nsteps, st) : tadj<double>(op, ts, lam, dt, nsteps, st);
}

int tracer_fused_rk4_run(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, hipStream_t st) {
  FF_CHECK(op->fused3, FFHIP_ERR_UNSUPPORTED, "no fused tracer path for this operator");
  return op->prec == FFHIP_F32 ? trun<float>(op, ts, sol, dt, nsteps, st) : trun<double>(op, ts, sol, dt, nsteps, st);
}

}  // namespace ffhip
