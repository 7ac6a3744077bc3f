// Discrete adjoint of the fused 2D vorticity ETDRK4 step, with
// checkpoint-and-recompute over the time window.  No counterpart in the
// reference (it has no adjoint); the FP64 oracle adjoint restates the same
// algebra.
//
// Inner product: Re sum conj(a) b over the stored rfft entries.  With it,
//   ldiv!^T(phi) = W * mul!(phi) / (nx ny),   mul!^T(Y) = (nx ny) ldiv!(Y / W),
// W = 1 on the kx = 0 column and 2 elsewhere (the Nyquist column is outside
// the kept band).  For N(s) = P(Dx mul(u z) + Dy mul(v z)) with
// u = ldiv(Du P s), v = ldiv(Dv P s), z = ldiv(P s), the transposed
// linearisation at s is
//   p = mul^T(conj(Dx) P lam), q = mul^T(conj(Dy) P lam)              [physical]
//   m = conj(Du) ldiv^T(p z) + conj(Dv) ldiv^T(q z) + ldiv^T(p u + q v)
// and it needs the forward fields u, v, z of the stage, which the recompute
// stores.  Per adjoint N the kernels are:
//   adjoint column kernel: forward y-FFTs of the three x-transformed terms,
//     m, the reverse ETDRK4 stage combination (FP64), inverse y-FFTs of
//     i kx s / W and i l s / W for the next stage's input s
//   adjoint row kernel: one inverse x-FFT of p + i q, the products with the
//     stored fields ((p + i q) zeta and Re((p + i q) conj(u + i v))), and the
//     forward x-FFTs back to half spectra, scaled by W / (nx ny).
// Reverse step (a, b, c the forward stage states, coefficients real):
//   ub = E lam; N1b = alpha lam; N2b = N3b = 2 beta lam; N4b = Gamma lam
//   cb = dN(c)^T N4b; ab = E/2 cb; N3b += 2 zeta cb; N1b -= zeta cb
//   bb = dN(b)^T N3b; ub += E/2 bb; N2b += zeta bb
//   ab += dN(a)^T N2b; ub += E/2 ab; N1b += zeta ab
//   lam_prev = ub + dN(u)^T N1b
#include <atomic>
#include <cstring>
#include <new>

#include "vort2d.h"

namespace ffhip {

struct VortAdjColArgs {
  const void* OA0;  // x-forward terms of the row kernel
  const void* OA1;
  const void* OA2;
  void* IA0;  // next inputs, y-inverted
  void* IA1;
  void* UB;
  void* N1B;
  void* N2B;
  void* AB;
  void* LAM;
  const double* E;
  const double* Eh;
  const double* Z;
  const double* Al;
  const double* Be;
  const double* Ga;
  const void* tw;
  int cont;  // STAGE 4: 1 = start the next (earlier) step, 0 = store lambda_0 in LAM
  const double* Fl;  // filter of a FilteredETDRK4 stepper (lambda <- filter lambda first), or null
};

struct VortAdjRowArgs {
  const void* IA0;
  const void* IA1;
  void* OA0;
  void* OA1;
  void* OA2;
  const void* F;  // stored fields of the forward stage: u + i v, then zeta
  const void* tw;
};

// STAGE 0: entry from LAM.  STAGE 1..4: after dN(c)^T, dN(b)^T, dN(a)^T, dN(u)^T.
#ifndef FF_ACOL_PAIR
#define FF_ACOL_PAIR 1  // FP32 adjoint column pass up to 1024: paired transforms
#endif
template <typename R, int N>
constexpr bool acol_pair() {
  return sizeof(R) == 4 && N <= 1024 && FF_ACOL_PAIR;
}

template <typename R, int NY, int P, int STAGE>
__global__ void __launch_bounds__(kCBK*(NY / P)) k_vacol(VortGeom g, VortAdjColArgs a) {
  constexpr int T = NY / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char vsm[];
  int tx, sub;
  xcd_split(blockIdx.x, kCSUB, &tx, &sub);
  if (tx >= g.TX) return;
  const int t = threadIdx.x;
  const int cc = t % kCBK;
  const int c = sub * kCBK + cc;
  const int tid = t / kCBK;
  constexpr int SL = team_slice(TeamSched<NY, P>::LDS, sizeof(cplx<R>), kCBK);
  constexpr bool pair = acol_pair<R, NY>();  // second buffers after the first, then the table
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(vsm) + cc * SL;
  cplx<R>* lds2 = pair ? reinterpret_cast<cplx<R>*>(vsm) + (kCBK + cc) * SL : nullptr;
  const cplx<R>* tw = team_tw_load<NY, P, R>(reinterpret_cast<cplx<R>*>(vsm) + (pair ? 2 : 1) * kCBK * SL, a.tw);
  const int ix = tx * CB + c;
  const bool colkept = ix < g.KX;
  const R kx = (R)(g.dk * (double)ix);
  const uint32_t ibase = (uint32_t)tx * NY * CB + c;  // 32-bit offsets: uniform base + VGPR offset loads
  const uint32_t sbase = (uint32_t)tx * g.KY * CB + c;

  // m = -i l/|k|^2 U + i kx/|k|^2 V + Z after the forward y-FFTs
  cplx<R> mb[P];
  if constexpr (STAGE > 0) {
    const cplx<R>* O[3] = {reinterpret_cast<const cplx<R>*>(a.OA0), reinterpret_cast<const cplx<R>*>(a.OA1),
                           reinterpret_cast<const cplx<R>*>(a.OA2)};
#pragma unroll
    for (int p = 0; p < P; ++p) mb[p] = mkc<R>(0, 0);
    if constexpr (pair) {  // U and V side by side, then Z
      cplx<R> v0[P], v1[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        v0[p] = O[0][ibase + (uint32_t)(tid + p * T) * CB];
        v1[p] = O[1][ibase + (uint32_t)(tid + p * T) * CB];
      }
      team_fft2<NY, P, -1, R>(v0, v1, lds, lds2, tid, tw);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int iy = tid + p * T;
        const R l = (R)(g.dl * (double)vfreq(iy, NY));
        const R K2 = kx * kx + l * l;
        const R invK = (ix == 0 && iy == 0) ? R(0) : R(1) / K2;
        const R q0 = l * invK, q1 = -kx * invK;
        mb[p] = mkc<R>(mb[p].x + q0 * v0[p].y, mb[p].y - q0 * v0[p].x);
        mb[p] = mkc<R>(mb[p].x + q1 * v1[p].y, mb[p].y - q1 * v1[p].x);
      }
    }
#pragma unroll 1
    for (int f = pair ? 2 : 0; f < 3; ++f) {
      cplx<R> v[P];
#pragma unroll
      for (int p = 0; p < P; ++p) v[p] = O[f][ibase + (uint32_t)(tid + p * T) * CB];
      team_fft<NY, P, -1, R>(v, lds, tid, tw);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int iy = tid + p * T;
        if (f == 2) {
          mb[p] = cadd<R>(mb[p], v[p]);
        } else {
          const R l = (R)(g.dl * (double)vfreq(iy, NY));
          const R K2 = kx * kx + l * l;
          const R invK = (ix == 0 && iy == 0) ? R(0) : R(1) / K2;
          const R q = (f == 0 ? l : -kx) * invK;
          // f = 0: -i q U  (q = l/|k|^2);  f = 1: +i kx/|k|^2 V = -i q V (q = -kx/|k|^2)
          mb[p] = mkc<R>(mb[p].x + q * v[p].y, mb[p].y - q * v[p].x);
        }
      }
    }
  }

  cplx<R> nx_[P];  // next adjoint-N input
  cplx<R>* UB = reinterpret_cast<cplx<R>*>(a.UB);
  cplx<R>* N1B = reinterpret_cast<cplx<R>*>(a.N1B);
  cplx<R>* N2B = reinterpret_cast<cplx<R>*>(a.N2B);
  cplx<R>* AB = reinterpret_cast<cplx<R>*>(a.AB);
  cplx<R>* LAM = reinterpret_cast<cplx<R>*>(a.LAM);
  bool any_next = true;
  if constexpr (STAGE == 4) any_next = a.cont != 0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int iy = tid + p * T;
    const bool kept = colkept && (iy < g.ylo || iy > g.yhi);
    nx_[p] = mkc<R>(0, 0);
    if (!kept) continue;
    const uint32_t si = sbase + (uint32_t)v_ky(iy, g) * CB;
    double lx = 0, ly = 0;  // lambda for the entry combination
    bool entry = false;
    if constexpr (STAGE == 0) {
      const cplx<R> l0 = LAM[si];
      lx = l0.x;
      ly = l0.y;
      entry = true;
    } else if constexpr (STAGE == 1) {
      const double cx = mb[p].x, cy = mb[p].y;
      const double eh = a.Eh[si], z = a.Z[si];
      AB[si] = mkc<R>((R)(eh * cx), (R)(eh * cy));
      const cplx<R> n2 = N2B[si], n1 = N1B[si];
      nx_[p] = mkc<R>((R)((double)n2.x + 2 * z * cx), (R)((double)n2.y + 2 * z * cy));  // N3b
      N1B[si] = mkc<R>((R)((double)n1.x - z * cx), (R)((double)n1.y - z * cy));
    } else if constexpr (STAGE == 2) {
      const double bx = mb[p].x, by = mb[p].y;
      const double eh = a.Eh[si], z = a.Z[si];
      const cplx<R> ub = UB[si], n2 = N2B[si];
      UB[si] = mkc<R>((R)((double)ub.x + eh * bx), (R)((double)ub.y + eh * by));
      nx_[p] = mkc<R>((R)((double)n2.x + z * bx), (R)((double)n2.y + z * by));  // N2b
    } else if constexpr (STAGE == 3) {
      const cplx<R> ab = AB[si], ub = UB[si], n1 = N1B[si];
      const double ax = (double)ab.x + (double)mb[p].x, ay = (double)ab.y + (double)mb[p].y;
      const double eh = a.Eh[si], z = a.Z[si];
      UB[si] = mkc<R>((R)((double)ub.x + eh * ax), (R)((double)ub.y + eh * ay));
      nx_[p] = mkc<R>((R)((double)n1.x + z * ax), (R)((double)n1.y + z * ay));  // N1b
    } else {
      const cplx<R> ub = UB[si];
      lx = (double)ub.x + (double)mb[p].x;
      ly = (double)ub.y + (double)mb[p].y;
      if (a.cont) {
        entry = true;
        // lambda is rounded to the state type, as a stored adjoint state would be
        lx = (double)(R)lx;
        ly = (double)(R)ly;
      } else {
        LAM[si] = mkc<R>((R)lx, (R)ly);
      }
    }
    if (entry) {
      if (a.Fl) {  // transpose of sol .*= filter
        lx *= a.Fl[si];
        ly *= a.Fl[si];
      }
      const double e = a.E[si], al = a.Al[si], be = 2.0 * a.Be[si], ga = a.Ga[si];
      UB[si] = mkc<R>((R)(e * lx), (R)(e * ly));
      N1B[si] = mkc<R>((R)(al * lx), (R)(al * ly));
      N2B[si] = mkc<R>((R)(be * lx), (R)(be * ly));
      nx_[p] = mkc<R>((R)(ga * lx), (R)(ga * ly));  // N4b
    }
  }
  if (!any_next) return;  // uniform over the block

  // inverse y-FFTs of i kx s / W and i l s / W (W = 1 at kx = 0, 2 otherwise)
  const R invW = ix == 0 ? R(1) : R(0.5);
  cplx<R>* dst[2] = {reinterpret_cast<cplx<R>*>(a.IA0), reinterpret_cast<cplx<R>*>(a.IA1)};
  if constexpr (pair) {
    cplx<R> v0[P], v1[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      const R q0 = kx * invW, q1 = (R)(g.dl * (double)vfreq(iy, NY)) * invW;
      v0[p] = mkc<R>(-q0 * nx_[p].y, q0 * nx_[p].x);  // i q s
      v1[p] = mkc<R>(-q1 * nx_[p].y, q1 * nx_[p].x);
    }
    team_fft2<NY, P, 1, R>(v0, v1, lds, lds2, tid, tw);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      dst[0][ibase + (uint32_t)(tid + p * T) * CB] = v0[p];
      dst[1][ibase + (uint32_t)(tid + p * T) * CB] = v1[p];
    }
    return;
  }
#pragma unroll 1
  for (int f = 0; f < 2; ++f) {
    cplx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      const R q = (f == 0 ? kx : (R)(g.dl * (double)vfreq(iy, NY))) * invW;
      v[p] = mkc<R>(-q * nx_[p].y, q * nx_[p].x);  // i q s
    }
    team_fft<NY, P, 1, R>(v, lds, tid, tw);
    cplx<R>* d = dst[f];
#pragma unroll
    for (int p = 0; p < P; ++p) d[ibase + (uint32_t)(tid + p * T) * CB] = v[p];
  }
}

template <typename R, int NX, int P>
__global__ void __launch_bounds__(kRRK*(NX / P)) k_varow(VortGeom g, VortAdjRowArgs a) {
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  extern __shared__ unsigned char vsm[];
  int grp, sub;
  xcd_split(blockIdx.x, kRSUB, &grp, &sub);
  if (grp >= g.ny / 4) return;
  const int t = threadIdx.x;
  const int r = t % kRRK;
  const int tid = t / kRRK;
  constexpr int SL = team_slice(TeamSched<NX, P>::LDS, sizeof(cplx<R>), kRRK);
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(vsm) + r * SL;
  const cplx<R>* tw = team_tw_load<NX, P, R>(reinterpret_cast<cplx<R>*>(vsm) + kRRK * SL, a.tw);
  const int y = grp * 4 + sub * kRRK + r;
  const int64_t ny = g.ny;
  auto iaddr = [&](int k) -> uint32_t { return ((uint32_t)(k / CB) * (uint32_t)ny + y) * CB + (k % CB); };
  const cplx<R>* Ph = reinterpret_cast<const cplx<R>*>(a.IA0);
  const cplx<R>* Qh = reinterpret_cast<const cplx<R>*>(a.IA1);
  const int KX = g.KX;

  // p + i q = c2r of the two half spectra, packed (unnormalised inverses are mul!^T exactly)
  cplx<R> z[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {  // branch-free: all loads issue before the first wait
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;  // direct / mirrored (conjugate) source
    const int kk = lo ? k : (hi ? NX - k : 0);
    const cplx<R> A = at32(Ph, iaddr(kk)), B = at32(Qh, iaddr(kk));
    const R ay = k == 0 ? R(0) : A.y, by = k == 0 ? R(0) : B.y;  // Im ignored at kx = 0
    z[p] = mkc<R>(lo ? A.x - by : (hi ? A.x + B.y : R(0)), lo ? ay + B.x : (hi ? B.x - A.y : R(0)));
  }
  team_fft<NX, P, 1, R>(z, lds, tid, tw);

  // products with the stored forward fields
  const cplx<R>* Fz = reinterpret_cast<const cplx<R>*>(a.F) + (int64_t)y * NX;
  const R* Fr = reinterpret_cast<const R*>(reinterpret_cast<const cplx<R>*>(a.F) + (int64_t)g.nx * g.ny) +
                (int64_t)y * NX;
  cplx<R> w2[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const cplx<R> uv = Fz[tid + p * T];
    const R zr = Fr[tid + p * T];
    w2[p] = mkc<R>(z[p].x * uv.x + z[p].y * uv.y, 0);  // p u + q v
    z[p] = mkc<R>(z[p].x * zr, z[p].y * zr);          // p zeta + i q zeta
  }
  team_fft<NX, P, -1, R>(z, lds, tid, tw);
  const R cw = (R)(1.0 / ((double)g.nx * (double)g.ny));  // W / (nx ny), W applied per k below
  cplx<R>* OU = reinterpret_cast<cplx<R>*>(a.OA0);
  cplx<R>* OV = reinterpret_cast<cplx<R>*>(a.OA1);
  cplx<R>* OZ = reinterpret_cast<cplx<R>*>(a.OA2);
#pragma unroll
  for (int p = 0; p < P; ++p) lds[lpad(tid + p * T)] = z[p];
  __syncthreads();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    cplx<R> Pk = mkc<R>(0, 0), Qk = mkc<R>(0, 0);
    if (k < KX) {
      const R s = (k == 0 ? R(1) : R(2)) * cw * R(0.5);
      const cplx<R> wk = z[p];
      const cplx<R> wc = cconj<R>(lds[lpad((NX - k) & (NX - 1))]);
      Pk = mkc<R>(s * (wk.x + wc.x), s * (wk.y + wc.y));
      Qk = mkc<R>(s * (wk.y - wc.y), -s * (wk.x - wc.x));
    }
    OU[iaddr(k)] = Pk;
    OV[iaddr(k)] = Qk;
  }
  __syncthreads();
  team_fft<NX, P, -1, R>(w2, lds, tid, tw);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    cplx<R> v = mkc<R>(0, 0);
    if (k < KX) v = cscale<R>(w2[p], (k == 0 ? R(1) : R(2)) * cw);
    OZ[iaddr(k)] = v;
  }
}

// ----------------------------------------------------------------- launchers
template <typename R, int N, int STAGE>
static int launch_acol(const VortGeom& g, const VortAdjColArgs& a, hipStream_t st) {
  constexpr int P = vcol_pts<R, N>();
  constexpr size_t sl = team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), kCBK);
  const size_t sh = ((acol_pair<R, N>() ? 2 : 1) * kCBK * sl + tw_slots<N, P>()) * sizeof(cplx<R>);
  auto kern = k_vacol<R, N, P, STAGE>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  static const char* names[5] = {"vort_adj_col_stage0", "vort_adj_col_stage1", "vort_adj_col_stage2",
                                 "vort_adj_col_stage3", "vort_adj_col_stage4"};
  const double cs = sizeof(cplx<R>), ds = sizeof(double);
  const double ni = (double)g.KX * g.ny, nc = (double)g.KX * g.KY;
  const double bytes = (STAGE == 4 && !a.cont ? 0 : 2) * ni * cs + (STAGE > 0 ? 3 * ni * cs : 0) +
                       nc * (STAGE == 0 ? 4 * cs + 4 * ds : STAGE == 1 ? 5 * cs + 2 * ds
                                          : STAGE == 2 ? 3 * cs + 2 * ds : STAGE == 3 ? 5 * cs + 2 * ds
                                          : (a.cont ? 4 * cs + 4 * ds : 2 * cs));
  const int groups = (g.TX + kXCD - 1) / kXCD * kXCD;
  prof_launch(names[STAGE], bytes, kern, dim3(groups * kCSUB), dim3(kCBK * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_vacol");
  return FFHIP_OK;
}

template <typename R, int N>
static int launch_arow(const VortGeom& g, const VortAdjRowArgs& a, hipStream_t st) {
  constexpr int P = vrow_pts<R, N>();
  const size_t sh =
      ((size_t)kRRK * team_slice(TeamSched<N, P>::LDS, sizeof(cplx<R>), kRRK) + tw_slots<N, P>()) * sizeof(cplx<R>);
  auto kern = k_varow<R, N, P>;
  static bool attr_set = false;
  if (!attr_set) {
    FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sh));
    attr_set = true;
  }
  const double bytes = 5.0 * g.KX * g.ny * sizeof(cplx<R>) + (double)vfield_bytes<R>(g);
  const int groups = (g.ny / 4 + kXCD - 1) / kXCD * kXCD;
  prof_launch("vort_adj_row", bytes, kern, dim3(groups * kRSUB), dim3(kRRK * (N / P)), sh, st, g, a);
  FF_LAUNCHED("k_varow");
  return FFHIP_OK;
}

template <typename R, int STAGE>
static int adispatch_col(const VortGeom& g, const VortAdjColArgs& a, hipStream_t st) {
  switch (g.ny) {
    case 64: return launch_acol<R, 64, STAGE>(g, a, st);
    case 128: return launch_acol<R, 128, STAGE>(g, a, st);
    case 256: return launch_acol<R, 256, STAGE>(g, a, st);
    case 512: return launch_acol<R, 512, STAGE>(g, a, st);
    case 1024: return launch_acol<R, 1024, STAGE>(g, a, st);
    case 2048: return launch_acol<R, 2048, STAGE>(g, a, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fused adjoint column kernel: ny = %d", g.ny);
  }
}

template <typename R>
static int adispatch_stage(int stage, const VortGeom& g, const VortAdjColArgs& a, hipStream_t st) {
  switch (stage) {
    case 0: return adispatch_col<R, 0>(g, a, st);
    case 1: return adispatch_col<R, 1>(g, a, st);
    case 2: return adispatch_col<R, 2>(g, a, st);
    case 3: return adispatch_col<R, 3>(g, a, st);
    default: return adispatch_col<R, 4>(g, a, st);
  }
}

template <typename R>
static int adispatch_row(const VortGeom& g, const VortAdjRowArgs& a, hipStream_t st) {
  switch (g.nx) {
    case 64: return launch_arow<R, 64>(g, a, st);
    case 128: return launch_arow<R, 128>(g, a, st);
    case 256: return launch_arow<R, 256>(g, a, st);
    case 512: return launch_arow<R, 512>(g, a, st);
    case 1024: return launch_arow<R, 1024>(g, a, st);
    case 2048: return launch_arow<R, 2048>(g, a, st);
    default: return fail(FFHIP_ERR_UNSUPPORTED, "fused adjoint row kernel: nx = %d", g.nx);
  }
}

// ----------------------------------------------------------------- driver
static int ensure_adjoint_buffers(VortState* s, int prec, int64_t ncheck, int64_t seg) {
  const VortGeom& g = s->g;
  const size_t cs = cplx_size(prec);
  const size_t sb = (size_t)g.TX * g.KY * kVCB * cs;
  const size_t ib = (size_t)g.TX * g.ny * kVCB * cs;
  void** comp[5] = {&s->UB, &s->N1B, &s->N2B, &s->AB, &s->LAM};
  for (void** p : comp)
    if (!*p) FF_TRY(dev_alloc(p, sb));
  for (auto& p : s->IA)
    if (!p) FF_TRY(dev_alloc(&p, ib));
  for (auto& p : s->OA)
    if (!p) FF_TRY(dev_alloc(&p, ib));
  if (s->ck_cap < ncheck) {
    dev_free(s->CK);
    s->CK = nullptr;
    s->ck_cap = 0;
    FF_TRY(dev_alloc(&s->CK, (size_t)ncheck * sb));
    s->ck_cap = ncheck;
  }
  const size_t fb = prec == FFHIP_F32 ? vfield_bytes<float>(g) : vfield_bytes<double>(g);
  if (s->fs_cap < seg) {
    dev_free(s->FS);
    s->FS = nullptr;
    s->fs_cap = 0;
    FF_TRY(dev_alloc(&s->FS, (size_t)seg * 4 * fb));
    s->fs_cap = seg;
  }
  return FFHIP_OK;
}

// run identifiers of checkpointed forward runs, unique over all operators of the process
static std::atomic<int64_t> g_ck_seq{0};

struct VortAdjPlan {
  int64_t nsteps = 0, ncheck = 0, seg = 0;
};

static VortColArgs fwd_col_args(VortState* s, const double* pc, size_t n, bool filtered) {
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
  ca.Fl = filtered ? pc + 6 * n : nullptr;
  return ca;
}

static VortRowArgs fwd_row_args(VortState* s) {
  VortRowArgs ra;
  ra.I0 = s->I[0];
  ra.I1 = s->I[1];
  ra.I2 = s->I[2];
  ra.O0 = s->O[0];
  ra.O1 = s->O[1];
  ra.tw = s->twx;
  ra.F = nullptr;
  return ra;
}

// forward steps [k0, k1) from the state in U (whose y-inverse is already in I);
// store_fields: keep the four stage fields of every step in FS
static int fwd_steps(int prec, VortState* s, VortColArgs& ca, VortRowArgs& ra, int64_t k0, int64_t k1,
                     bool last_inverse, bool store_fields, const size_t fb, hipStream_t st) {
  for (int64_t k = k0; k < k1; ++k) {
    for (int stage = 1; stage <= 4; ++stage) {
      ra.F = store_fields ? (void*)((char*)s->FS + ((k - k0) * 4 + (stage - 1)) * fb) : nullptr;
      FF_TRY(vort_row(prec, s->g, ra, st));
      ca.do_inverse = (stage < 4 || k + 1 < k1 || last_inverse) ? 1 : 0;
      FF_TRY(vort_col(prec, stage, s->g, ca, st));
    }
  }
  ra.F = nullptr;
  return FFHIP_OK;
}

template <typename R>
static int vforward_ckpt(ffhip_op* op, ffhip_etdrk4* ts, const void* sol0, void* solT, int64_t nsteps,
                         int64_t ncheck, hipStream_t st) {
  VortState* s = reinterpret_cast<VortState*>(op->fused);
  const VortGeom& g = s->g;
  const int prec = op->prec;
  const int64_t seg = (nsteps + ncheck - 1) / ncheck;
  s->ck_ts = nullptr;
  s->ck_nsteps = -1;
  s->ck_id = 0;
  FF_TRY(ensure_adjoint_buffers(s, prec, ncheck, seg));
  FF_TRY(vort_ensure_packed(op, ts, g, st));
  const size_t n = (size_t)g.TX * g.KY * kVCB;
  const size_t sb = n * sizeof(cplx<R>);
  const double* pc = reinterpret_cast<const double*>(ts->packed);
  VortColArgs ca = fwd_col_args(s, pc, n, ts->filter != nullptr);
  VortRowArgs ra = fwd_row_args(s);
  const size_t fb = vfield_bytes<R>(g);
  FF_TRY(vort_import(prec, g, sol0, s->U, st));
  FF_TRY(vort_col(prec, 0, g, ca, st));
  s->fs_last = false;
  for (int64_t k0 = 0, c = 0; k0 < nsteps; k0 += seg, ++c) {
    FF_HIP(hipMemcpyAsync((char*)s->CK + c * sb, s->U, sb, hipMemcpyDeviceToDevice, st));
    const int64_t k1 = k0 + seg < nsteps ? k0 + seg : nsteps;
    // the last segment keeps its stage fields, so the adjoint sweep starts without a recompute
    FF_TRY(fwd_steps(prec, s, ca, ra, k0, k1, k1 < nsteps, k1 == nsteps, fb, st));
  }
  s->fs_last = true;
  if (solT != sol0)
    FF_HIP(hipMemcpyAsync(solT, sol0, (size_t)g.ny * g.nkr * sizeof(cplx<R>), hipMemcpyDeviceToDevice, st));
  FF_TRY(vort_export(prec, g, solT, s->U, reinterpret_cast<const double*>(ts->coef[4]),
                     reinterpret_cast<const double*>(ts->filter), nsteps, st));
  s->ck_ts = ts;
  s->ck_nsteps = nsteps;
  s->ck_ncheck = ncheck;
  s->ck_id = ++g_ck_seq;
  return FFHIP_OK;
}

template <typename R>
static int vadjoint(ffhip_op* op, ffhip_etdrk4* ts, const void* lamT, void* lam0, int64_t nsteps, int64_t ncheck,
                    hipStream_t st) {
  VortState* s = reinterpret_cast<VortState*>(op->fused);
  const VortGeom& g = s->g;
  const int prec = op->prec;
  const int64_t seg = (nsteps + ncheck - 1) / ncheck;
  FF_CHECK(s->CK && s->ck_cap >= ncheck && s->FS && s->fs_cap >= seg && s->ck_ts == ts &&
               s->ck_nsteps == nsteps && s->ck_ncheck == ncheck,
           FFHIP_ERR_ARG, "the adjoint needs the checkpointed forward run of this operator and stepper with the same "
                          "nsteps and ncheck just before it");
  FF_TRY(vort_ensure_packed(op, ts, g, st));
  const size_t n = (size_t)g.TX * g.KY * kVCB;
  const size_t sb = n * sizeof(cplx<R>);
  const double* pc = reinterpret_cast<const double*>(ts->packed);
  VortColArgs ca = fwd_col_args(s, pc, n, ts->filter != nullptr);
  VortRowArgs ra = fwd_row_args(s);
  const size_t fb = vfield_bytes<R>(g);

  VortAdjColArgs aa;
  aa.OA0 = s->OA[0];
  aa.OA1 = s->OA[1];
  aa.OA2 = s->OA[2];
  aa.IA0 = s->IA[0];
  aa.IA1 = s->IA[1];
  aa.UB = s->UB;
  aa.N1B = s->N1B;
  aa.N2B = s->N2B;
  aa.AB = s->AB;
  aa.LAM = s->LAM;
  aa.Z = pc + 0 * n;
  aa.Al = pc + 1 * n;
  aa.Be = pc + 2 * n;
  aa.Ga = pc + 3 * n;
  aa.E = pc + 4 * n;
  aa.Eh = pc + 5 * n;
  aa.tw = s->twy;
  aa.cont = 1;
  aa.Fl = ts->filter ? pc + 6 * n : nullptr;
  VortAdjRowArgs ar;
  ar.IA0 = s->IA[0];
  ar.IA1 = s->IA[1];
  ar.OA0 = s->OA[0];
  ar.OA1 = s->OA[1];
  ar.OA2 = s->OA[2];
  ar.tw = s->twx;
  ar.F = nullptr;

  FF_TRY(vort_import(prec, g, lamT, s->LAM, st));
  FF_TRY((adispatch_stage<R>(0, g, aa, st)));
  const int64_t nseg = (nsteps + seg - 1) / seg;
  for (int64_t c = nseg - 1; c >= 0; --c) {
    const int64_t k0 = c * seg;
    const int64_t k1 = k0 + seg < nsteps ? k0 + seg : nsteps;
    // recompute the segment, keeping every stage's fields (the forward run kept the last one's)
    if (!(c == nseg - 1 && s->fs_last)) {
      FF_HIP(hipMemcpyAsync(s->U, (char*)s->CK + c * sb, sb, hipMemcpyDeviceToDevice, st));
      ca.do_inverse = 1;
      FF_TRY(vort_col(prec, 0, g, ca, st));
      FF_TRY(fwd_steps(prec, s, ca, ra, k0, k1, false, true, fb, st));
    }
    s->fs_last = false;
    // reverse sweep over the segment
    for (int64_t k = k1 - 1; k >= k0; --k) {
      const char* Fk = (const char*)s->FS + (k - k0) * 4 * fb;
      for (int stage = 1; stage <= 4; ++stage) {
        ar.F = Fk + (4 - stage) * fb;  // fields of c, b, a, u
        FF_TRY((adispatch_row<R>(g, ar, st)));
        aa.cont = (stage < 4 || k > 0) ? 1 : 0;
        FF_TRY((adispatch_stage<R>(stage, g, aa, st)));
      }
    }
  }
  if (lam0 != lamT)
    FF_HIP(hipMemcpyAsync(lam0, lamT, (size_t)g.ny * g.nkr * sizeof(cplx<R>), hipMemcpyDeviceToDevice, st));
  FF_TRY(vort_export(prec, g, lam0, s->LAM, reinterpret_cast<const double*>(ts->coef[4]),
                     reinterpret_cast<const double*>(ts->filter), nsteps, st));
  return FFHIP_OK;
}

int fused_forward_ckpt(ffhip_op* op, ffhip_etdrk4* ts, const void* sol0, void* solT, int64_t nsteps,
                       int64_t ncheck, hipStream_t st) {
  FF_CHECK(op->fused, FFHIP_ERR_UNSUPPORTED, "the adjoint needs the fused path (power-of-two 2D vorticity grid)");
  return op->prec == FFHIP_F32 ? vforward_ckpt<float>(op, ts, sol0, solT, nsteps, ncheck, st)
                               : vforward_ckpt<double>(op, ts, sol0, solT, nsteps, ncheck, st);
}

int64_t fused_checkpoint_id(const ffhip_op* op) {
  const VortState* s = reinterpret_cast<const VortState*>(op->fused);
  return s ? s->ck_id : 0;
}

int fused_adjoint(ffhip_op* op, ffhip_etdrk4* ts, const void* lamT, void* lam0, int64_t nsteps, int64_t ncheck,
                  hipStream_t st) {
  FF_CHECK(op->fused, FFHIP_ERR_UNSUPPORTED, "the adjoint needs the fused path (power-of-two 2D vorticity grid)");
  return op->prec == FFHIP_F32 ? vadjoint<float>(op, ts, lamT, lam0, nsteps, ncheck, st)
                               : vadjoint<double>(op, ts, lamT, lam0, nsteps, ncheck, st);
}

}  // namespace ffhip
