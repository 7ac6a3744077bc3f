// Device bodies of the fused 2D vorticity passes and the layout-conversion
// kernels, shared by the single-grid kernels (vort2d_fused.hip) and the
// ensemble work-list kernels (vort2d_ensemble.hip).
#pragma once

#include "vort2d.h"

namespace ffhip {

// ETDRK4 stage combination at one kept mode (blocked index si) given this
// stage's N there (src/timesteppers.jl:501-516, in FP64 with FP64
// coefficients; 2 N3 - N1 and N2 + N3 formed in the state type as the
// reference's broadcasts do).  STAGE 0 only reads the state.
template <typename R, int STAGE>
__device__ __forceinline__ cplx<R> vstage_point(const VortColArgs& a, const int64_t si, const cplx<R> n) {
  cplx<R>* U = reinterpret_cast<cplx<R>*>(a.U);
  cplx<R>* N1 = reinterpret_cast<cplx<R>*>(a.N1);
  cplx<R>* N2 = reinterpret_cast<cplx<R>*>(a.N2);
  cplx<R>* As = reinterpret_cast<cplx<R>*>(a.A);
  if constexpr (STAGE == 0) {
    return U[si];
  } else if constexpr (STAGE == 1 || STAGE == 2) {
    // ETDRK4substep12!: exp(dt L/2) u + zeta N
    const cplx<R> u = U[si];
    const double eh = a.Eh[si], z = a.Z[si];
    const cplx<R> w = mkc<R>((R)(eh * (double)u.x + z * (double)n.x), (R)(eh * (double)u.y + z * (double)n.y));
    if constexpr (STAGE == 1) {
      N1[si] = n;
      As[si] = w;
    } else {
      N2[si] = n;
    }
    return w;
  } else if constexpr (STAGE == 3) {
    // ETDRK4substep3!: exp(dt L/2) a + zeta (2 N3 - N1); keep N2 + N3 for the update
    const cplx<R> av = As[si];
    const cplx<R> n1 = N1[si];
    const cplx<R> d = mkc<R>(R(2) * n.x - n1.x, R(2) * n.y - n1.y);
    const double eh = a.Eh[si], z = a.Z[si];
    const cplx<R> w = mkc<R>((R)(eh * (double)av.x + z * (double)d.x), (R)(eh * (double)av.y + z * (double)d.y));
    const cplx<R> n2 = N2[si];
    N2[si] = mkc<R>(n2.x + n.x, n2.y + n.y);
    return w;
  } else {
    // ETDRK4update!: exp(dt L) u + alpha N1 + 2 beta (N2 + N3) + Gamma N4
    const cplx<R> u = U[si];
    const cplx<R> n1 = N1[si];
    const cplx<R> s23 = N2[si];
    const double e = a.E[si], al = a.Al[si], be = 2.0 * a.Be[si], ga = a.Ga[si];
    double rx = e * (double)u.x + al * (double)n1.x;
    double ry = e * (double)u.y + al * (double)n1.y;
    rx = rx + be * (double)s23.x;
    ry = ry + be * (double)s23.y;
    rx = rx + ga * (double)n.x;
    ry = ry + ga * (double)n.y;
    cplx<R> w = mkc<R>((R)rx, (R)ry);
    if (a.Fl) {  // sol .*= filter (src/timesteppers.jl:549-558), a second rounding as there
      const double f = a.Fl[si];
      w = mkc<R>((R)(f * (double)w.x), (R)(f * (double)w.y));
    }
    U[si] = w;
    return w;
  }
}

// i q z for the three inverse fields of the column pass: f = 0: zeta^,
// f = 1: u^ = i l/|k|^2 zeta^, f = 2: v^ = -i kr/|k|^2 zeta^
template <typename R>
__device__ __forceinline__ cplx<R> vfield(int f, const cplx<R> z, const R kx, const R l, const bool origin) {
  if (f == 0) return z;
  const R K2 = kx * kx + l * l;
  const R invK = origin ? R(0) : R(1) / K2;
  if (f == 1) {
    const R q = l * invK;
    return mkc<R>(-q * z.y, q * z.x);
  }
  const R q = kx * invK;
  return mkc<R>(q * z.y, -q * z.x);
}

// One y-column (kx index tx*4 + c) of the column pass, run by a team of NY/P
// threads (team index tid) with its own LDS slice.  live = false marks a
// padding column beyond the last tile: it joins the team barriers but reads
// zeros and writes nothing.
// PAIR: independent transforms run side by side (team_fft2) in lds and lds2.
template <typename R, int NY, int P, int STAGE, bool PAIR = false>
__device__ __forceinline__ void vcol_body(const VortGeom& g, const VortColArgs& a, const int tx, const int c,
                                          const int tid, cplx<R>* lds, const cplx<R>* tw, const bool live,
                                          cplx<R>* lds2 = nullptr) {
  constexpr int T = NY / P;
  constexpr int CB = kVCB;
  const int ix = tx * CB + c;
  const bool colkept = live && ix < g.KX;
  const R kx = (R)(g.dk * (double)ix);
  const uint32_t ibase = (uint32_t)tx * NY * CB + c;  // 32-bit element offsets (at32)
  const uint32_t sbase = (uint32_t)tx * g.KY * CB + c;
  // the five transforms share one set of twiddles: FP32 keeps the first
  // pass's factors in registers (all passes would cost the second wavefront)
  constexpr bool treg = sizeof(R) == 4 ? FF_COL_TWREG : FF_COL_TWREG_F64;
  TwRegs<NY, P, R, 1> twr;
  if constexpr (treg) twr_fill(twr, tw, tid);
  auto fft = [&](cplx<R>(&x)[P], auto dir) {
    if constexpr (treg)
      team_fft<NY, P, decltype(dir)::value, R>(x, lds, tid, tw, twr);
    else
      team_fft<NY, P, decltype(dir)::value, R>(x, lds, tid, tw);
  };
  using Inv = std::integral_constant<int, 1>;
  using Fwd = std::integral_constant<int, -1>;

  cplx<R> nn[P];
  if constexpr (STAGE == 0) {
#pragma unroll
    for (int p = 0; p < P; ++p) nn[p] = mkc<R>(0, 0);
  }
  if constexpr (STAGE > 0) {
    cplx<R> vb[P];
    const cplx<R>* O0 = reinterpret_cast<const cplx<R>*>(a.O0);
    const cplx<R>* O1 = reinterpret_cast<const cplx<R>*>(a.O1);
#pragma unroll
    for (int p = 0; p < P; ++p) nn[p] = live ? at32(O0, ibase + (uint32_t)(tid + p * T) * CB) : mkc<R>(0, 0);
    if constexpr (sizeof(R) == 8 && FF_COL_O1PRE) {
      // FP64: the second input is loaded before the first transform, which
      // hides its load (measured at 2048^2: stages 1-4 about 1% faster; in
      // FP32 2% slower)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int p = 0; p < P; ++p) vb[p] = live ? at32(O1, ibase + (uint32_t)(tid + p * T) * CB) : mkc<R>(0, 0);
      if constexpr (!PAIR) fft(nn, Fwd{});
    } else {
      if constexpr (!PAIR) fft(nn, Fwd{});
#pragma unroll
      for (int p = 0; p < P; ++p) vb[p] = live ? at32(O1, ibase + (uint32_t)(tid + p * T) * CB) : mkc<R>(0, 0);
    }
    if constexpr (PAIR) {
      if constexpr (treg)
        team_fft2<NY, P, -1, R>(nn, vb, lds, lds2, tid, tw, twr);
      else
        team_fft2<NY, P, -1, R>(nn, vb, lds, lds2, tid, tw);
    } else {
      fft(vb, Fwd{});
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      const R l = (R)(g.dl * (double)vfreq(iy, NY));
      const cplx<R> A = nn[p], B = vb[p];
      nn[p] = mkc<R>(kx * A.y + l * B.y, -(kx * A.x) - l * B.x);  // -i kr A - i l B
    }
  }

  // stage combination on the kept band
  cplx<R> nw[P];
#pragma unroll FF_STAGE_UNROLL
  for (int p = 0; p < P; ++p) {
    const int iy = tid + p * T;
    const bool kept = colkept && (iy < g.ylo || iy > g.yhi);
    nw[p] = mkc<R>(0, 0);
    if (!kept) continue;
    const int64_t si = (int64_t)sbase + (int64_t)v_ky(iy, g) * CB;
    nw[p] = vstage_point<R, STAGE>(a, si, nn[p]);
  }
  if (!a.do_inverse) return;  // uniform over the block

  // inverse y-FFTs of zeta^, u^ = i l/|k|^2 zeta^, v^ = -i kr/|k|^2 zeta^
  cplx<R>* dst[3] = {reinterpret_cast<cplx<R>*>(a.I0), reinterpret_cast<cplx<R>*>(a.I1),
                     reinterpret_cast<cplx<R>*>(a.I2)};
  if constexpr (PAIR) {  // u^ and v^ side by side, then zeta^ (the last use of nw)
    cplx<R> v1[P], v2[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      const R l = (R)(g.dl * (double)vfreq(iy, NY));
      v1[p] = vfield<R>(1, nw[p], kx, l, ix == 0 && iy == 0);
      v2[p] = vfield<R>(2, nw[p], kx, l, ix == 0 && iy == 0);
    }
    if constexpr (treg)
      team_fft2<NY, P, 1, R>(v1, v2, lds, lds2, tid, tw, twr);
    else
      team_fft2<NY, P, 1, R>(v1, v2, lds, lds2, tid, tw);
    if (live) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        at32(dst[1], ibase + (uint32_t)(tid + p * T) * CB) = v1[p];
        at32(dst[2], ibase + (uint32_t)(tid + p * T) * CB) = v2[p];
      }
    }
    fft(nw, Inv{});
    if (live) {
#pragma unroll
      for (int p = 0; p < P; ++p) at32(dst[0], ibase + (uint32_t)(tid + p * T) * CB) = nw[p];
    }
    return;
  }
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    cplx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      v[p] = vfield<R>(f, nw[p], kx, (R)(g.dl * (double)vfreq(iy, NY)), ix == 0 && iy == 0);
    }
    fft(v, Inv{});
    cplx<R>* d = dst[f];
    if (live) {
#pragma unroll
      for (int p = 0; p < P; ++p) at32(d, ibase + (uint32_t)(tid + p * T) * CB) = v[p];
    }
  }
}

// The column pass of one y-column split over three teams of T = NY/P
// threads (T a multiple of 64, so every wavefront belongs to one team and
// the team branches below are wave-uniform).  Teams 0 and 1 run the two
// forward FFTs side by side while team 2 only keeps the barrier count; all
// threads then form N and the stage combination; each team then runs one of
// the three inverse FFTs.  lds holds three team regions of
// TeamSched<NY, P>::LDS slots.  The serial chain per column is two FFTs
// instead of five.
template <typename R, int NY, int P, int STAGE>
__device__ __forceinline__ void vcol3_body(const VortGeom& g, const VortColArgs& a, const int tx, const int c,
                                           const int team, const int tid, cplx<R>* lds, const cplx<R>* tw) {
  constexpr int T = NY / P;
  constexpr int CB = kVCB;
  constexpr int L = TeamSched<NY, P>::LDS;
  constexpr int NBAR = 2 * (TeamSched<NY, P>::NPASS - 1);  // barriers inside one team_fft
  static_assert(T % 64 == 0, "three-team column pass needs whole wavefronts per team");
  cplx<R>* mine = lds + team * L;
  cplx<R>* RA = lds;          // forward (u zeta)^ after the y-FFT, natural order
  cplx<R>* RB = lds + L;      // forward (v zeta)^
  cplx<R>* RW = lds + 2 * L;  // new stage state, natural order
  const int ix = tx * CB + c;
  const bool colkept = ix < g.KX;
  const R kx = (R)(g.dk * (double)ix);
  const uint32_t ibase = (uint32_t)tx * NY * CB + c;  // 32-bit offsets: uniform base + VGPR offset loads
  const uint32_t sbase = (uint32_t)tx * g.KY * CB + c;
  const int t = team * T + tid;  // thread index within the three teams

  if constexpr (STAGE > 0) {
    if (team < 2) {
      const cplx<R>* O = reinterpret_cast<const cplx<R>*>(team == 0 ? a.O0 : a.O1);
      cplx<R> v[P];
#pragma unroll
      for (int p = 0; p < P; ++p) v[p] = O[ibase + (uint32_t)(tid + p * T) * CB];
      team_fft<NY, P, -1, R>(v, mine, tid, tw);
#pragma unroll
      for (int p = 0; p < P; ++p) mine[lpad(tid + p * T)] = v[p];
    } else {
#pragma unroll
      for (int b = 0; b < NBAR; ++b) __syncthreads();
    }
    __syncthreads();
  }

  // N and the stage combination, all three teams over the y points
#pragma unroll 2
  for (int iy = t; iy < NY; iy += 3 * T) {
    const bool kept = colkept && (iy < g.ylo || iy > g.yhi);
    cplx<R> w = mkc<R>(0, 0);
    if (kept) {
      cplx<R> n = mkc<R>(0, 0);
      if constexpr (STAGE > 0) {
        const R l = (R)(g.dl * (double)vfreq(iy, NY));
        const cplx<R> A = RA[lpad(iy)], B = RB[lpad(iy)];
        n = mkc<R>(kx * A.y + l * B.y, -(kx * A.x) - l * B.x);  // -i kr A - i l B
      }
      w = vstage_point<R, STAGE>(a, sbase + (uint32_t)v_ky(iy, g) * CB, n);
    }
    RW[lpad(iy)] = w;
  }
  __syncthreads();
  if (!a.do_inverse) return;  // uniform over the block

  // team f: inverse y-FFT of field f (zeta^, u^, v^)
  cplx<R> v[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int iy = tid + p * T;
    v[p] = vfield<R>(team, RW[lpad(iy)], kx, (R)(g.dl * (double)vfreq(iy, NY)), ix == 0 && iy == 0);
  }
  __syncthreads();  // every team has read RW before team 2's transform reuses it
  team_fft<NY, P, 1, R>(v, mine, tid, tw);
  cplx<R>* d = reinterpret_cast<cplx<R>*>(team == 0 ? a.I0 : (team == 1 ? a.I1 : a.I2));
#pragma unroll
  for (int p = 0; p < P; ++p) d[ibase + (uint32_t)(tid + p * T) * CB] = v[p];
}

// One x-row y of the row pass, run by a team of NX/P threads (team index
// tid) with its own LDS slice.
// PAIR (FP32 only): the two inverse transforms run side by side through
// team_fft2 in lds and lds2, with every twiddle in registers filled from the
// roots of unity tw (global memory); otherwise tw is the LDS table and lds2
// is unused.
template <typename R, int NX, int P, bool PAIR = false>
__device__ __forceinline__ void vrow_body(const VortGeom& g, const VortRowArgs& a, const int y, const int tid,
                                          cplx<R>* lds, const cplx<R>* tw, cplx<R>* lds2 = nullptr) {
  static_assert(!PAIR || sizeof(R) == 4, "paired row transforms are FP32 only");
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  const int64_t ny = g.ny;
  // 32-bit element offsets from the array bases (the blocked arrays hold
  // fewer than 2^32 elements): saddr loads, no 64-bit address arithmetic
  auto iaddr = [&](int k) -> uint32_t { return ((uint32_t)(k / CB) * (uint32_t)ny + (uint32_t)y) * CB + (k % CB); };
  const cplx<R>* Zh = reinterpret_cast<const cplx<R>*>(a.I0);
  const cplx<R>* Uh = reinterpret_cast<const cplx<R>*>(a.I1);
  const cplx<R>* Vh = reinterpret_cast<const cplx<R>*>(a.I2);
  const int KX = g.KX;
  // the three transforms share one set of twiddles: FP32 keeps it in registers
  constexpr bool treg = PAIR || (sizeof(R) == 4 ? FF_ROW_TWREG : FF_ROW_TWREG_F64);
  TwRegs<NX, P, R> twr;
  if constexpr (PAIR)
    twr_fill_roots<NX, P, 1, R>(twr, tw, tid);
  else if constexpr (treg)
    twr_fill(twr, tw, tid);
  auto fft = [&](cplx<R>(&x)[P], auto dir) {
    if constexpr (treg)
      team_fft<NX, P, decltype(dir)::value, R>(x, lds, tid, tw, twr);
    else
      team_fft<NX, P, decltype(dir)::value, R>(x, lds, tid, tw);
  };
  using Inv = std::integral_constant<int, 1>;
  using Fwd = std::integral_constant<int, -1>;

  // u + i v: Hermitian-extended half spectra packed into one complex inverse.
  // Branch-free: every lane loads (the zero band reads column 0 and discards
  // it), so the compiler issues all loads before the first wait instead of
  // one round trip per point.
  cplx<R> z[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;  // direct / mirrored (conjugate) source
    const int kk = lo ? k : (hi ? NX - k : 0);
    const cplx<R> U = at32(Uh, iaddr(kk)), V = at32(Vh, iaddr(kk));
    const R uy = k == 0 ? R(0) : U.y, vy = k == 0 ? R(0) : V.y;  // c2r ignores Im at kx = 0
    const R sx = lo ? U.x - vy : (hi ? U.x + V.y : R(0));
    const R sy = lo ? uy + V.x : (hi ? V.x - U.y : R(0));  // conj(U) + i conj(V) when mirrored
    z[p] = mkc<R>(sx, sy);
  }
  // FP64: issue the zeta^ loads before the u + i v transform (measured faster);
  // FP32: transform first (with the twiddles in registers, measured 3% faster)
  constexpr bool prefetch = PAIR || (sizeof(R) == 8 ? FF_ROW_PREFETCH_F64 : FF_ROW_PREFETCH_F32);
  if constexpr (!prefetch) fft(z, Inv{});

  cplx<R> w[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;
    const int kk = lo ? k : (hi ? NX - k : 0);
    const cplx<R> Z = at32(Zh, iaddr(kk));
    const R sx = (lo || hi) ? Z.x : R(0);
    const R sy = lo ? (k == 0 ? R(0) : Z.y) : (hi ? -Z.y : R(0));
    w[p] = mkc<R>(sx, sy);
  }
  if constexpr (PAIR) {
    team_fft2<NX, P, 1, R>(z, w, lds, lds2, tid, tw, twr);
  } else {
    if constexpr (prefetch) fft(z, Inv{});
    fft(w, Inv{});
  }

  // products u zeta + i v zeta with the 1/(nx ny)^2 normalisation
  const R s2 = (R)g.s2;
  if (a.F) {  // keep u + i v and zeta (normalised) for the adjoint sweep
    const R s1 = (R)(1.0 / ((double)g.nx * (double)g.ny));
    cplx<R>* Fz = reinterpret_cast<cplx<R>*>(a.F) + (int64_t)y * NX;
    R* Fr = reinterpret_cast<R*>(reinterpret_cast<cplx<R>*>(a.F) + (int64_t)g.nx * g.ny) + (int64_t)y * NX;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      Fz[tid + p * T] = mkc<R>(z[p].x * s1, z[p].y * s1);
      Fr[tid + p * T] = w[p].x * s1;
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const R zz = w[p].x * s2;
    w[p] = mkc<R>(z[p].x * zz, z[p].y * zz);
  }
  fft(w, Fwd{});

  // split W into the half spectra of its real and imaginary parts
#pragma unroll
  for (int p = 0; p < P; ++p) lds[lpad(tid + p * T)] = w[p];
  __syncthreads();
  cplx<R>* O0 = reinterpret_cast<cplx<R>*>(a.O0);
  cplx<R>* O1 = reinterpret_cast<cplx<R>*>(a.O1);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    if (k >= g.KXp) continue;
    cplx<R> Pk = mkc<R>(0, 0), Qk = mkc<R>(0, 0);
    if (k < KX) {
      const cplx<R> wk = w[p];
      const cplx<R> wc = cconj<R>(lds[lpad((NX - k) & (NX - 1))]);
      Pk = mkc<R>(R(0.5) * (wk.x + wc.x), R(0.5) * (wk.y + wc.y));
      Qk = mkc<R>(R(0.5) * (wk.y - wc.y), -R(0.5) * (wk.x - wc.x));  // (wk - wc) / (2i)
    }
    at32(O0, iaddr(k)) = Pk;
    at32(O1, iaddr(k)) = Qk;
  }
}

// The row pass of one x-row y by one wavefront (FP32), with half-length
// transforms: every transform is an M = NX/2-point complex FFT of Q = M/64
// points per lane, exchanged through the wavefront's own LDS buffers with
// wavefront-level barriers only, so the rows of a workgroup never wait for
// each other.  Lane l holds the sample pairs x = 2n, 2n+1 for n = l + 64 q.
//   zeta:   the real inverse as one M-point transform of
//           Z[k] = X[k] + conj(X[M-k]) + i (X[k] - conj(X[M-k])) w^k,
//           giving zeta(2n) + i zeta(2n+1)              (w = exp(2 pi i/NX))
//   u + iv: decimation in frequency, A[k] = S[k] + S[k+M] and
//           B[k] = (S[k] - S[k+M]) w^k, two M-point inverses giving
//           (u + iv)(2n) and (u + iv)(2n+1)
//   forward: decimation in time on the products' even and odd samples, two
//           M-point transforms E, O and W[k] = E + w^-k O, W[k+M] = E - w^-k O
//   split:  P[k], Q[k] from W[k] and conj(W[NX-k]); W[NX-k] sits in lane
//           (64 - l) mod 64, fetched with ds_bpermute (__shfl)
// Five M-point transforms per row instead of three NX-point ones.  lds1 and
// lds2 are the row's two buffers of TeamSched<M, Q>::LDS slots.
// Team size T = M/Q: 64 (FP32) is one wavefront, with every twiddle in
// registers (filled from the NX roots of unity in tw) and wavefront-level
// exchanges; a larger team (128 or 256 threads, two or four wavefronts)
// takes its factors from the LDS table ltw, exchanges through workgroup
// barriers (all rows of the workgroup in step) and fetches the split's
// mirrored coefficient from LDS instead of another lane.  In the text above
// read "thread" for "lane" and T for 64 then.  Such a team may pass
// lds2 = null: its paired transforms then run one after the other in lds1.
template <typename R, int NX, int Q>
__device__ __forceinline__ void vrowh_body(const VortGeom& g, const VortRowArgs& a, const int y, const int lane,
                                           cplx<R>* lds1, cplx<R>* lds2, const cplx<R>* ltw = nullptr) {
  constexpr int M = NX / 2;
  constexpr int T = M / Q;
  constexpr bool WAVE = T == 64;
  static_assert(T % 64 == 0, "whole wavefronts per row");
  // zeta^, u^, v^ loaded up front in 16 B chunks (FP32 wavefront rows)
  constexpr bool PRE = WAVE && FF_ROWH_UVPRE;
  constexpr int CB = kVCB;
  const uint32_t ny = (uint32_t)g.ny;
  auto iaddr = [&](int k) -> uint32_t { return ((uint32_t)(k / CB) * ny + (uint32_t)y) * CB + (k % CB); };
  const cplx<R>* roots = reinterpret_cast<const cplx<R>*>(a.tw);  // exp(-2 pi i k/NX)
  const cplx<R>* Zh = reinterpret_cast<const cplx<R>*>(a.I0);
  const cplx<R>* Uh = reinterpret_cast<const cplx<R>*>(a.I1);
  const cplx<R>* Vh = reinterpret_cast<const cplx<R>*>(a.I2);
  const int KX = g.KX;
  TwRegs<M, Q, R> twr;  // wavefront rows only
  if constexpr (WAVE) twr_fill_roots_half<M, Q, 1, R, TwRegs<M, Q, R>::NR>(twr, roots, lane);
  auto fft1 = [&](cplx<R>(&x)[Q]) {
    if constexpr (WAVE)
      wave_fft<M, Q, 1, R>(x, lds1, lane, twr);
    else
      team_fft<M, Q, 1, R>(x, lds1, lane, ltw);
  };
  auto fft2 = [&](cplx<R>(&x1)[Q], cplx<R>(&x2)[Q], auto dir) {
    if constexpr (WAVE)
      wave_fft2<M, Q, decltype(dir)::value, R>(x1, x2, lds1, lds2, lane, twr);
    else if (lds2)
      team_fft2<M, Q, decltype(dir)::value, R>(x1, x2, lds1, lds2, lane, ltw);
    else {  // one buffer (lds2 null, uniform): the two transforms one after the other
      team_fft<M, Q, decltype(dir)::value, R>(x1, lds1, lane, ltw);
      team_fft<M, Q, decltype(dir)::value, R>(x2, lds1, lane, ltw);
    }
  };
  using Inv = std::integral_constant<int, 1>;
  using Fwd = std::integral_constant<int, -1>;

  // The row's zeta^, u^ and v^ (k < KXp) are loaded in 16 B chunks
  // c = lane + 64 i holding k = 2c, 2c + 1, each element once, zeta^ first;
  // u^ and v^ stay in registers across the zeta transform, which hides their
  // load behind it.  They reach the lanes that combine k with M - k through
  // LDS.
  constexpr int NCH = ((NX / 3 + 4) / 2 + 63) / 64;  // KXp <= NX/3 + 4
  using f4 = float __attribute__((ext_vector_type(4)));  // a native vector: stays in registers
  const int nch = g.KXp >> 1;
  f4 rz[NCH], ru[NCH], rv[NCH];
  cplx<R> omz[Q];  // the roots the zeta combination needs, loaded ahead of u^, v^ (loads retire in order)
  if constexpr (PRE) {
#pragma unroll
    for (int q = 0; q < Q; ++q) omz[q] = roots[lane + 64 * q];
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
      const int c = min(lane + 64 * i, nch - 1);  // the tail lanes repeat the last chunk
      rz[i] = at32(reinterpret_cast<const f4*>(Zh), iaddr(2 * c) >> 1);
    }
    __builtin_amdgcn_sched_barrier(0);  // issue order: zeta^ before u^, v^
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
      const int c = min(lane + 64 * i, nch - 1);
      ru[i] = at32(reinterpret_cast<const f4*>(Uh), iaddr(2 * c) >> 1);
      rv[i] = at32(reinterpret_cast<const f4*>(Vh), iaddr(2 * c) >> 1);
    }
#pragma unroll
    for (int i = 0; i < NCH; ++i) reinterpret_cast<f4*>(lds1)[min(lane + 64 * i, nch - 1)] = rz[i];
    team_sync<true>();
  }

  // zeta (2n) + i zeta (2n + 1).  Branch-free loads: out-of-band modes read
  // column 0 and are discarded; c2r ignores Im at kx = 0 and the Nyquist
  // column (k = M) is never kept.
  cplx<R> zr[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = lane + T * q, m = M - k;
    const bool okk = k < KX, okm = m < KX;
    cplx<R> Xk, Xm, om;
    if constexpr (PRE) {
      Xk = lds1[okk ? k : 0];
      Xm = lds1[okm ? m : 0];
      om = omz[q];
    } else {
      Xk = at32(Zh, iaddr(okk ? k : 0));
      Xm = at32(Zh, iaddr(okm ? m : 0));
      om = roots[k];
    }
    const R xkx = okk ? Xk.x : R(0), xky = (okk && k != 0) ? Xk.y : R(0);
    const R xmx = okm ? Xm.x : R(0), xmy = okm ? Xm.y : R(0);
    const R dx = xkx - xmx, dy = xky + xmy;                           // X[k] - conj(X[M-k])
    const R aox = dx * om.x + dy * om.y, aoy = dy * om.x - dx * om.y;  // times w^k = conj(root k)
    zr[q] = mkc<R>(xkx + xmx - aoy, xky - xmy + aox);                 // Ae + i Ao
  }
  if constexpr (PRE) team_sync<true>();  // every lane's reads of zeta^ before the transform overwrites them
  fft1(zr);

  // u + i v at the even (xa) and odd (xb) samples
  if constexpr (PRE) {
#pragma unroll
    for (int i = 0; i < NCH; ++i) {  // u^ into lds1, v^ into lds2 (free after the transform)
      const int c = min(lane + 64 * i, nch - 1);  // tail lanes rewrite the last chunk with its own value
      reinterpret_cast<f4*>(lds1)[c] = ru[i];
      reinterpret_cast<f4*>(lds2)[c] = rv[i];
    }
    team_sync<true>();
  }
  cplx<R> xa[Q], xb[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = lane + T * q, m = M - k;
    const bool okk = k < KX, okm = m < KX;
    cplx<R> U, V, Um, Vm;
    if constexpr (PRE) {
      const int ik = okk ? k : 0, im = okm ? m : 0;
      U = lds1[ik];
      V = lds2[ik];
      Um = lds1[im];
      Vm = lds2[im];
    } else {
      const uint32_t ik = iaddr(okk ? k : 0), im = iaddr(okm ? m : 0);
      U = at32(Uh, ik);
      V = at32(Vh, ik);
      Um = at32(Uh, im);
      Vm = at32(Vh, im);
    }
    const cplx<R> om = roots[k];
    const R uy = k == 0 ? R(0) : U.y, vy = k == 0 ? R(0) : V.y;
    const R slx = okk ? U.x - vy : R(0), sly = okk ? uy + V.x : R(0);       // S[k] = U + i V
    const R shx = okm ? Um.x + Vm.y : R(0), shy = okm ? Vm.x - Um.y : R(0);  // S[k+M] = conj(U) + i conj(V) at M-k
    xa[q] = mkc<R>(slx + shx, sly + shy);
    const R dx = slx - shx, dy = sly - shy;
    xb[q] = mkc<R>(dx * om.x + dy * om.y, dy * om.x - dx * om.y);
  }
  if constexpr (PRE) team_sync<true>();  // every lane's reads of u^, v^ before the transform overwrites them
  fft2(xa, xb, Inv{});

  // products u zeta + i v zeta with the 1/(nx ny)^2 normalisation
  const R s2 = (R)g.s2;
  if (a.F) {  // keep u + i v and zeta (normalised) for the adjoint sweep
    const R s1 = (R)(1.0 / ((double)g.nx * (double)g.ny));
    cplx<R>* Fz = reinterpret_cast<cplx<R>*>(a.F) + (int64_t)y * NX;
    R* Fr = reinterpret_cast<R*>(reinterpret_cast<cplx<R>*>(a.F) + (int64_t)g.nx * g.ny) + (int64_t)y * NX;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int n = lane + T * q;  // samples 2n and 2n + 1
      if constexpr (sizeof(R) == 4) {
        reinterpret_cast<float4*>(Fz)[n] = make_float4(xa[q].x * s1, xa[q].y * s1, xb[q].x * s1, xb[q].y * s1);
        reinterpret_cast<float2*>(Fr)[n] = make_float2(zr[q].x * s1, zr[q].y * s1);
      } else {
        Fz[2 * n] = mkc<R>(xa[q].x * s1, xa[q].y * s1);
        Fz[2 * n + 1] = mkc<R>(xb[q].x * s1, xb[q].y * s1);
        Fr[2 * n] = zr[q].x * s1;
        Fr[2 * n + 1] = zr[q].y * s1;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const R ze = zr[q].x * s2, zo = zr[q].y * s2;
    xa[q] = mkc<R>(xa[q].x * ze, xa[q].y * ze);
    xb[q] = mkc<R>(xb[q].x * zo, xb[q].y * zo);
  }
  fft2(xa, xb, Fwd{});
#pragma unroll
  for (int q = 0; q < Q; ++q) {  // W[k] in xa, W[k + M] in xb
    const cplx<R> t = cmul<R>(roots[lane + T * q], xb[q]);
    const cplx<R> e = xa[q];
    xa[q] = mkc<R>(e.x + t.x, e.y + t.y);
    xb[q] = mkc<R>(e.x - t.x, e.y - t.y);
  }

  // split W into the half spectra of its real and imaginary parts.  For
  // k = l + T q >= 1, NX - k = M + (M - k): W[NX - k] is element M - k of
  // xb.  One wavefront: M - k = (64 - l) + 64 (Q-1-q) (lane l >= 1) or
  // 64 (Q - q) (lane 0), fetched from that lane; larger teams: through LDS.
  cplx<R>* O0 = reinterpret_cast<cplx<R>*>(a.O0);
  cplx<R>* O1 = reinterpret_cast<cplx<R>*>(a.O1);
  if constexpr (!WAVE) {  // the transform ended with its LDS reads behind a barrier
#pragma unroll
    for (int q = 0; q < Q; ++q) lds1[lpad(lane + T * q)] = xb[q];
    __syncthreads();
  }
  const int partner = (64 - lane) & 63;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (T * q >= g.KXp) break;  // uniform over the workgroup: the shuffles below run in every lane
    const int k = lane + T * q;
    cplx<R> mir;
    if constexpr (WAVE) {
      mir.x = __shfl(xb[Q - 1 - q].x, partner, 64);
      mir.y = __shfl(xb[Q - 1 - q].y, partner, 64);
      if (lane == 0) mir = q == 0 ? xa[0] : xb[q == 0 ? 0 : Q - q];
    } else {
      mir = k == 0 ? xa[0] : lds1[lpad(M - k)];
    }
    if (k < g.KXp) {
      cplx<R> Pk = mkc<R>(0, 0), Qk = mkc<R>(0, 0);
      if (k < KX) {
        const cplx<R> wk = xa[q], wc = cconj<R>(mir);
        Pk = mkc<R>(R(0.5) * (wk.x + wc.x), R(0.5) * (wk.y + wc.y));
        Qk = mkc<R>(R(0.5) * (wk.y - wc.y), -R(0.5) * (wk.x - wc.x));  // (wk - wc) / (2i)
      }
      at32(O0, iaddr(k)) = Pk;
      at32(O1, iaddr(k)) = Qk;
    }
  }
}

// The row pass of two x-rows (y0, y0 + 1) by three teams of T = NX/P threads
// (T a multiple of 64: wave-uniform team branches).  Team 0 inverts u + i v
// of row y0, team 1 of row y0 + 1, team 2 zeta(y0) + i zeta(y0 + 1) -- three
// complex transforms for the five real inverse fields of the two rows.
// Teams 0 and 1 then form their row's products and run its forward
// transform while team 2 keeps the barrier count.  Five transforms per two
// rows instead of six, two in the serial chain instead of three.
template <typename R, int NX, int P>
__device__ __forceinline__ void vrow3_body(const VortGeom& g, const VortRowArgs& a, const int y0, const int team,
                                           const int tid, cplx<R>* lds, const cplx<R>* tw) {
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  constexpr int L = TeamSched<NX, P>::LDS;
  constexpr int NBAR = 2 * (TeamSched<NX, P>::NPASS - 1);  // barriers inside one team_fft
  static_assert(T % 64 == 0, "three-team row pass needs whole wavefronts per team");
  const int64_t ny = g.ny;
  const int KX = g.KX;
  cplx<R>* mine = lds + team * L;
  cplx<R>* RZ = lds + 2 * L;  // zeta(y0) + i zeta(y0 + 1), natural order
  const int y = y0 + (team & 1);
  auto iaddr = [&](int k, int yy) -> uint32_t { return ((uint32_t)(k / CB) * (uint32_t)ny + yy) * CB + (k % CB); };

  // team 0/1: u + i v of its row from (U, V) at row y; team 2: zeta(y0) + i zeta(y0 + 1)
  // from (Z(y0), Z(y0 + 1)).  Both are A + i B with A, B Hermitian-extended half spectra.
  const cplx<R>* SA = reinterpret_cast<const cplx<R>*>(team < 2 ? a.I1 : a.I0);
  const cplx<R>* SB = reinterpret_cast<const cplx<R>*>(team < 2 ? a.I2 : a.I0);
  const int ya = team < 2 ? y : y0, yb = team < 2 ? y : y0 + 1;
  cplx<R> v[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    const bool lo = k < KX, hi = k > NX - KX;
    const int kk = lo ? k : (hi ? NX - k : 0);  // branch-free: the zero band reads column 0
    const cplx<R> A = at32(SA, iaddr(kk, ya)), B = at32(SB, iaddr(kk, yb));
    // c2r ignores Im at kx = 0; mirrored modes are conjugated
    const R ay = k == 0 ? R(0) : (hi && !lo ? -A.y : A.y), by = k == 0 ? R(0) : (hi && !lo ? -B.y : B.y);
    const bool any = lo || hi;
    v[p] = mkc<R>(any ? A.x - by : R(0), any ? ay + B.x : R(0));
  }
  team_fft<NX, P, 1, R>(v, mine, tid, tw);
  if (team == 2) {
#pragma unroll
    for (int p = 0; p < P; ++p) RZ[lpad(tid + p * T)] = v[p];
  }
  __syncthreads();

  if (team < 2) {
    const R s2 = (R)g.s2;
    const R s1 = (R)(1.0 / ((double)g.nx * (double)g.ny));
    cplx<R>* Fz = a.F ? reinterpret_cast<cplx<R>*>(a.F) + (int64_t)y * NX : nullptr;
    R* Fr = a.F ? reinterpret_cast<R*>(reinterpret_cast<cplx<R>*>(a.F) + (int64_t)g.nx * g.ny) + (int64_t)y * NX
                : nullptr;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const cplx<R> zz = RZ[lpad(tid + p * T)];
      const R zr = team == 0 ? zz.x : zz.y;
      if (Fz) {  // keep u + i v and zeta (normalised) for the adjoint sweep
        Fz[tid + p * T] = mkc<R>(v[p].x * s1, v[p].y * s1);
        Fr[tid + p * T] = zr * s1;
      }
      const R zs = zr * s2;
      v[p] = mkc<R>(v[p].x * zs, v[p].y * zs);
    }
    team_fft<NX, P, -1, R>(v, mine, tid, tw);
#pragma unroll
    for (int p = 0; p < P; ++p) mine[lpad(tid + p * T)] = v[p];
  } else {
#pragma unroll
    for (int b = 0; b < NBAR; ++b) __syncthreads();
  }
  __syncthreads();

  if (team < 2) {
    // split W into the half spectra of its real and imaginary parts
    cplx<R>* O0 = reinterpret_cast<cplx<R>*>(a.O0);
    cplx<R>* O1 = reinterpret_cast<cplx<R>*>(a.O1);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int k = tid + p * T;
      if (k >= g.KXp) continue;
      cplx<R> Pk = mkc<R>(0, 0), Qk = mkc<R>(0, 0);
      if (k < KX) {
        const cplx<R> wk = v[p];
        const cplx<R> wc = cconj<R>(mine[lpad((NX - k) & (NX - 1))]);
        Pk = mkc<R>(R(0.5) * (wk.x + wc.x), R(0.5) * (wk.y + wc.y));
        Qk = mkc<R>(R(0.5) * (wk.y - wc.y), -R(0.5) * (wk.x - wc.x));  // (wk - wc) / (2i)
      }
      O0[iaddr(k, y)] = Pk;
      O1[iaddr(k, y)] = Qk;
    }
  }
}

// ----------------------------------------------------------------- layout conversion
template <typename R>
__global__ void k_vimport(const cplx<R>* __restrict__ sol, cplx<R>* __restrict__ U, VortGeom g) {
  const int64_t total = (int64_t)g.TX * g.KY * kVCB;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % kVCB);
    const int64_t rest = idx / kVCB;
    const int ky = (int)(rest % g.KY);
    const int tx = (int)(rest / g.KY);
    const int ix = tx * kVCB + c;
    U[idx] = ix < g.KX ? sol[(int64_t)v_iy(ky, g) * g.nkr + ix] : mkc<R>(0, 0);
  }
}

static __global__ void k_vpack(const double* __restrict__ src, double* __restrict__ dst, VortGeom g) {
  const int64_t total = (int64_t)g.TX * g.KY * kVCB;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % kVCB);
    const int64_t rest = idx / kVCB;
    const int ky = (int)(rest % g.KY);
    const int tx = (int)(rest / g.KY);
    const int ix = tx * kVCB + c;
    dst[idx] = ix < g.KX ? src[(int64_t)v_iy(ky, g) * g.nkr + ix] : 0.0;
  }
}

// Scatter the kept band back; modes outside it evolve as u <- exp(dt L) u per
// step (N is zero there), then u <- filter u when F is set, done here as
// nsteps roundings in the state type.
template <typename R>
__global__ void k_vexport(cplx<R>* __restrict__ sol, const cplx<R>* __restrict__ U, const double* __restrict__ E,
                          const double* __restrict__ F, int64_t nsteps, VortGeom g) {
  const int64_t total = (int64_t)g.ny * g.nkr;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(idx % g.nkr);
    const int iy = (int)(idx / g.nkr);
    const bool kept = ix < g.KX && (iy < g.ylo || iy > g.yhi);
    if (kept) {
      const int tx = ix / kVCB, c = ix % kVCB;
      sol[idx] = U[((int64_t)tx * g.KY + v_ky(iy, g)) * kVCB + c];
    } else {
      cplx<R> u = sol[idx];
      if (u.x != R(0) || u.y != R(0)) {
        const double e = E[idx];
        const double f = F ? F[idx] : 1.0;
        for (int64_t s = 0; s < nsteps; ++s) {
          u = mkc<R>((R)(e * (double)u.x), (R)(e * (double)u.y));
          if (F) u = mkc<R>((R)(f * (double)u.x), (R)(f * (double)u.y));
        }
        sol[idx] = u;
      }
    }
  }
}

// Algorithmic bytes of one launch: only the kept band (KX columns, KY kept
// rows of the state, all ny rows of the y-transformed intermediates) counts;
// the zero padding of the last column block does not.
template <typename R>
inline double col_bytes(const VortGeom& g, int stage, int do_inverse) {
  const double cs = sizeof(cplx<R>), ds = sizeof(double);
  const double ni = (double)g.KX * g.ny, nc = (double)g.KX * g.KY;
  double b = (do_inverse ? 3 : 0) * ni * cs;
  switch (stage) {
    case 0: return b + nc * cs;
    case 1: return b + 2 * ni * cs + nc * (3 * cs + 2 * ds);
    case 2: return b + 2 * ni * cs + nc * (2 * cs + 2 * ds);
    case 3: return b + 2 * ni * cs + nc * (4 * cs + 2 * ds);
    default: return b + 2 * ni * cs + nc * (4 * cs + 4 * ds);
  }
}

template <typename R>
inline double row_bytes(const VortGeom& g) {
  return 5.0 * g.KX * g.ny * sizeof(cplx<R>);
}

}  // namespace ffhip
