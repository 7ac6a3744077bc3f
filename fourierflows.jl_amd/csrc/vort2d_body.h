// Device bodies of the fused 2D vorticity passes and the layout-conversion
// kernels, shared by the single-grid kernels (vort2d_fused.hip) and the
// ensemble work-list kernels (vort2d_ensemble.hip).
#pragma once

#include "vort2d.h"

namespace ffhip {

// One y-column (kx index tx*4 + c) of the column pass, run by a team of NY/P
// threads (team index tid) with its own LDS slice.  live = false marks a
// padding column beyond the last tile: it joins the team barriers but reads
// zeros and writes nothing.
template <typename R, int NY, int P, int STAGE>
__device__ __forceinline__ void vcol_body(const VortGeom& g, const VortColArgs& a, const int tx, const int c,
                                          const int tid, cplx<R>* lds, const bool live) {
  constexpr int T = NY / P;
  constexpr int CB = kVCB;
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(a.tw);
  const int ix = tx * CB + c;
  const bool colkept = live && ix < g.KX;
  const R kx = (R)(g.dk * (double)ix);
  const int64_t ibase = (int64_t)tx * NY * CB + c;
  const int64_t sbase = (int64_t)tx * g.KY * CB + c;

  cplx<R> nn[P];
  if constexpr (STAGE > 0) {
    cplx<R> vb[P];
    const cplx<R>* O0 = reinterpret_cast<const cplx<R>*>(a.O0);
    const cplx<R>* O1 = reinterpret_cast<const cplx<R>*>(a.O1);
#pragma unroll
    for (int p = 0; p < P; ++p) nn[p] = live ? O0[ibase + (int64_t)(tid + p * T) * CB] : mkc<R>(0, 0);
    team_fft<NY, P, -1, R>(nn, lds, tid, tw);
#pragma unroll
    for (int p = 0; p < P; ++p) vb[p] = live ? O1[ibase + (int64_t)(tid + p * T) * CB] : mkc<R>(0, 0);
    team_fft<NY, P, -1, R>(vb, lds, tid, tw);
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
  cplx<R>* U = reinterpret_cast<cplx<R>*>(a.U);
  cplx<R>* N1 = reinterpret_cast<cplx<R>*>(a.N1);
  cplx<R>* N2 = reinterpret_cast<cplx<R>*>(a.N2);
  cplx<R>* As = reinterpret_cast<cplx<R>*>(a.A);
#pragma unroll FF_STAGE_UNROLL
  for (int p = 0; p < P; ++p) {
    const int iy = tid + p * T;
    const bool kept = colkept && (iy < g.ylo || iy > g.yhi);
    nw[p] = mkc<R>(0, 0);
    if (!kept) continue;
    const int64_t si = sbase + (int64_t)v_ky(iy, g) * CB;
    if constexpr (STAGE == 0) {
      nw[p] = U[si];
    } else if constexpr (STAGE == 1 || STAGE == 2) {
      // ETDRK4substep12!: exp(dt L/2) u + zeta N
      const cplx<R> u = U[si];
      const double eh = a.Eh[si], z = a.Z[si];
      const cplx<R> n = nn[p];
      nw[p] = mkc<R>((R)(eh * (double)u.x + z * (double)n.x), (R)(eh * (double)u.y + z * (double)n.y));
      if constexpr (STAGE == 1) {
        N1[si] = n;
        As[si] = nw[p];
      } else {
        N2[si] = n;
      }
    } else if constexpr (STAGE == 3) {
      // ETDRK4substep3!: exp(dt L/2) a + zeta (2 N3 - N1); keep N2 + N3 for the update
      const cplx<R> av = As[si];
      const cplx<R> n1 = N1[si];
      const cplx<R> n3 = nn[p];
      const cplx<R> d = mkc<R>(R(2) * n3.x - n1.x, R(2) * n3.y - n1.y);
      const double eh = a.Eh[si], z = a.Z[si];
      nw[p] = mkc<R>((R)(eh * (double)av.x + z * (double)d.x), (R)(eh * (double)av.y + z * (double)d.y));
      const cplx<R> n2 = N2[si];
      N2[si] = mkc<R>(n2.x + n3.x, n2.y + n3.y);
    } else {
      // ETDRK4update!: exp(dt L) u + alpha N1 + 2 beta (N2 + N3) + Gamma N4
      const cplx<R> u = U[si];
      const cplx<R> n1 = N1[si];
      const cplx<R> s23 = N2[si];
      const cplx<R> n4 = nn[p];
      const double e = a.E[si], al = a.Al[si], be = 2.0 * a.Be[si], ga = a.Ga[si];
      double rx = e * (double)u.x + al * (double)n1.x;
      double ry = e * (double)u.y + al * (double)n1.y;
      rx = rx + be * (double)s23.x;
      ry = ry + be * (double)s23.y;
      rx = rx + ga * (double)n4.x;
      ry = ry + ga * (double)n4.y;
      nw[p] = mkc<R>((R)rx, (R)ry);
      U[si] = nw[p];
    }
  }
  if (!a.do_inverse) return;  // uniform over the block

  // inverse y-FFTs of zeta^, u^ = i l/|k|^2 zeta^, v^ = -i kr/|k|^2 zeta^
  cplx<R>* dst[3] = {reinterpret_cast<cplx<R>*>(a.I0), reinterpret_cast<cplx<R>*>(a.I1),
                     reinterpret_cast<cplx<R>*>(a.I2)};
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    cplx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int iy = tid + p * T;
      if (f == 0) {
        v[p] = nw[p];
      } else {
        const R l = (R)(g.dl * (double)vfreq(iy, NY));
        const R K2 = kx * kx + l * l;
        const R invK = (ix == 0 && iy == 0) ? R(0) : R(1) / K2;
        const cplx<R> z = nw[p];
        if (f == 1) {
          const R q = l * invK;
          v[p] = mkc<R>(-q * z.y, q * z.x);
        } else {
          const R q = kx * invK;
          v[p] = mkc<R>(q * z.y, -q * z.x);
        }
      }
    }
    team_fft<NY, P, 1, R>(v, lds, tid, tw);
    cplx<R>* d = dst[f];
    if (live) {
#pragma unroll
      for (int p = 0; p < P; ++p) d[ibase + (int64_t)(tid + p * T) * CB] = v[p];
    }
  }
}

// One x-row y of the row pass, run by a team of NX/P threads (team index
// tid) with its own LDS slice.
template <typename R, int NX, int P>
__device__ __forceinline__ void vrow_body(const VortGeom& g, const VortRowArgs& a, const int y, const int tid,
                                          cplx<R>* lds) {
  constexpr int T = NX / P;
  constexpr int CB = kVCB;
  const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(a.tw);
  const int64_t ny = g.ny;
  auto iaddr = [&](int k) -> int64_t { return ((int64_t)(k / CB) * ny + y) * CB + (k % CB); };
  const cplx<R>* Zh = reinterpret_cast<const cplx<R>*>(a.I0);
  const cplx<R>* Uh = reinterpret_cast<const cplx<R>*>(a.I1);
  const cplx<R>* Vh = reinterpret_cast<const cplx<R>*>(a.I2);
  const int KX = g.KX;

  // u + i v: Hermitian-extended half spectra packed into one complex inverse
  cplx<R> z[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    cplx<R> s = mkc<R>(0, 0);
    if (k == 0) {
      const cplx<R> U = Uh[iaddr(0)], V = Vh[iaddr(0)];
      s = mkc<R>(U.x, V.x);  // c2r ignores Im at kx = 0
    } else if (k < KX) {
      const cplx<R> U = Uh[iaddr(k)], V = Vh[iaddr(k)];
      s = mkc<R>(U.x - V.y, U.y + V.x);
    } else if (k > NX - KX) {
      const int k2 = NX - k;
      const cplx<R> U = Uh[iaddr(k2)], V = Vh[iaddr(k2)];
      s = mkc<R>(U.x + V.y, V.x - U.y);  // conj(U) + i conj(V)
    }
    z[p] = s;
  }
  // FP64: issue the zeta^ loads before the u + i v transform (measured faster);
  // FP32: transform first (the extra live registers cost more than they hide)
  constexpr bool prefetch = sizeof(R) == 8 ? FF_ROW_PREFETCH_F64 : FF_ROW_PREFETCH_F32;
  if constexpr (!prefetch) team_fft<NX, P, 1, R>(z, lds, tid, tw);

  cplx<R> w[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = tid + p * T;
    cplx<R> s = mkc<R>(0, 0);
    if (k == 0) {
      s = mkc<R>(Zh[iaddr(0)].x, 0);
    } else if (k < KX) {
      s = Zh[iaddr(k)];
    } else if (k > NX - KX) {
      s = cconj<R>(Zh[iaddr(NX - k)]);
    }
    w[p] = s;
  }
  if constexpr (prefetch) team_fft<NX, P, 1, R>(z, lds, tid, tw);
  team_fft<NX, P, 1, R>(w, lds, tid, tw);

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
  team_fft<NX, P, -1, R>(w, lds, tid, tw);

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
    O0[iaddr(k)] = Pk;
    O1[iaddr(k)] = Qk;
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
// step (N is zero there), done here as nsteps roundings in the state type.
template <typename R>
__global__ void k_vexport(cplx<R>* __restrict__ sol, const cplx<R>* __restrict__ U, const double* __restrict__ E,
                          int64_t nsteps, VortGeom g) {
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
        for (int64_t s = 0; s < nsteps; ++s) u = mkc<R>((R)(e * (double)u.x), (R)(e * (double)u.y));
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
