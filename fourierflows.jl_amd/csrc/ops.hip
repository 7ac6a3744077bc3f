// Built-in nonlinear operators (the calcN! bodies the hot path needs) and
// their time-step drivers.
//
// Generic path: the reference's dataflow (spectral multiplies, ldiv! through
// the plan, physical products, mul!, assembly with dealiasing), one kernel per
// stage, any grid the plan supports.  The 2D vorticity operator additionally
// has a fused path (vort2d_fused.hip) used for power-of-two grids.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ops.h"
#include "reduce.h"

namespace ffhip {

constexpr int kOpThreads = 256;
static unsigned op_blocks(int64_t n) { return blocks_for(n, kOpThreads, 2048 * 8); }

struct Kept {
  int64_t lo0, hi0, lo1, hi1, lo2, hi2;  // 0-based inclusive aliased ranges (empty: lo > hi)
  __device__ __forceinline__ bool operator()(int64_t i0, int64_t i1, int64_t i2) const {
    return !((i0 >= lo0 && i0 <= hi0) || (i1 >= lo1 && i1 <= hi1) || (i2 >= lo2 && i2 <= hi2));
  }
};

static Kept kept_of(const ffhip_op* op) {
  Kept k;
  k.lo0 = op->lo[0] - 1;
  k.hi0 = op->hi[0] - 1;
  k.lo1 = op->lo[1] - 1;
  k.hi1 = op->hi[1] - 1;
  k.lo2 = op->lo[2] - 1;
  k.hi2 = op->hi[2] - 1;
  return k;
}

// signed wavenumber index of fftfreq ordering
__device__ __forceinline__ int64_t freq_index(int64_t i, int64_t n) { return i < n / 2 ? i : i - n; }

// ----------------------------------------------------------------- 2D vorticity (generic)
template <typename R>
__global__ void k_vort_prep(const cplx<R>* __restrict__ sol, cplx<R>* __restrict__ uh, cplx<R>* __restrict__ vh,
                            cplx<R>* __restrict__ zh, int64_t nkr, int64_t ny, double dk, double dl, Kept kept) {
  const int64_t total = nkr * ny;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr, iy = idx / nkr;
    const cplx<R> z = kept(ix, iy, 0) ? sol[idx] : mkc<R>(0, 0);
    const R kr = (R)(dk * (double)ix);
    const R l = (R)(dl * (double)freq_index(iy, ny));
    const R K2 = kr * kr + l * l;
    const R invK = (ix == 0 && iy == 0) ? R(0) : R(1) / K2;
    const R a = l * invK, b = kr * invK;
    uh[idx] = mkc<R>(-a * z.y, a * z.x);   // i l invK zeta
    vh[idx] = mkc<R>(b * z.y, -b * z.x);   // -i kr invK zeta
    zh[idx] = z;
  }
}

template <typename R>
__global__ void k_mul2(R* __restrict__ u, R* __restrict__ v, const R* __restrict__ z, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const R zz = z[i];
    u[i] *= zz;
    v[i] *= zz;
  }
}

template <typename R>
__global__ void k_vort_assemble(cplx<R>* __restrict__ N, const cplx<R>* __restrict__ A,
                                const cplx<R>* __restrict__ B, int64_t nkr, int64_t ny, double dk, double dl,
                                Kept kept) {
  const int64_t total = nkr * ny;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr, iy = idx / nkr;
    if (!kept(ix, iy, 0)) {
      N[idx] = mkc<R>(0, 0);
      continue;
    }
    const R kr = (R)(dk * (double)ix);
    const R l = (R)(dl * (double)freq_index(iy, ny));
    const cplx<R> a = A[idx], b = B[idx];
    // -i kr A - i l B
    N[idx] = mkc<R>(kr * a.y + l * b.y, -(kr * a.x) - l * b.x);
  }
}

// ----------------------------------------------------------------- 3D tracer (generic)
template <typename R>
__global__ void k_mask_copy(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int64_t d0, int64_t d1,
                            int64_t total, Kept kept) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = idx % d0;
    const int64_t r = idx / d0;
    const int64_t i1 = r % d1, i2 = r / d1;
    out[idx] = kept(i0, i1, i2) ? in[idx] : mkc<R>(0, 0);
  }
}

// tables: sinx, cosx (nx) | siny, cosy (ny) | sinz, cosz (nz), FP64
template <typename R>
__global__ void k_tracer_products(const R* __restrict__ c, R* __restrict__ uc, R* __restrict__ vc,
                                  R* __restrict__ wc, int64_t nx, int64_t ny, int64_t nz,
                                  const double* __restrict__ tab, double A, double B, double C) {
  const int64_t total = nx * ny * nz;
  const double* sx = tab;
  const double* cx = tab + nx;
  const double* sy = tab + 2 * nx;
  const double* cy = sy + ny;
  const double* sz = tab + 2 * nx + 2 * ny;
  const double* cz = sz + nz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nx;
    const int64_t r = idx / nx;
    const int64_t iy = r % ny, iz = r / ny;
    const R u = (R)(A * sz[iz] + C * cy[iy]);
    const R v = (R)(B * sx[ix] + A * cz[iz]);
    const R w = (R)(C * sy[iy] + B * cx[ix]);
    const R cc = c[idx];
    uc[idx] = u * cc;
    vc[idx] = v * cc;
    wc[idx] = w * cc;
  }
}

template <typename R>
__global__ void k_tracer_assemble(cplx<R>* __restrict__ N, const cplx<R>* __restrict__ U,
                                  const cplx<R>* __restrict__ V, const cplx<R>* __restrict__ W, int64_t nkr,
                                  int64_t ny, int64_t nz, double dk, double dl, double dm, Kept kept) {
  const int64_t total = nkr * ny * nz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr;
    const int64_t r = idx / nkr;
    const int64_t iy = r % ny, iz = r / ny;
    if (!kept(ix, iy, iz)) {
      N[idx] = mkc<R>(0, 0);
      continue;
    }
    const R kr = (R)(dk * (double)ix);
    const R l = (R)(dl * (double)freq_index(iy, ny));
    const R m = (R)(dm * (double)freq_index(iz, nz));
    const cplx<R> u = U[idx], v = V[idx], w = W[idx];
    const R sx = kr * u.x + l * v.x + m * w.x;
    const R sy = kr * u.y + l * v.y + m * w.y;
    N[idx] = mkc<R>(sy, -sx);  // -i (kr U + l V + m W)
  }
}

// ----------------------------------------------------------------- tracer adjoint
// The tracer N is linear in c (steady velocity), so the adjoint of an RK4
// step needs no forward states.  For the real inner product on the stored
// rfft entries (W = 1 at kx = 0, nx/2, else 2; Nt = nx ny nz):
//   N^T(x) = P ldiv!^T( u mul!^T(i kr P x) + v mul!^T(i l P x) + w mul!^T(i m P x) )
//   mul!^T(Y) = Nt ldiv!(Y / W),  ldiv!^T(phi) = W mul!(phi) / Nt.
// The 1/W is folded into the prep pass, W/Nt into the finish pass.
template <typename R>
__global__ void k_tracer_adj_prep(const cplx<R>* __restrict__ x, cplx<R>* __restrict__ s0, cplx<R>* __restrict__ s1,
                                  cplx<R>* __restrict__ s2, int64_t nkr, int64_t ny, int64_t nz, double dk, double dl,
                                  double dm, Kept kept) {
  const int64_t total = nkr * ny * nz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr;
    const int64_t r = idx / nkr;
    const int64_t iy = r % ny, iz = r / ny;
    if (!kept(ix, iy, iz)) {
      s0[idx] = s1[idx] = s2[idx] = mkc<R>(0, 0);
      continue;
    }
    const double iw = (ix == 0 || ix == nkr - 1) ? 1.0 : 0.5;
    const R kr = (R)(dk * (double)ix * iw);
    const R l = (R)(dl * (double)freq_index(iy, ny) * iw);
    const R m = (R)(dm * (double)freq_index(iz, nz) * iw);
    const cplx<R> v = x[idx];
    s0[idx] = mkc<R>(-kr * v.y, kr * v.x);  // i kr x / W
    s1[idx] = mkc<R>(-l * v.y, l * v.x);
    s2[idx] = mkc<R>(-m * v.y, m * v.x);
  }
}

template <typename R>
__global__ void k_tracer_adj_products(const R* __restrict__ p0, const R* __restrict__ p1, const R* __restrict__ p2,
                                      R* __restrict__ c, int64_t nx, int64_t ny, int64_t nz,
                                      const double* __restrict__ tab, double A, double B, double C) {
  const int64_t total = nx * ny * nz;
  const double* sx = tab;
  const double* cx = tab + nx;
  const double* sy = tab + 2 * nx;
  const double* cy = sy + ny;
  const double* sz = tab + 2 * nx + 2 * ny;
  const double* cz = sz + nz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nx;
    const int64_t r = idx / nx;
    const int64_t iy = r % ny, iz = r / ny;
    const R u = (R)(A * sz[iz] + C * cy[iy]);
    const R v = (R)(B * sx[ix] + A * cz[iz]);
    const R w = (R)(C * sy[iy] + B * cx[ix]);
    c[idx] = u * p0[idx] + v * p1[idx] + w * p2[idx];
  }
}

// out = P (W / Nt) S + L x   (L real, FP64)
template <typename R>
__global__ void k_tracer_adj_finish(cplx<R>* __restrict__ out, const cplx<R>* __restrict__ S,
                                    const cplx<R>* __restrict__ x, const double* __restrict__ L, int64_t nkr,
                                    int64_t ny, int64_t nz, double inv_nt, Kept kept) {
  const int64_t total = nkr * ny * nz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr;
    const int64_t r = idx / nkr;
    const int64_t iy = r % ny, iz = r / ny;
    const cplx<R> v = x[idx];
    const double l = L[idx];
    double ox = l * (double)v.x, oy = l * (double)v.y;
    if (kept(ix, iy, iz)) {
      const double w = ((ix == 0 || ix == nkr - 1) ? 1.0 : 2.0) * inv_nt;
      const cplx<R> s = S[idx];
      ox += w * (double)s.x;
      oy += w * (double)s.y;
    }
    out[idx] = mkc<R>((R)ox, (R)oy);
  }
}

// mode 0: kb = cl lam | 1: ub = lam + o, kb = cl lam + cu o | 2: ub += o, kb = cl lam + cu o | 3: lam = ub + o
template <typename R>
__global__ void k_rk4_adj_combo(cplx<R>* __restrict__ ub, cplx<R>* __restrict__ kb, cplx<R>* __restrict__ lam,
                                const cplx<R>* __restrict__ o, int mode, double cl, double cu, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cplx<R> l = lam[i];
    if (mode == 0) {
      kb[i] = mkc<R>((R)(cl * (double)l.x), (R)(cl * (double)l.y));
      continue;
    }
    const cplx<R> v = o[i];
    if (mode == 3) {
      const cplx<R> u = ub[i];
      lam[i] = mkc<R>(u.x + v.x, u.y + v.y);
      continue;
    }
    const cplx<R> base = mode == 1 ? l : ub[i];
    ub[i] = mkc<R>(base.x + v.x, base.y + v.y);
    kb[i] = mkc<R>((R)(cl * (double)l.x + cu * (double)v.x), (R)(cl * (double)l.y + cu * (double)v.y));
  }
}

// ----------------------------------------------------------------- diagnostics
// out: [sum w |u|^2, sum w |u|^2 invK^2] for vorticity; [sum w |u|^2, Re u(0)] for tracer
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_diag2(const cplx<R>* __restrict__ u, int64_t nkr, int64_t ny,
                                                       int64_t nz, double dk, double dl, double dm,
                                                       double2* __restrict__ partial) {
  const int64_t total = nkr * ny * nz;
  double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = idx % nkr;
    const int64_t r = idx / nkr;
    const int64_t iy = r % ny, iz = r / ny;
    const cplx<R> v = u[idx];
    double2 a2 = dd_add(dd_sq((double)v.x), dd_sq((double)v.y));
    const double w = (ix == 0 || ix == nkr - 1) ? 1.0 : 2.0;
    a2 = make_double2(w * a2.x, w * a2.y);
    const double kr = dk * (double)ix, l = dl * (double)freq_index(iy, ny), m = dm * (double)freq_index(iz, nz);
    const double K2 = kr * kr + l * l + m * m;
    const double inv = K2 == 0 ? 0.0 : 1.0 / K2;
    s0 = dd_add(s0, a2);
    s1 = dd_add(s1, make_double2(a2.x * inv, a2.y * inv));
  }
  s0 = block_dd_sum(s0);
  __syncthreads();
  s1 = block_dd_sum(s1);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_diag2_final(const double2* __restrict__ partial, int n,
                                                             double2* __restrict__ out) {
  double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s0 = dd_add(s0, partial[2 * i]);
    s1 = dd_add(s1, partial[2 * i + 1]);
  }
  s0 = block_dd_sum(s0);
  __syncthreads();
  s1 = block_dd_sum(s1);
  if (threadIdx.x == 0) {
    out[0] = s0;
    out[1] = s1;
  }
}

// ----------------------------------------------------------------- generic calcN
template <typename R>
static int calcN_generic(ffhip_op* op, void* N, const void* sol, hipStream_t st) {
  const Kept kept = kept_of(op);
  cplx<R>* s0 = reinterpret_cast<cplx<R>*>(op->spec[0]);
  cplx<R>* s1 = reinterpret_cast<cplx<R>*>(op->spec[1]);
  cplx<R>* s2 = reinterpret_cast<cplx<R>*>(op->spec[2]);
  R* p0 = reinterpret_cast<R*>(op->phys[0]);
  R* p1 = reinterpret_cast<R*>(op->phys[1]);
  R* p2 = reinterpret_cast<R*>(op->phys[2]);
  const double twopi = 6.283185307179586476925286766559;
  if (op->kind == FFHIP_OP_VORTICITY2D) {
    const int64_t nkr = op->nkr, ny = op->n[1];
    const double dk = twopi / op->box[0], dl = twopi / op->box[1];
    hipLaunchKernelGGL((k_vort_prep<R>), dim3(op_blocks(nkr * ny)), dim3(kOpThreads), 0, st,
                       (const cplx<R>*)sol, s0, s1, s2, nkr, ny, dk, dl, kept);
    FF_LAUNCHED("k_vort_prep");
    FF_TRY(ffhip_plan_inverse(op->plan, s0, p0, st));  // u
    FF_TRY(ffhip_plan_inverse(op->plan, s1, p1, st));  // v
    FF_TRY(ffhip_plan_inverse(op->plan, s2, p2, st));  // zeta
    hipLaunchKernelGGL((k_mul2<R>), dim3(op_blocks(op->phys_count)), dim3(kOpThreads), 0, st, p0, p1,
                       (const R*)p2, op->phys_count);
    FF_LAUNCHED("k_mul2");
    FF_TRY(ffhip_plan_forward(op->plan, p0, s0, st));
    FF_TRY(ffhip_plan_forward(op->plan, p1, s1, st));
    hipLaunchKernelGGL((k_vort_assemble<R>), dim3(op_blocks(nkr * ny)), dim3(kOpThreads), 0, st, (cplx<R>*)N,
                       (const cplx<R>*)s0, (const cplx<R>*)s1, nkr, ny, dk, dl, kept);
    FF_LAUNCHED("k_vort_assemble");
    return FFHIP_OK;
  }
  // tracer
  const int64_t nkr = op->nkr, nx = op->n[0], ny = op->n[1], nz = op->n[2];
  hipLaunchKernelGGL((k_mask_copy<R>), dim3(op_blocks(op->spec_count)), dim3(kOpThreads), 0, st,
                     (const cplx<R>*)sol, s0, nkr, ny, op->spec_count, kept);
  FF_LAUNCHED("k_mask_copy");
  R* c = reinterpret_cast<R*>(op->phys[3]);
  FF_TRY(ffhip_plan_inverse(op->plan, s0, c, st));
  hipLaunchKernelGGL((k_tracer_products<R>), dim3(op_blocks(op->phys_count)), dim3(kOpThreads), 0, st,
                     (const R*)c, p0, p1, p2, nx, ny, nz, (const double*)op->tables, op->params[0], op->params[1],
                     op->params[2]);
  FF_LAUNCHED("k_tracer_products");
  FF_TRY(ffhip_plan_forward(op->plan, p0, s0, st));
  FF_TRY(ffhip_plan_forward(op->plan, p1, s1, st));
  FF_TRY(ffhip_plan_forward(op->plan, p2, s2, st));
  hipLaunchKernelGGL((k_tracer_assemble<R>), dim3(op_blocks(op->spec_count)), dim3(kOpThreads), 0, st,
                     (cplx<R>*)N, (const cplx<R>*)s0, (const cplx<R>*)s1, (const cplx<R>*)s2, nkr, ny, nz,
                     twopi / op->box[0], twopi / op->box[1], twopi / op->box[2], kept);
  FF_LAUNCHED("k_tracer_assemble");
  return FFHIP_OK;
}

int op_calcN(ffhip_op* op, void* N, const void* sol, hipStream_t st) {
  return op->prec == FFHIP_F32 ? calcN_generic<float>(op, N, sol, st) : calcN_generic<double>(op, N, sol, st);
}

static int etd_run_generic(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, hipStream_t st) {
  FF_TRY(etd_scratch(ts));
  void* s1 = ts->scratch[0];
  void* s2 = ts->scratch[1];
  void* N1 = ts->scratch[2];
  void* N2 = ts->scratch[3];
  void* N3 = ts->scratch[4];
  void* N4 = ts->scratch[5];
  for (int64_t k = 0; k < nsteps; ++k) {
    FF_TRY(op_calcN(op, N1, sol, st));
    FF_TRY(etd_substep12(ts, s1, sol, N1, st));
    FF_TRY(op_calcN(op, N2, s1, st));
    FF_TRY(etd_substep12(ts, s2, sol, N2, st));
    FF_TRY(op_calcN(op, N3, s2, st));
    FF_TRY(etd_substep3(ts, s2, s1, N1, N3, st));
    FF_TRY(op_calcN(op, N4, s2, st));
    FF_TRY(etd_update(ts, sol, N1, N2, N3, N4, st));
    FF_TRY(apply_filter(ts->prec, ts->count, ts->filter, sol, st));
  }
  return FFHIP_OK;
}

static int rk4_run_generic(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, hipStream_t st) {
  FF_TRY(rk4_scratch(ts));
  void* s1 = ts->scratch[0];
  void* R1 = ts->scratch[1];
  void* R2 = ts->scratch[2];
  void* R3 = ts->scratch[3];
  void* R4 = ts->scratch[4];
  const double h2 = op->prec == FFHIP_F32 ? (double)((float)dt / 2.0f) : dt / 2;
  for (int64_t k = 0; k < nsteps; ++k) {
    FF_TRY(op_calcN(op, R1, sol, st));
    FF_TRY(rk4_addlinear(ts, R1, sol, st));
    FF_TRY(rk4_substepsol(ts, s1, sol, R1, h2, st));
    FF_TRY(op_calcN(op, R2, s1, st));
    FF_TRY(rk4_addlinear(ts, R2, s1, st));
    FF_TRY(rk4_substepsol(ts, s1, sol, R2, h2, st));
    FF_TRY(op_calcN(op, R3, s1, st));
    FF_TRY(rk4_addlinear(ts, R3, s1, st));
    FF_TRY(rk4_substepsol(ts, s1, sol, R3, dt, st));
    FF_TRY(op_calcN(op, R4, s1, st));
    FF_TRY(rk4_addlinear(ts, R4, s1, st));
    FF_TRY(rk4_update(ts, sol, R1, R2, R3, R4, dt, st));
    FF_TRY(apply_filter(ts->prec, ts->count, ts->filter, sol, st));
  }
  return FFHIP_OK;
}

// LSRK54 coefficients (src/timesteppers.jl:330-345), Carpenter & Kennedy (1994)
static const double kLsrkA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                 -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double kLsrkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                 1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                 2277821191437.0 / 14882151754819.0};

// ForwardEuler (kind 0), AB3 (kind 1) and LSRK54 (kind 2), filtered when ts
// carries a filter, with a built-in operator (src/timesteppers.jl:112-149,
// 389-414, 630-666).  step0 = clock.step at entry (AB3 starts with 3 Euler
// steps); AB3's RHS_-1 and RHS_-2 live in the caller's arrays m1, m2.
static int explicit_run_generic(ffhip_op* op, ffhip_rk4* ts, int kind, void* sol, double dt, int64_t nsteps,
                                int64_t step0, void* m1, void* m2, hipStream_t st) {
  FF_TRY(rk4_scratch(ts));
  void* A = ts->scratch[1];
  const size_t bytes = (size_t)ts->count * cplx_size(ts->prec);
  const bool f32 = op->prec == FFHIP_F32;
  for (int64_t k = 0; k < nsteps; ++k) {
    if (kind == 0) {
      FF_TRY(op_calcN(op, A, sol, st));
      FF_TRY(euler_update(ts, sol, A, dt, st));
    } else if (kind == 1) {
      FF_TRY(op_calcN(op, A, sol, st));
      FF_TRY(rk4_addlinear(ts, A, sol, st));
      FF_TRY(ab3_update(ts, sol, A, m1, m2, dt, step0 + k, st));
      FF_HIP(hipMemcpyAsync(m2, m1, bytes, hipMemcpyDeviceToDevice, st));  // RHS_-2 = RHS_-1
      FF_HIP(hipMemcpyAsync(m1, A, bytes, hipMemcpyDeviceToDevice, st));   // RHS_-1 = RHS
    } else {
      void* S2 = ts->scratch[4];  // the first stage (A = 0) overwrites it
      for (int i = 0; i < 5; ++i) {
        FF_TRY(op_calcN(op, A, sol, st));
        FF_TRY(rk4_addlinear(ts, A, sol, st));
        const double a = f32 ? (double)(float)kLsrkA[i] : kLsrkA[i];
        const double b = f32 ? (double)(float)kLsrkB[i] : kLsrkB[i];
        FF_TRY(lsrk54_stage(ts, sol, S2, A, a, b, dt, st));
      }
      FF_TRY(apply_filter(ts->prec, ts->count, ts->filter, sol, st));
    }
  }
  return FFHIP_OK;
}

// out = (N + L)^T x for the tracer (out, x distinct from the operator's scratch)
template <typename R>
static int tracer_rhs_adjoint(ffhip_op* op, const ffhip_rk4* ts, void* out, const void* x, hipStream_t st) {
  const Kept kept = kept_of(op);
  const int64_t nkr = op->nkr, nx = op->n[0], ny = op->n[1], nz = op->n[2];
  const double twopi = 6.283185307179586476925286766559;
  cplx<R>* s0 = reinterpret_cast<cplx<R>*>(op->spec[0]);
  cplx<R>* s1 = reinterpret_cast<cplx<R>*>(op->spec[1]);
  cplx<R>* s2 = reinterpret_cast<cplx<R>*>(op->spec[2]);
  R* p0 = reinterpret_cast<R*>(op->phys[0]);
  R* p1 = reinterpret_cast<R*>(op->phys[1]);
  R* p2 = reinterpret_cast<R*>(op->phys[2]);
  R* c = reinterpret_cast<R*>(op->phys[3]);
  const double nt = (double)nx * (double)ny * (double)nz;
  hipLaunchKernelGGL((k_tracer_adj_prep<R>), dim3(op_blocks(op->spec_count)), dim3(kOpThreads), 0, st,
                     (const cplx<R>*)x, s0, s1, s2, nkr, ny, nz, twopi / op->box[0], twopi / op->box[1],
                     twopi / op->box[2], kept);
  FF_LAUNCHED("k_tracer_adj_prep");
  FF_TRY(plan_inverse_scaled(op->plan, s0, p0, nt, st));
  FF_TRY(plan_inverse_scaled(op->plan, s1, p1, nt, st));
  FF_TRY(plan_inverse_scaled(op->plan, s2, p2, nt, st));
  hipLaunchKernelGGL((k_tracer_adj_products<R>), dim3(op_blocks(op->phys_count)), dim3(kOpThreads), 0, st,
                     (const R*)p0, (const R*)p1, (const R*)p2, c, nx, ny, nz, (const double*)op->tables,
                     op->params[0], op->params[1], op->params[2]);
  FF_LAUNCHED("k_tracer_adj_products");
  FF_TRY(ffhip_plan_forward(op->plan, c, s0, st));
  hipLaunchKernelGGL((k_tracer_adj_finish<R>), dim3(op_blocks(op->spec_count)), dim3(kOpThreads), 0, st,
                     (cplx<R>*)out, (const cplx<R>*)s0, (const cplx<R>*)x, (const double*)ts->L, nkr, ny, nz,
                     1.0 / nt, kept);
  FF_LAUNCHED("k_tracer_adj_finish");
  return FFHIP_OK;
}

// lam <- (d sol_T / d sol_0)^T lam over nsteps RK4 steps (reverse order of the
// forward stages of rk4_run_generic, same step sizes)
template <typename R>
static int rk4_adjoint_tracer(ffhip_op* op, ffhip_rk4* ts, void* lam, double dt, int64_t nsteps, hipStream_t st) {
  FF_TRY(rk4_scratch(ts));
  cplx<R>* ub = reinterpret_cast<cplx<R>*>(ts->scratch[0]);
  cplx<R>* kb = reinterpret_cast<cplx<R>*>(ts->scratch[1]);
  cplx<R>* o = reinterpret_cast<cplx<R>*>(ts->scratch[2]);
  cplx<R>* l = reinterpret_cast<cplx<R>*>(lam);
  const double h = dt;
  const double h2 = op->prec == FFHIP_F32 ? (double)((float)dt / 2.0f) : dt / 2;
  const double h6 = dt / 6, h3 = 2 * (dt / 6);
  const int64_t n = ts->count;
  auto combo = [&](int mode, double cl, double cu) -> int {
    hipLaunchKernelGGL((k_rk4_adj_combo<R>), dim3(op_blocks(n)), dim3(kOpThreads), 0, st, ub, kb, l,
                       (const cplx<R>*)o, mode, cl, cu, n);
    FF_LAUNCHED("k_rk4_adj_combo");
    return FFHIP_OK;
  };
  for (int64_t k = 0; k < nsteps; ++k) {
    FF_TRY(combo(0, h6, 0));                              // R4bar = h/6 lam
    FF_TRY((tracer_rhs_adjoint<R>(op, ts, o, kb, st)));   // stage-4 input bar
    FF_TRY(combo(1, h3, h));                              // ubar = lam + o; R3bar = h/3 lam + h o
    FF_TRY((tracer_rhs_adjoint<R>(op, ts, o, kb, st)));
    FF_TRY(combo(2, h3, h2));                             // R2bar = h/3 lam + h/2 o
    FF_TRY((tracer_rhs_adjoint<R>(op, ts, o, kb, st)));
    FF_TRY(combo(2, h6, h2));                             // R1bar = h/6 lam + h/2 o
    FF_TRY((tracer_rhs_adjoint<R>(op, ts, o, kb, st)));
    FF_TRY(combo(3, 0, 0));                               // lam = ubar + o
  }
  return FFHIP_OK;
}

static void op_free(ffhip_op* op) {
  if (!op) return;
  fused_destroy(op);
  tracer_fused_destroy(op);
  if (op->plan) ffhip_plan_destroy(op->plan);
  for (auto& p : op->spec) dev_free(p);
  for (auto& p : op->phys) dev_free(p);
  dev_free(op->tables);
  delete op;
}

static bool same_layout(const ffhip_op* op, int ndim, const int64_t* dims, int prec) {
  if (prec != op->prec || ndim != op->ndim) return false;
  if (dims[0] != op->nkr) return false;
  for (int d = 1; d < ndim; ++d)
    if (dims[d] != op->n[d]) return false;
  return true;
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_op_create(ffhip_op** out, int kind, int prec, int ndim, const int64_t* n, const double* box,
                    double aliased_fraction, const double* params, int nparams) {
  clear_error();
  FF_CHECK(out, FFHIP_ERR_ARG, "null output pointer");
  *out = nullptr;
  FF_CHECK(kind == FFHIP_OP_VORTICITY2D || kind == FFHIP_OP_TRACER3D, FFHIP_ERR_ARG, "unknown operator kind %d",
           kind);
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  FF_CHECK(n && box, FFHIP_ERR_ARG, "null size or box");
  const int want = kind == FFHIP_OP_VORTICITY2D ? 2 : 3;
  FF_CHECK(ndim == want, FFHIP_ERR_ARG, "operator kind %d needs ndim = %d", kind, want);
  for (int d = 0; d < ndim; ++d) {
    FF_CHECK(n[d] >= 2, FFHIP_ERR_ARG, "grid sizes must be >= 2");
    FF_CHECK(n[d] % 2 == 0, FFHIP_ERR_DOMAIN, "grid sizes must be even");
    FF_CHECK(box[d] > 0, FFHIP_ERR_ARG, "domain lengths must be positive");
  }
  FF_CHECK(aliased_fraction >= 0 && aliased_fraction < 1, FFHIP_ERR_ARG, "aliased_fraction must be in [0, 1)");
  if (kind == FFHIP_OP_TRACER3D) FF_CHECK(params && nparams >= 6, FFHIP_ERR_ARG, "tracer needs A,B,C,x0,y0,z0");
  ffhip_op* op = new (std::nothrow) ffhip_op();
  FF_CHECK(op, FFHIP_ERR_OOM, "host allocation failed");
  op->kind = kind;
  op->prec = prec;
  op->ndim = ndim;
  for (int d = 0; d < 3; ++d) {
    op->n[d] = d < ndim ? n[d] : 1;
    op->box[d] = d < ndim ? box[d] : 1.0;
  }
  op->af = aliased_fraction;
  for (int i = 0; i < nparams && i < 8; ++i) op->params[i] = params[i];
  op->nparams = nparams;
  op->nkr = n[0] / 2 + 1;
  op->spec_count = op->nkr * op->n[1] * op->n[2];
  op->phys_count = op->n[0] * op->n[1] * op->n[2];
  // alias ranges: x uses kralias, y/z use kalias-style ranges (src/domains.jl:210-211, 351-353)
  int rc = ffhip_aliased_wavenumbers(op->n[0], op->nkr, aliased_fraction, nullptr, nullptr, &op->lo[0], &op->hi[0]);
  for (int d = 1; d < 3 && rc == FFHIP_OK; ++d) {
    if (d < ndim)
      rc = ffhip_aliased_wavenumbers(op->n[d], op->n[d] / 2 + 1, aliased_fraction, &op->lo[d], &op->hi[d], nullptr,
                                     nullptr);
    else {
      op->lo[d] = 1;
      op->hi[d] = 0;
    }
  }
  if (rc == FFHIP_OK) rc = ffhip_plan_rfft(&op->plan, ndim, n, 1, prec);
  const size_t sb = (size_t)op->spec_count * cplx_size(prec), pb = (size_t)op->phys_count * real_size(prec);
  for (int i = 0; i < 3 && rc == FFHIP_OK; ++i) rc = dev_alloc(&op->spec[i], sb);
  for (int i = 0; i < (kind == FFHIP_OP_TRACER3D ? 4 : 3) && rc == FFHIP_OK; ++i) rc = dev_alloc(&op->phys[i], pb);
  if (rc == FFHIP_OK && kind == FFHIP_OP_TRACER3D) {
    const int64_t nx = op->n[0], ny = op->n[1], nz = op->n[2];
    std::vector<double> tab(2 * (nx + ny + nz));
    const double x0[3] = {params[3], params[4], params[5]};
    const int64_t ns[3] = {nx, ny, nz};
    size_t off = 0;
    for (int d = 0; d < 3; ++d) {
      const double dx = op->box[d] / (double)ns[d];
      for (int64_t i = 0; i < ns[d]; ++i) {
        const double x = x0[d] + dx * (double)i;  // range(x0, step=dx) (src/domains.jl:333-335)
        tab[off + i] = std::sin(x);
        tab[off + ns[d] + i] = std::cos(x);
      }
      off += 2 * ns[d];
    }
    rc = dev_alloc(&op->tables, tab.size() * sizeof(double));
    if (rc == FFHIP_OK) {
      hipError_t e = hipMemcpy(op->tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = fail(FFHIP_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
    }
  }
  if (rc == FFHIP_OK) rc = fused_create(op);
  if (rc == FFHIP_OK) rc = tracer_fused_create(op);
  if (rc != FFHIP_OK) {
    op_free(op);
    return rc;
  }
  *out = op;
  return FFHIP_OK;
}

int ffhip_op_destroy(ffhip_op* op) {
  clear_error();
  op_free(op);
  return FFHIP_OK;
}

int ffhip_op_calcN(ffhip_op* op, void* N, const void* sol, double t, void* stream) {
  clear_error();
  (void)t;  // both built-in operators are autonomous
  FF_CHECK(op && N && sol, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(N != sol, FFHIP_ERR_ARG, "N and sol must be distinct arrays");
  return op_calcN(op, N, sol, as_stream(stream));
}

int ffhip_op_etdrk4_run(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, void* stream) {
  clear_error();
  FF_CHECK(op && ts && sol, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 0, FFHIP_ERR_ARG, "nsteps must be >= 0");
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  if (nsteps == 0) return FFHIP_OK;
  hipStream_t st = as_stream(stream);
  if (op->fused && !ts->cplx_coef) return fused_etdrk4_run(op, ts, sol, nsteps, st);
  return etd_run_generic(op, ts, sol, nsteps, st);
}

int ffhip_op_rk4_run(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, void* stream) {
  clear_error();
  FF_CHECK(op && ts && sol, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 0, FFHIP_ERR_ARG, "nsteps must be >= 0");
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  const double h = op->prec == FFHIP_F32 ? (double)(float)dt : dt;
  if (nsteps == 0) return FFHIP_OK;
  if (op->fused3 && !ts->cplx_coef && !ts->filter) return tracer_fused_rk4_run(op, ts, sol, h, nsteps, as_stream(stream));
  return rk4_run_generic(op, ts, sol, h, nsteps, as_stream(stream));
}

int ffhip_op_etdrk4_forward_ckpt(ffhip_op* op, ffhip_etdrk4* ts, const void* sol0, void* solT, int64_t nsteps,
                                 int64_t ncheck, void* stream) {
  clear_error();
  FF_CHECK(op && ts && sol0 && solT, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 1 && ncheck >= 1 && ncheck <= nsteps, FFHIP_ERR_ARG,
           "need nsteps >= 1 and 1 <= ncheck <= nsteps (got %lld, %lld)", (long long)nsteps, (long long)ncheck);
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  FF_CHECK(op->fused && !ts->cplx_coef, FFHIP_ERR_UNSUPPORTED,
           "the adjoint runs on the fused path only (power-of-two 2D vorticity grid, real L)");
  return fused_forward_ckpt(op, ts, sol0, solT, nsteps, ncheck, as_stream(stream));
}

int ffhip_op_etdrk4_adjoint(ffhip_op* op, ffhip_etdrk4* ts, const void* lamT, void* lam0, int64_t nsteps,
                            int64_t ncheck, void* stream) {
  clear_error();
  FF_CHECK(op && ts && lamT && lam0, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 1 && ncheck >= 1 && ncheck <= nsteps, FFHIP_ERR_ARG,
           "need nsteps >= 1 and 1 <= ncheck <= nsteps (got %lld, %lld)", (long long)nsteps, (long long)ncheck);
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  FF_CHECK(op->fused && !ts->cplx_coef, FFHIP_ERR_UNSUPPORTED,
           "the adjoint runs on the fused path only (power-of-two 2D vorticity grid, real L)");
  return fused_adjoint(op, ts, lamT, lam0, nsteps, ncheck, as_stream(stream));
}

int ffhip_op_checkpoint_id(const ffhip_op* op, int64_t* id) {
  clear_error();
  FF_CHECK(op && id, FFHIP_ERR_ARG, "null argument");
  *id = op->fused ? fused_checkpoint_id(op) : 0;
  return FFHIP_OK;
}

int ffhip_op_rk4_adjoint(ffhip_op* op, ffhip_rk4* ts, const void* lamT, void* lam0, double dt, int64_t nsteps,
                         void* stream) {
  clear_error();
  FF_CHECK(op && ts && lamT && lam0, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 0, FFHIP_ERR_ARG, "nsteps must be >= 0");
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  FF_CHECK(op->kind == FFHIP_OP_TRACER3D && !ts->cplx_coef, FFHIP_ERR_UNSUPPORTED,
           "the RK4 adjoint is implemented for the tracer operator with real L");
  hipStream_t st = as_stream(stream);
  if (lam0 != lamT)
    FF_HIP(hipMemcpyAsync(lam0, lamT, (size_t)op->spec_count * cplx_size(op->prec), hipMemcpyDeviceToDevice, st));
  const double h = op->prec == FFHIP_F32 ? (double)(float)dt : dt;
  if (nsteps == 0) return FFHIP_OK;
  if (op->fused3 && !ts->filter) return tracer_fused_rk4_adjoint(op, ts, lam0, h, nsteps, st);
  return op->prec == FFHIP_F32 ? rk4_adjoint_tracer<float>(op, ts, lam0, h, nsteps, st)
                               : rk4_adjoint_tracer<double>(op, ts, lam0, h, nsteps, st);
}

int ffhip_op_explicit_run(ffhip_op* op, ffhip_rk4* ts, int kind, void* sol, double dt, int64_t nsteps,
                          int64_t step0, void* rhs_m1, void* rhs_m2, void* stream) {
  clear_error();
  FF_CHECK(op && ts && sol, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(kind >= 0 && kind <= 2, FFHIP_ERR_ARG, "explicit stepper kind must be 0 (ForwardEuler), 1 (AB3) or "
                                                  "2 (LSRK54), got %d", kind);
  FF_CHECK(nsteps >= 0 && step0 >= 0, FFHIP_ERR_ARG, "nsteps and step0 must be >= 0");
  FF_CHECK(same_layout(op, ts->ndim, ts->dims, ts->prec), FFHIP_ERR_ARG,
           "stepper shape/precision does not match the operator's rfft layout");
  FF_CHECK(kind != 1 || (rhs_m1 && rhs_m2 && rhs_m1 != rhs_m2), FFHIP_ERR_ARG,
           "AB3 needs two distinct arrays for RHS_-1 and RHS_-2");
  const double h = op->prec == FFHIP_F32 ? (double)(float)dt : dt;
  return explicit_run_generic(op, ts, kind, sol, h, nsteps, step0, rhs_m1, rhs_m2, as_stream(stream));
}

int ffhip_op_is_fused(const ffhip_op* op, int* fused) {
  clear_error();
  FF_CHECK(op && fused, FFHIP_ERR_ARG, "null argument");
  *fused = op->fused != nullptr || op->fused3 != nullptr;
  return FFHIP_OK;
}

int ffhip_op_diagnostics(ffhip_op* op, const void* sol, double* out, void* stream) {
  clear_error();
  FF_CHECK(op && sol && out, FFHIP_ERR_ARG, "null argument");
  hipStream_t st = as_stream(stream);
  double2* buf = nullptr;
  FF_TRY(red_buffer(&buf));
  const int64_t total = op->spec_count;
  int blocks = (int)((total + kRedThreads - 1) / kRedThreads);
  if (blocks > kRedBlocks / 2 - 1) blocks = kRedBlocks / 2 - 1;  // partials stay below the result slots
  const double twopi = 6.283185307179586476925286766559;
  const double dk = twopi / op->box[0], dl = twopi / op->box[1], dm = op->ndim > 2 ? twopi / op->box[2] : 0.0;
  if (op->prec == FFHIP_F32)
    hipLaunchKernelGGL((k_diag2<float>), dim3(blocks), dim3(kRedThreads), 0, st, (const float2*)sol, op->nkr,
                       op->n[1], op->n[2], dk, dl, dm, buf);
  else
    hipLaunchKernelGGL((k_diag2<double>), dim3(blocks), dim3(kRedThreads), 0, st, (const double2*)sol, op->nkr,
                       op->n[1], op->n[2], dk, dl, dm, buf);
  FF_LAUNCHED("k_diag2");
  hipLaunchKernelGGL(k_diag2_final, dim3(1), dim3(kRedThreads), 0, st, buf, blocks, buf + kRedBlocks - 2);
  FF_LAUNCHED("k_diag2_final");
  double2 h[2];
  FF_HIP(hipMemcpyAsync(h, buf + kRedBlocks - 2, sizeof h, hipMemcpyDeviceToHost, st));
  double c0[2] = {0, 0};
  if (op->kind == FFHIP_OP_TRACER3D) FF_HIP(hipMemcpyAsync(c0, sol, cplx_size(op->prec), hipMemcpyDeviceToHost, st));
  FF_HIP(hipStreamSynchronize(st));
  double N2 = 1, V = 1;
  for (int d = 0; d < op->ndim; ++d) {
    N2 *= (double)op->n[d] * (double)op->n[d];
    V *= op->box[d];
  }
  const double norm = V / N2;  // parsevalsum2 normalisation (src/utils.jl:122)
  const double s0 = (h[0].x + h[0].y) * norm, s1 = (h[1].x + h[1].y) * norm;
  if (op->kind == FFHIP_OP_VORTICITY2D) {
    out[0] = 0.5 * s1 / V;  // energy    1/2 <|u|^2> = 1/2 sum |k|^2 |psi|^2
    out[1] = 0.5 * s0 / V;  // enstrophy 1/2 <zeta^2>
  } else {
    double re0 = op->prec == FFHIP_F32 ? (double)reinterpret_cast<float*>(c0)[0] : c0[0];
    out[0] = 0.5 * s0 / V;
    out[1] = re0 / (double)op->phys_count;
  }
  if (!std::isfinite(out[0]) || !std::isfinite(out[1])) return fail(FFHIP_ERR_NONFINITE, "non-finite diagnostics");
  return FFHIP_OK;
}

}  // extern "C"
