/*
 * ffhip.h -- C ABI of libffhip.so, an MI355X (gfx950) implementation of
 * FourierFlows.jl's pseudospectral step.
 *
 * Every entry point returns an int status (FFHIP_OK = 0, negative on error)
 * and sets a thread-local message readable with ffhip_last_error().  No
 * pointer passed in is retained after a call returns, except by the objects
 * the create functions return.  Device pointers are plain HIP device
 * addresses (from ffhip_malloc, hipMalloc or a framework allocator).  A
 * `stream` argument of NULL means the HIP null stream.
 *
 * Array layout: Julia column-major with x fastest.  `dims[0]` is the x
 * extent (nx for physical fields, nkr = nx/2+1 for rfft spectral fields),
 * `dims[1]` is y, `dims[2]` is z.  Complex arrays are interleaved (re, im).
 * `prec` is FFHIP_F32 (Float32 / ComplexF32) or FFHIP_F64.
 *
 * Reference interfaces replaced (FourierFlows.jl v0.10.7):
 *   plan_flows_rfft / plan_flows_fft ........ src/domains.jl:2-5, :206-208, :347-349
 *   mul!(fh, rfftplan, f) ................... src/domains.jl:5 (call: src/diffusion.jl:139)
 *   ldiv!(f, rfftplan, fh) .................. src/diffusion.jl:137 (AbstractFFTs scaled inverse)
 *   getaliasedwavenumbers / dealias! ........ src/domains.jl:408-421, :428-476
 *   getexpLs / getetdcoeffs ................. src/timesteppers.jl:673-721
 *   ETDRK4TimeStepper and its substeps ...... src/timesteppers.jl:434-463, :501-547
 *   RK4TimeStepper and its substeps ......... src/timesteppers.jl:180-197, :225-274
 *   parsevalsum2 / parsevalsum .............. src/utils.jl:113-183
 *   stepforward!(prob, nsteps) .............. src/timesteppers.jl:6-20 (fused built-in operators)
 *   zeros(dev, T, dims) / device_array ...... src/utils.jl:79-80, :329-332
 */
#ifndef FFHIP_H
#define FFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define FFHIP_OK 0
#define FFHIP_ERR_DOMAIN (-1)      /* bad grid size (odd n): Julia DomainError, src/domains.jl:66,179,316 */
#define FFHIP_ERR_ARG (-2)         /* bad argument (null pointer, bad precision code, bad rank) */
#define FFHIP_ERR_OOM (-3)         /* device allocation failed */
#define FFHIP_ERR_HIP (-4)         /* HIP runtime / launch failure */
#define FFHIP_ERR_UNSUPPORTED (-5) /* valid request this build does not implement */
#define FFHIP_ERR_NONFINITE (-6)   /* NaN/Inf found by a checked reduction */

#define FFHIP_F32 32
#define FFHIP_F64 64

typedef struct ffhip_plan ffhip_plan;
typedef struct ffhip_etdrk4 ffhip_etdrk4;
typedef struct ffhip_rk4 ffhip_rk4;
typedef struct ffhip_op ffhip_op;
typedef struct ffhip_ensemble ffhip_ensemble;

/* ---------------------------------------------------------------- runtime
 * Device selection and memory: the HIPArray analogue of `zeros(dev,T,dims)`
 * and `device_array(dev)` (src/utils.jl:79-80, :329-332). */
int ffhip_init(int device);
const char* ffhip_last_error(void);
int ffhip_version(int* major, int* minor, int* patch);
int ffhip_malloc(void** ptr, size_t bytes);            /* zero-filled, like zeros(dev, T, dims) */
int ffhip_free(void* ptr);
int ffhip_memcpy_htod(void* dst, const void* src, size_t bytes);
int ffhip_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int ffhip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int ffhip_memset_zero(void* dst, size_t bytes, void* stream);
int ffhip_synchronize(void);

/* ---------------------------------------------------------------- FFT plans
 * plan_flows_rfft (src/domains.jl:3,5) / plan_flows_fft (:2,4).
 * n[0..ndim-1]: transform sizes, x first.  `batch` independent transforms
 * are stored back to back.  Any even n (rfft x axis) / any n >= 1 (other
 * axes) is accepted; factors 2,3,4,5,8 use unrolled butterflies, other
 * primes a generic DFT stage. */
int ffhip_plan_rfft(ffhip_plan** plan, int ndim, const int64_t* n, int64_t batch, int prec);
int ffhip_plan_fft(ffhip_plan** plan, int ndim, const int64_t* n, int64_t batch, int prec);
int ffhip_plan_destroy(ffhip_plan* plan);
/* mul!(out, plan, in): unnormalised forward transform (sign -1).
 * rfft: in real (n0,n1,..), out complex (n0/2+1, n1, ..).  In place is not allowed. */
int ffhip_plan_forward(ffhip_plan* plan, const void* in, void* out, void* stream);
/* ldiv!(out, plan, in): inverse transform scaled by 1/prod(n).  Unlike
 * FFTW/CUFFT the input is preserved, so callers that deepcopy before ldiv!
 * (src/diffusion.jl:158-159) are unaffected.  rfft: the imaginary parts of
 * the x-DC and x-Nyquist entries are ignored after the y/z passes, as FFTW. */
int ffhip_plan_inverse(ffhip_plan* plan, const void* in, void* out, void* stream);
/* Transposes of the two transforms for the real inner products sum(a*b)
 * (physical) and Re sum conj(a) b (spectral, stored entries); no counterpart
 * in the reference, they carry gradients through mul!/ldiv!.
 * forward_adjoint:  mul!^T(Y) = prod(n) ldiv!(Y / W)      (spectral in, physical out)
 * inverse_adjoint:  ldiv!^T(phi) = W mul!(phi) / prod(n)  (physical in, spectral out)
 * W = 1 on the kx = 0 and kx = nx/2 entries of an rfft layout and 2
 * elsewhere; W = 1 for complex plans. */
int ffhip_plan_forward_adjoint(ffhip_plan* plan, const void* in, void* out, void* stream);
int ffhip_plan_inverse_adjoint(ffhip_plan* plan, const void* in, void* out, void* stream);
/* plan query: ndim, sizes, batch, precision, kind (0 = rfft, 1 = fft). */
int ffhip_plan_info(const ffhip_plan* plan, int* ndim, int64_t* n, int64_t* batch, int* prec, int* kind);

/* ---------------------------------------------------------------- dealiasing
 * getaliasedwavenumbers (src/domains.jl:408-421), host only: 1-based
 * inclusive index ranges; lo > hi means "nothing" (aliased_fraction == 0).
 * Returns FFHIP_ERR_ARG if aliased_fraction >= 1 (:415). */
int ffhip_aliased_wavenumbers(int64_t nk, int64_t nkr, double aliased_fraction,
                              int64_t* kalias_lo, int64_t* kalias_hi,
                              int64_t* kralias_lo, int64_t* kralias_hi);
/* dealias!(fh, grid) (src/domains.jl:428-476): zero fh[lo[d]:hi[d]] along each
 * dimension d (1-based inclusive ranges as above; lo > hi skips d). */
int ffhip_dealias(void* fh, int prec, int ndim, const int64_t* dims, int64_t batch,
                  const int64_t* lo, const int64_t* hi, void* stream);

/* ---------------------------------------------------------------- ETDRK4
 * getetdcoeffs (src/timesteppers.jl:689-718), host only, FP64: contour mean
 * over ncirc points of radius rcirc.  L_im == NULL means real L (then only
 * the real outputs are written; the *_im outputs may be NULL). */
int ffhip_etdcoeffs(double dt, const double* L_re, const double* L_im, int64_t count,
                    int ncirc, double rcirc,
                    double* zeta_re, double* alpha_re, double* beta_re, double* gamma_re,
                    double* zeta_im, double* alpha_im, double* beta_im, double* gamma_im);
/* ETDRK4TimeStepper(equation, dt, dev) (src/timesteppers.jl:456-463).
 * L (host, FP64, `count` = prod(dims) entries, already broadcast to dims).
 * dt is first rounded to the state's float type (:457).  The coefficient
 * arrays zeta, alpha, beta, Gamma, exp(dt L), exp(dt L/2) are kept in FP64 on
 * the device; the scratch arrays sol1, sol2, N1..N4 are in `prec`. */
int ffhip_etdrk4_create(ffhip_etdrk4** ts, const double* L_re, const double* L_im,
                        int ndim, const int64_t* dims, double dt, int prec);
int ffhip_etdrk4_destroy(ffhip_etdrk4* ts);
/* which: 0 sol1, 1 sol2, 2 N1, 3 N2, 4 N3, 5 N4 (state-precision complex). */
int ffhip_etdrk4_scratch(ffhip_etdrk4* ts, int which, void** ptr);
/* which: 0 zeta, 1 alpha, 2 beta, 3 Gamma, 4 expLdt, 5 exp_half_Ldt.
 * FP64, real (is_complex = 0) or complex interleaved (is_complex = 1). */
int ffhip_etdrk4_coeff(ffhip_etdrk4* ts, int which, void** ptr, int* is_complex);
/* ETDRK4substep12! (:506-510): out = exp_half_Ldt*sol + zeta*N */
int ffhip_etdrk4_substep12(ffhip_etdrk4* ts, void* out, const void* sol, const void* N, void* stream);
/* ETDRK4substep3! (:512-516): out = exp_half_Ldt*sol1 + zeta*(2*N3 - N1) */
int ffhip_etdrk4_substep3(ffhip_etdrk4* ts, void* out, const void* sol1, const void* N1,
                          const void* N3, void* stream);
/* ETDRK4update! (:501-504): sol = expLdt*sol + alpha*N1 + 2*beta*(N2+N3) + Gamma*N4 */
int ffhip_etdrk4_update(ffhip_etdrk4* ts, void* sol, const void* N1, const void* N2,
                        const void* N3, const void* N4, void* stream);

/* ---------------------------------------------------------------- RK4
 * RK4TimeStepper (src/timesteppers.jl:180-197); L kept in FP64 on device. */
int ffhip_rk4_create(ffhip_rk4** ts, const double* L_re, const double* L_im, int ndim,
                     const int64_t* dims, int prec);
int ffhip_rk4_destroy(ffhip_rk4* ts);
/* which: 0 sol1, 1 RHS1, 2 RHS2, 3 RHS3, 4 RHS4 */
int ffhip_rk4_scratch(ffhip_rk4* ts, int which, void** ptr);
/* addlinearterm! (:225-229): RHS += L*sol */
int ffhip_rk4_addlinearterm(ffhip_rk4* ts, void* RHS, const void* sol, void* stream);
/* substepsol! (:231-235): newsol = sol + dt*RHS */
int ffhip_rk4_substepsol(ffhip_rk4* ts, void* newsol, const void* sol, const void* RHS,
                         double dt, void* stream);
/* RK4update! (:260-264): sol += dt/6*(RHS1 + 2*RHS2 + 2*RHS3 + RHS4) */
int ffhip_rk4_update(ffhip_rk4* ts, void* sol, const void* R1, const void* R2, const void* R3,
                     const void* R4, double dt, void* stream);

/* ---------------------------------------------------------------- filters and explicit steppers
 * makefilter (src/domains.jl:506-543) is computed by the host; attaching it
 * to a stepper makes it the Filtered* variant: sol *= filter after every
 * step (src/timesteppers.jl:142-149, 276-285, 405-414, 549-558, 653-666).
 * filter: host FP64, prod(dims) entries; NULL detaches.  The fused and
 * generic drivers (ffhip_op_*_run) apply it. */
int ffhip_etdrk4_set_filter(ffhip_etdrk4* ts, const double* filter);
int ffhip_rk4_set_filter(ffhip_rk4* ts, const double* filter);
int ffhip_etdrk4_apply_filter(ffhip_etdrk4* ts, void* sol, void* stream);
int ffhip_rk4_apply_filter(ffhip_rk4* ts, void* sol, void* stream);
/* ForwardEuler / AB3 / LSRK54 (src/timesteppers.jl:98-150, 318-414, 565-667)
 * share the RK4 object (L, filter, scratch).
 * euler_update (:112-118, :142-149): sol = [filter*](sol + dt (L sol + N)).
 * ab3_update (AB3update!, :630-638): step < 3: sol += dt RHS, else
 *   sol += dt (23/12 RHS - 16/12 RHSm1 + 5/12 RHSm2); then the filter.
 * lsrk54_stage (:389-397): S2 = A S2 + dt RHS; sol += B S2 (A = 0 does not
 *   read S2, which is what the reference's S2 = 0 start gives). */
int ffhip_euler_update(ffhip_rk4* ts, void* sol, const void* N, double dt, void* stream);
int ffhip_ab3_update(ffhip_rk4* ts, void* sol, const void* RHS, const void* RHSm1, const void* RHSm2, double dt,
                     int64_t step, void* stream);
int ffhip_lsrk54_stage(ffhip_rk4* ts, void* sol, void* S2, const void* RHS, double A, double B, double dt,
                       void* stream);

/* ---------------------------------------------------------------- reductions
 * parsevalsum2 / parsevalsum (src/utils.jl:113-183), compensated FP64.
 * rfft != 0: dims[0] == nkr and the kx = 0 and kx = nx/2 entries get weight 1,
 * the rest weight 2 (:114-117).  The sum is multiplied by `norm`
 * (Lx*Ly/(nx^2 ny^2) in 2D, :122).  Returns FFHIP_ERR_NONFINITE if the field
 * holds NaN/Inf (the sum is still written). */
int ffhip_parsevalsum2(const void* uh, int prec, int ndim, const int64_t* dims, int rfft,
                       int64_t nx, double norm, double* out, void* stream);
int ffhip_parsevalsum(const void* uh, int prec, int ndim, const int64_t* dims, int rfft,
                      int64_t nx, double norm, double* out, void* stream);
/* Allocation-free variant of the two sums above.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * mode 0 = parsevalsum2, 1 = parsevalsum, same arguments and arithmetic; the
 * result (the same double) is written to device memory out_dev on `stream`
 * with no host synchronisation, so no non-finite check is made -- the caller
 * tests the value.  Partial sums use a fixed 16 KiB scratch array per GPU
 * (1024 double-double partials; inputs of any size fit, the grid-stride pass
 * never makes more partials); the device address is looked up by ffhip_init,
 * or by the first call on a device ffhip_init did not set up.  Calls that
 * overlap on one GPU race on that array. */
int ffhip_parseval_noalloc(int mode, const void* uh, int prec, int ndim, const int64_t* dims,
                           int rfft, int64_t nx, double norm, double* out_dev, void* stream);

/* ---------------------------------------------------------------- built-in operators
 * The reference's calcN! is arbitrary user Julia.  A C ABI cannot compile
 * it, so the physics that the hot path needs is built in and selected by
 * kind; each operator also runs whole fused time steps
 * (stepforward!(prob, nsteps), src/timesteppers.jl:14-20).
 *
 * FFHIP_OP_VORTICITY2D: N = -J(psi, zeta), psi_hat = -zeta_hat/|k|^2, i.e.
 *   N_hat = -i kr (u zeta)^ - i l (v zeta)^, u = -psi_y, v = psi_x, dealiased.
 *   ndim = 2; params: none.
 * FFHIP_OP_TRACER3D: N = -div(u c) for the steady ABC flow
 *   u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x), dealiased.
 *   ndim = 3; params = {A, B, C, x0, y0, z0}.
 * Both evaluate N on the dealiased state (modes dealias! would zero are
 * read as zero) and return a dealiased N.  n: grid sizes, box: domain
 * lengths (Lx, Ly[, Lz]). */
#define FFHIP_OP_VORTICITY2D 1
#define FFHIP_OP_TRACER3D 2
int ffhip_op_create(ffhip_op** op, int kind, int prec, int ndim, const int64_t* n,
                    const double* box, double aliased_fraction, const double* params, int nparams);
int ffhip_op_destroy(ffhip_op* op);
/* calcN!(N, sol, t, ...) for the built-in operator. */
int ffhip_op_calcN(ffhip_op* op, void* N, const void* sol, double t, void* stream);
/* nsteps fused ETDRK4 steps of `sol` (in place) with the coefficients of ts
 * (created for this operator's grid and precision). */
int ffhip_op_etdrk4_run(ffhip_op* op, ffhip_etdrk4* ts, void* sol, int64_t nsteps, void* stream);
/* nsteps fused RK4 steps with the linear term of ts and step dt. */
int ffhip_op_rk4_run(ffhip_op* op, ffhip_rk4* ts, void* sol, double dt, int64_t nsteps, void* stream);
/* nsteps of ForwardEuler (kind 0), AB3 (1) or LSRK54 (2) with the built-in
 * operator, filtered if ts has a filter; step0 = clock.step at entry (AB3
 * takes forward Euler steps while step < 3).  AB3 reads and updates the
 * caller's RHS_-1, RHS_-2 arrays (rhs_m1, rhs_m2; ignored otherwise). */
int ffhip_op_explicit_run(ffhip_op* op, ffhip_rk4* ts, int kind, void* sol, double dt, int64_t nsteps,
                          int64_t step0, void* rhs_m1, void* rhs_m2, void* stream);
/* Diagnostics on the rfft state: out[0] = energy, out[1] = enstrophy
 * (vorticity: 1/2<|u|^2>, 1/2<zeta^2>; tracer: out[0] = 1/2<c^2>, out[1] = <c>). */
int ffhip_op_diagnostics(ffhip_op* op, const void* sol, double* out, void* stream);

/* Discrete adjoint of nsteps fused ETDRK4 steps.  No counterpart in the
 * reference (it has no adjoint); for gradient-based estimation of initial
 * conditions, on the fused path only (power-of-two 2D vorticity, real L).
 * forward_ckpt: solT = ETDRK4^nsteps(sol0) exactly as ffhip_op_etdrk4_run
 *   computes it (solT may equal sol0), keeping ncheck uniformly spaced
 *   checkpoints of the kept band on the device (segment ceil(nsteps/ncheck)).
 * adjoint: lam0 = (d solT / d sol0)^T lamT, with the transpose taken for the
 *   real inner product <a, b> = Re sum conj(a) b over the stored rfft entries.
 *   Recomputes each segment from its checkpoint, keeping the stage fields
 *   (nx*ny*3 reals per stage) of one segment.  Needs the forward_ckpt call of
 *   the same op, ts, nsteps and ncheck just before it; lam0 may equal lamT. */
int ffhip_op_etdrk4_forward_ckpt(ffhip_op* op, ffhip_etdrk4* ts, const void* sol0, void* solT, int64_t nsteps,
                                 int64_t ncheck, void* stream);
int ffhip_op_etdrk4_adjoint(ffhip_op* op, ffhip_etdrk4* ts, const void* lamT, void* lam0, int64_t nsteps,
                            int64_t ncheck, void* stream);
/* Identifier of the checkpointed run the operator currently holds (0: none).
 * Every forward_ckpt call takes a new one, unique in the process, so a
 * caller that keeps several differentiable solves on one operator (chained
 * solves, a loss over two solves) can tell whether the checkpoints are still
 * those of its own run, and re-run forward_ckpt from its saved sol0 if not. */
int ffhip_op_checkpoint_id(const ffhip_op* op, int64_t* id);

/* Discrete adjoint of nsteps RK4 steps of the tracer operator (no
 * counterpart in the reference): lam0 = (d solT / d sol0)^T lamT for the
 * steps ffhip_op_rk4_run takes with the same ts and dt, same inner product
 * as above.  The tracer N is linear in c (steady velocity), so no forward
 * states, checkpoints or recompute are needed.  lam0 may equal lamT. */
int ffhip_op_rk4_adjoint(ffhip_op* op, ffhip_rk4* ts, const void* lamT, void* lam0, double dt, int64_t nsteps,
                         void* stream);

/* Is the fused path active for this operator (power-of-two 2D vorticity grid)? */
int ffhip_op_is_fused(const ffhip_op* op, int* fused);

/* ---------------------------------------------------------------- ensembles
 * No counterpart in the reference (it steps one Problem at a time): a batch
 * of independent 2D vorticity problems (the FFHIP_OP_VORTICITY2D physics,
 * ETDRK4 with a common dt) on power-of-two grids of mixed sizes in
 * 64..1024, stepped together: one kernel launch per pass covers every
 * member, and steps are replayed from a captured HIP graph.
 * n[2m], n[2m+1]: member m's nx, ny; box[2m], box[2m+1]: its Lx, Ly;
 * L[m]: its linear operator (host FP64, nkr*ny entries, x fastest, real).
 * autotune != 0: on the first run, time each valid tile/radix variant on
 * the ensemble and keep the fastest (otherwise the single-grid kernels'
 * radix schedule is used).  Each member's result agrees with the one
 * ffhip_op_etdrk4_run gives to a few ulps. */
int ffhip_ensemble_create(ffhip_ensemble** e, int prec, int count, const int64_t* n, const double* box,
                          double aliased_fraction, const double* const* L, double dt, int autotune);
int ffhip_ensemble_destroy(ffhip_ensemble* e);
/* nsteps ETDRK4 steps of every member; sols[m] is member m's rfft state on
 * the device (nkr x ny, state precision), updated in place. */
int ffhip_ensemble_run(ffhip_ensemble* e, void* const* sols, int64_t nsteps, void* stream);
/* the variant in use: threads per workgroup, the points per thread of every
 * grid size and pass -- *prule bit s (n = 64 << s) set: 16 points per thread
 * in the column passes of size-n members, else 8; bit 8 + s the same for the
 * row passes -- and, if tuned, its measured ms per ensemble step. */
int ffhip_ensemble_info(const ffhip_ensemble* e, int* threads, int* prule, double* ms_per_step);

/* ---------------------------------------------------------------- profiling
 * No counterpart in the reference.  Between start and stop the fused step
 * brackets every kernel launch with HIP events on its stream; stop
 * synchronises and writes a JSON summary {"kernels": [{"name", "count",
 * "total_ms", "bytes"}]} where bytes are the algorithmic bytes each launch
 * must move (summed over launches). */
int ffhip_profile_start(void);
int ffhip_profile_stop(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* FFHIP_H */
