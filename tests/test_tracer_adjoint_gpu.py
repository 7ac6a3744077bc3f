"""GPU discrete adjoint of the RK4 tracer step.

* FP64 gradient equals the FP64 oracle's RK4 adjoint gradient;
* dot-product identity of the n-step map and its transpose;
* central-FD Taylor test (the misfit is quadratic: the FD is exact up to
  roundoff for every h);
* the 256^3 FP32 workload (300 RK4 steps, kappa = 1e-3) against the FP64
  GPU gradient.  dt = 0.0025: with the ABC flow's |u| up to 2 per component
  and kept wavenumbers up to 85 per axis, RK4 at dt = 0.01 is outside its
  stability region (|k.u| dt reaches ~5 > 2.83) and the run overflows in
  FP64 as in FP32; dt = 0.0025 keeps |k.u| dt below ~1.3.
"""

import numpy as np
import pytest

from oracle import adjoint as oa
from oracle import grids as og
from oracle import physics as oph

pytestmark = pytest.mark.gpu


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _setup(ff, n, kappa, dt, nsteps, T):
    import torch

    prob = ff.TracerProblem(n=n, kappa=kappa, dt=dt, T=T)
    g = prob.grid
    c0 = ff.tracer_ic(g, seed=1, ngauss=16, sigma=0.3)
    s = torch.zeros_like(prob.sol)
    ff.mul(s, g.rfftplan, ff.tracer_ic(g, seed=2, ngauss=16, sigma=0.3))
    ff.dealias(s, g)
    prob.sol.copy_(s)
    ff.stepforward(prob, nsteps)
    cobs = torch.zeros_like(c0)
    ff.ldiv(cobs, g.rfftplan, prob.sol)
    return prob, c0, cobs


def test_tracer_ic_matches_oracle(ff):
    from oracle import ics as oics

    g = ff.ThreeDGrid(nx=24, Lx=2 * np.pi, x0=0.0, y0=0.0, z0=0.0)
    o = og.ThreeDGrid(24, 2 * np.pi, x0=0, y0=0, z0=0)
    c = _host(ff.tracer_ic(g, seed=3, ngauss=5, sigma=0.4))
    np.testing.assert_allclose(c, oics.tracer_ic(o, seed=3, ngauss=5, sigma=0.4), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,nsteps", [(16, 10), (24, 7)])
def test_fp64_gradient_matches_oracle(ff, n, nsteps):
    kappa, dt = 1e-2, 0.05
    prob, c0, cobs = _setup(ff, n, kappa, dt, nsteps, np.float64)
    J, grad = ff.adjoint.tracer_misfit_gradient(prob, c0, cobs, nsteps)
    g = og.ThreeDGrid(n, 2 * np.pi, x0=0, y0=0, z0=0)
    params = oph.TracerParams(kappa=kappa)
    adj = oa.RK4TracerAdjoint(g, oph.abc_velocity(g, params), oph.tracer_linear(g, kappa), dt)
    Jr, gr = adj.gradient(_host(c0), _host(cobs), nsteps)
    assert abs(J - Jr) < 1e-12 * Jr, (J, Jr)
    err = np.linalg.norm(_host(grad) - gr) / np.linalg.norm(gr)
    assert err < 1e-11, err


def test_dot_product_identity(ff):
    import torch

    from fourierflows_hip import _lib

    n, nsteps = 32, 6
    prob = ff.TracerProblem(n=n, kappa=1e-2, dt=0.05, T=np.float64)
    g, op, ts = prob.grid, prob.eqn.calcN, prob.timestepper
    gen = torch.Generator(device="cuda").manual_seed(4)
    d = torch.randn(prob.sol.shape, generator=gen, device="cuda", dtype=torch.complex128)
    lam = torch.randn(prob.sol.shape, generator=gen, device="cuda", dtype=torch.complex128)
    ff.dealias(d, g)
    Sd = d.clone()
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("ffhip_op_rk4_run", op.handle, ts.handle, Sd.data_ptr(), prob.clock.dt, nsteps, st)
    STl = torch.zeros_like(lam)
    _lib.call("ffhip_op_rk4_adjoint", op.handle, ts.handle, lam.data_ptr(), STl.data_ptr(), prob.clock.dt, nsteps,
              st)
    lhs = float(torch.sum(Sd.real * lam.real + Sd.imag * lam.imag))
    rhs = float(torch.sum(d.real * STl.real + d.imag * STl.imag))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs), (lhs, rhs)


def _taylor(ff, prob, c0, cobs, grad, nsteps, ndir, hs, seed):
    import torch

    g = prob.grid
    gen = torch.Generator(device="cuda").manual_seed(seed)

    def J(x):
        s = torch.zeros_like(prob.sol)
        ff.mul(s, g.rfftplan, x)
        ff.dealias(s, g)
        prob.sol.copy_(s)
        ff.stepforward(prob, nsteps)
        cT = torch.zeros_like(x)
        ff.ldiv(cT, g.rfftplan, prob.sol)
        return 0.5 * float(torch.sum((cT.double() - cobs.double()) ** 2))

    out = []
    for _ in range(ndir):
        d = torch.randn(c0.shape, generator=gen, device="cuda", dtype=torch.float64).to(c0.dtype)
        gd = float(torch.sum(grad.double() * d.double()))
        out.append((gd, [abs((J(c0 + h * d) - J(c0 - h * d)) / (2 * h) - gd) for h in hs]))
    return out


def test_taylor_fp64_64(ff):
    n, kappa, dt, nsteps = 64, 1e-3, 0.01, 60
    prob, c0, cobs = _setup(ff, n, kappa, dt, nsteps, np.float64)
    J, grad = ff.adjoint.tracer_misfit_gradient(prob, c0, cobs, nsteps)
    for gd, errs in _taylor(ff, prob, c0, cobs, grad, nsteps, 8, [1e-1, 1e-2, 1e-3, 1e-4], seed=9):
        print(f"tracer taylor fp64: <g,d>={gd:.6e} errs={['%.3e' % e for e in errs]}")
        assert max(errs) < 1e-8 * abs(gd) + 1e-9, (gd, errs)


def test_fp32_gradient_256_300_steps(ff):
    """The tracer adjoint workload: 256^3 FP32, 300 RK4 steps (dt = 0.0025,
    kappa = 1e-3), against the FP64 GPU gradient of the same misfit."""
    import torch

    n, kappa, dt, nsteps = 256, 1e-3, 0.0025, 300
    p32, c0, cobs = _setup(ff, n, kappa, dt, nsteps, np.float32)
    J32, g32 = ff.adjoint.tracer_misfit_gradient(p32, c0, cobs, nsteps)
    del p32
    p64, c064, cobs64 = _setup(ff, n, kappa, dt, nsteps, np.float64)
    J64, g64 = ff.adjoint.tracer_misfit_gradient(p64, c064, cobs64, nsteps)
    err = float(torch.linalg.vector_norm(g32.double() - g64) / torch.linalg.vector_norm(g64))
    print(f"tracer 256^3: J fp32 {J32!r} fp64 {J64!r} gradient rel l2 {err:.3e}")
    assert abs(J32 - J64) < 1e-4 * J64, (J32, J64)
    assert err < 1e-4, err


def test_fused_adjoint_dot_product_and_generic(ff, monkeypatch):
    """The fused RK4 tracer adjoint (64^3, power of two) satisfies the
    dot-product identity with the fused forward map and equals the generic
    adjoint of the library (FP64)."""
    import torch

    from fourierflows_hip import _lib

    n, nsteps = 64, 5
    pf = ff.TracerProblem(n=n, kappa=1e-2, dt=0.01, T=np.float64)
    monkeypatch.setenv("FFHIP_DISABLE_FUSED", "1")
    pg = ff.TracerProblem(n=n, kappa=1e-2, dt=0.01, T=np.float64)
    monkeypatch.delenv("FFHIP_DISABLE_FUSED", raising=False)
    assert pf.eqn.calcN.fused and not pg.eqn.calcN.fused
    gen = torch.Generator(device="cuda").manual_seed(11)
    d = torch.randn(pf.sol.shape, generator=gen, device="cuda", dtype=torch.complex128)
    lam = torch.randn(pf.sol.shape, generator=gen, device="cuda", dtype=torch.complex128)
    ff.dealias(d, pf.grid)
    st = torch.cuda.current_stream().cuda_stream
    Sd = d.clone()
    _lib.call("ffhip_op_rk4_run", pf.eqn.calcN.handle, pf.timestepper.handle, Sd.data_ptr(), pf.clock.dt, nsteps, st)
    out = {}
    for name, p in (("fused", pf), ("generic", pg)):
        o = torch.zeros_like(lam)
        _lib.call("ffhip_op_rk4_adjoint", p.eqn.calcN.handle, p.timestepper.handle, lam.data_ptr(), o.data_ptr(),
                  p.clock.dt, nsteps, st)
        out[name] = o
    lhs = float(torch.sum(Sd.real * lam.real + Sd.imag * lam.imag))
    rhs = float(torch.sum(d.real * out["fused"].real + d.imag * out["fused"].imag))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs), (lhs, rhs)
    err = float(torch.linalg.vector_norm(out["fused"] - out["generic"]) / torch.linalg.vector_norm(out["generic"]))
    assert err < 1e-12, err
