"""GPU parity of the explicit and filtered steppers (src/timesteppers.jl).

* the reference's own timestepper tests (test/test_timesteppers.jl,
  test/runtests.jl:215-227) for all ten steppers: Gaussian diffusion with
  constant and varying kappa to rtol = step * 1e-12, and step_until! for the
  fully explicit ones;
* the built-in vorticity operator stepped by each stepper (library drivers,
  fused for FilteredETDRK4) against the FP64 oracle;
* the filtered fused ETDRK4 adjoint against central differences of the
  forward map.
"""

import numpy as np
import pytest

from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph
from oracle import timesteppers as ots
from oracle import utils as ou

pytestmark = pytest.mark.gpu

STEPPERS = ["ForwardEuler", "FilteredForwardEuler", "AB3", "FilteredAB3", "RK4", "FilteredRK4", "LSRK54",
            "FilteredLSRK54", "ETDRK4", "FilteredETDRK4"]


def _dev(x):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x)).cuda()


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _gaussian(x, t, c0=0.01, sigma=0.2, kappa=1e-2):
    return c0 * sigma / np.sqrt(sigma**2 + 2 * kappa * t) * np.exp(-(x**2) / (2 * (sigma**2 + 2 * kappa * t)))


@pytest.mark.parametrize("varying", [False, True])
@pytest.mark.parametrize("stepper", STEPPERS)
def test_diffusion_all_steppers(ff, stepper, varying):
    kappa = 1e-2
    dt = 1e-9 / kappa
    kap = kappa * np.ones(128) if varying else kappa
    prob = ff.Diffusion.Problem(nx=128, Lx=2 * np.pi, kappa=kap, dt=dt, stepper=stepper)
    ff.Diffusion.set_c(prob, _gaussian(prob.grid.x, 0))
    ff.stepforward(prob, 1000)
    ff.Diffusion.updatevars(prob)
    c_final = _gaussian(prob.grid.x, 1000 * dt)
    assert ou.isapprox(c_final, _host(prob.vars.c), rtol=prob.clock.step * 1e-12)


@pytest.mark.parametrize("stepper", [s for s in STEPPERS if "ETDRK4" not in s])
def test_step_until_explicit(ff, stepper):
    kappa, dt = 1e-2, 1e-7
    prob = ff.Diffusion.Problem(nx=128, kappa=kappa, dt=dt, stepper=stepper)
    t_final = 1000 * dt + 1e-6 / np.pi
    ff.Diffusion.set_c(prob, _gaussian(prob.grid.x, 0))
    ff.step_until(prob, t_final)
    ff.Diffusion.updatevars(prob)
    assert ou.isapprox(_gaussian(prob.grid.x, t_final), _host(prob.vars.c), rtol=prob.clock.step * 1e-12)


def test_isexplicit_list(ff):
    assert [s for s in STEPPERS if ff.isexplicit(s)] == [s for s in STEPPERS if "ETDRK4" not in s]


@pytest.mark.parametrize("stepper", STEPPERS)
def test_vorticity_builtin_all_steppers(ff, stepper):
    n, nu, dt, steps = 64, 2e-3, 0.004, 12
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=21, k0=5.0)
    p = ots.Problem(ots.Equation(oph.vorticity_linear(o, nu=nu), oph.vorticity_calcN, o), stepper, dt, o)
    p.sol[...] = zh0
    p.stepforward(steps)
    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, stepper=stepper)
    prob.sol.copy_(_dev(zh0))
    ff.stepforward(prob, 5)
    ff.stepforward(prob, steps - 5)  # AB3 continues from the stored history
    err = ou.rel_l2(_host(prob.sol), p.sol)
    assert err < 1e-12, (stepper, err)


def test_filtered_etdrk4_fused_fp32(ff):
    n, nu, dt, steps = 128, 1e-3, 0.004, 20
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=22, k0=8.0)
    p = ots.Problem(ots.Equation(oph.vorticity_linear(o, nu=nu), oph.vorticity_calcN, o), "FilteredETDRK4", dt, o)
    p.sol[...] = zh0
    p.stepforward(steps)
    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, stepper="FilteredETDRK4", T=np.float32)
    assert prob.eqn.calcN.fused
    prob.sol.copy_(_dev(zh0.astype(np.complex64)))
    ff.stepforward(prob, steps)
    err = ou.rel_l2(_host(prob.sol).astype(np.complex128), p.sol)
    assert err < 2e-6, err


def test_filtered_etdrk4_adjoint_dot_product(ff):
    """<(S(u + h d) - S(u - h d)) / 2h, lam> = <d, S^T lam> for the filtered
    fused step map S over 6 steps (FP64)."""
    import torch

    from fourierflows_hip import _lib

    n, nu, dt, nsteps = 64, 5e-3, 0.01, 6
    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, stepper="FilteredETDRK4")
    op, ts, g = prob.eqn.calcN, prob.timestepper, prob.grid
    u = ff.vorticity_ic(g, seed=3, k0=5.0)
    gen = torch.Generator(device="cuda").manual_seed(9)
    d = torch.randn(u.shape, generator=gen, device="cuda", dtype=torch.complex128)
    lam = torch.randn(u.shape, generator=gen, device="cuda", dtype=torch.complex128)
    ff.dealias(d, g)
    st = torch.cuda.current_stream().cuda_stream

    def S(x):
        y = x.clone()
        _lib.call("ffhip_op_etdrk4_run", op.handle, ts.handle, y.data_ptr(), nsteps, st)
        return y

    h = 1e-6
    tan = (S(u + h * d) - S(u - h * d)) / (2 * h)
    lhs = float(torch.sum(tan.real * lam.real + tan.imag * lam.imag))
    y = u.clone()
    _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, y.data_ptr(), y.data_ptr(), nsteps, 2, st)
    lt = lam.clone()
    _lib.call("ffhip_op_etdrk4_adjoint", op.handle, ts.handle, lt.data_ptr(), lt.data_ptr(), nsteps, 2, st)
    rhs = float(torch.sum(d.real * lt.real + d.imag * lt.imag))
    assert abs(lhs - rhs) < 1e-7 * abs(lhs), (lhs, rhs)
