"""GPU parity of the built-in 3D tracer operator (ABC flow) and RK4 runs."""

import numpy as np
import pytest

from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph
from oracle import timesteppers as ots
from oracle import utils as ou

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x)).cuda()


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("n", [16, 32, 30])
@pytest.mark.parametrize("prec", [64, 32])
def test_tracer_calcN_matches_oracle(ff, n, prec):
    T = np.float64 if prec == 64 else np.float32
    prob = ff.TracerProblem(n=n, kappa=1e-2, T=T)
    o = og.ThreeDGrid(n, 2 * np.pi, x0=0, y0=0, z0=0)
    params = oph.TracerParams(kappa=1e-2)
    vars = oph.TracerVars(o, params)
    c0 = og.dealias(o.mul(oics.tracer_ic(o, seed=3, ngauss=4, sigma=0.6)), o)
    ref = np.zeros_like(c0)
    oph.tracer_calcN(ref, c0, 0.0, None, vars, params, o)
    CT = np.complex128 if prec == 64 else np.complex64
    N = ff.zeros(prob.grid.device, CT, prob.grid.spec_dims)
    prob.eqn.calcN(N, _dev(c0.astype(CT)), 0.0)
    tol = 1e-12 if prec == 64 else 2e-5
    assert ou.rel_l2(_host(N), ref) < tol


def test_tracer_closed_form(ff):
    """c = sin x: N = -u·∇c = -(A sin z + C cos y) cos x."""
    prob = ff.TracerProblem(n=16, A=1.0, B=0.5, C=0.25)
    o = og.ThreeDGrid(16, 2 * np.pi, x0=0, y0=0, z0=0)
    x, y, z = o.x[None, None, :], o.y[None, :, None], o.z[:, None, None]
    c = np.sin(x) + 0 * y + 0 * z
    N = ff.zeros(prob.grid.device, np.complex128, prob.grid.spec_dims)
    prob.eqn.calcN(N, _dev(o.mul(c)), 0.0)
    expect = -(1.0 * np.sin(z) + 0.25 * np.cos(y)) * np.cos(x)
    assert ou.isapprox(o.ldiv(_host(N)), expect, rtol=1e-12)


def test_tracer_rk4_run_matches_oracle(ff):
    n, dt, steps = 32, 0.01, 20
    prob = ff.TracerProblem(n=n, kappa=1e-3, dt=dt, T=np.float64)
    o = og.ThreeDGrid(n, 2 * np.pi, x0=0, y0=0, z0=0)
    params = oph.TracerParams(kappa=1e-3)
    vars = oph.TracerVars(o, params)
    c0 = og.dealias(o.mul(oics.tracer_ic(o, seed=1, ngauss=16, sigma=0.3)), o)
    eqn = ots.Equation(oph.tracer_linear(o, 1e-3), oph.tracer_calcN, o)
    p = ots.Problem(eqn, "RK4", dt, o, vars, params)
    p.sol[...] = c0
    p.stepforward(steps)
    prob.sol.copy_(_dev(c0))
    ff.stepforward(prob, steps)
    assert ou.rel_l2(_host(prob.sol), p.sol) < 1e-12
    var, mean = prob.eqn.calcN.variance_mean(prob.sol)
    vref = 0.5 * ou.parsevalsum2(p.sol, o) / (2 * np.pi) ** 3
    assert abs(var - vref) <= 1e-12 * vref
    assert abs(mean - p.sol.flat[0].real / n**3) <= 1e-12 * abs(mean)


def _tracer_pair(ff, monkeypatch, n, T, **kw):
    pf = ff.TracerProblem(n=n, T=T, **kw)
    monkeypatch.setenv("FFHIP_DISABLE_FUSED", "1")
    pg = ff.TracerProblem(n=n, T=T, **kw)
    monkeypatch.delenv("FFHIP_DISABLE_FUSED", raising=False)
    assert pf.eqn.calcN.fused and not pg.eqn.calcN.fused
    return pf, pg


def test_fused_tracer_fp64_matches_oracle(ff):
    """The fused RK4 tracer path (64^3, kept band only, 4 kernels per stage)
    against the FP64 oracle; the IC has nonzero aliased modes, which take
    RK4 steps of du/dt = L u outside the band."""
    n, kappa, dt, steps = 64, 2e-3, 0.01, 8
    prob = ff.TracerProblem(n=n, kappa=kappa, dt=dt, T=np.float64)
    assert prob.eqn.calcN.fused
    o = og.ThreeDGrid(n, 2 * np.pi, x0=0, y0=0, z0=0)
    params = oph.TracerParams(kappa=kappa)
    rng = np.random.default_rng(4)
    c0 = o.mul(oics.tracer_ic(o, seed=5, ngauss=6, sigma=0.5))
    c0 += 1e-3 * (rng.standard_normal(c0.shape) + 1j * rng.standard_normal(c0.shape))
    c0[..., 0] = c0[..., 0].real
    p = ots.Problem(ots.Equation(oph.tracer_linear(o, kappa), oph.tracer_calcN, o), "RK4", dt, o,
                    oph.TracerVars(o, params), params)
    p.sol[...] = c0
    p.stepforward(steps)
    prob.sol.copy_(_dev(c0))
    ff.stepforward(prob, 3)
    ff.stepforward(prob, steps - 3)
    err = ou.rel_l2(_host(prob.sol), p.sol)
    assert err < 1e-12, err


@pytest.mark.parametrize("n,T,tol", [(128, np.float64, 1e-12), (256, np.float32, 2e-6)])
def test_fused_tracer_matches_generic(ff, monkeypatch, n, T, tol):
    pf, pg = _tracer_pair(ff, monkeypatch, n, T, kappa=1e-3, dt=0.0025)
    g = pf.grid
    import torch

    s = torch.zeros_like(pf.sol)
    ff.mul(s, g.rfftplan, ff.tracer_ic(g, seed=2, ngauss=16, sigma=0.3))
    ff.dealias(s, g)
    for p in (pf, pg):
        p.sol.copy_(s)
        ff.stepforward(p, 10)
    a, b = _host(pf.sol).astype(np.complex128), _host(pg.sol).astype(np.complex128)
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert err < tol, err
