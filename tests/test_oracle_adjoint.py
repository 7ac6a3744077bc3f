"""Pins for the oracle's discrete adjoints.

* dot-product identities <A x, y> = <x, A^T y> for ldiv!, mul!, the
  vorticity N tangent, and the tracer N
* Taylor tests: central finite differences of J converge to <grad J, d>
  at second order, and the adjoint gradient matches them to FD accuracy
* checkpoint-and-recompute gives the same gradient as storing every step
"""

import numpy as np
import pytest

from oracle import adjoint as oa
from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph


def _rand_spec(g, rng):
    return rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)


@pytest.mark.parametrize("shape", [(16, 12), (32, 32), (6, 8, 10)])
def test_transform_transposes(shape):
    rng = np.random.default_rng(0)
    if len(shape) == 2:
        g = og.TwoDGrid(shape[1], 2 * np.pi, shape[0], 3.0)
    else:
        g = og.ThreeDGrid(shape[2], 2 * np.pi, shape[1], 3.0, shape[0], 1.0)
    X = _rand_spec(g, rng)  # non-Hermitian on purpose: Im at kx=0, nx/2 is ignored by ldiv!
    phi = rng.standard_normal(g.phys_shape)
    lhs = float(np.sum(g.ldiv(X) * phi))
    assert abs(lhs - oa.sdot(X, oa.ldiv_T(phi, g))) < 1e-12 * max(1, abs(lhs))
    Y = _rand_spec(g, rng)
    lhs2 = oa.sdot(g.mul(phi), Y)
    assert abs(lhs2 - float(np.sum(phi * oa.mul_T(Y, g)))) < 1e-12 * max(1, abs(lhs2))


def test_vorticity_N_tangent_and_adjoint():
    g = og.TwoDGrid(32, 2 * np.pi, x0=0, y0=0)
    ops = oa.VorticityOps(g)
    rng = np.random.default_rng(1)
    s = oics.vorticity_ic(g, seed=3, k0=4.0)
    d = og.dealias(_rand_spec(g, rng), g)
    flds = ops.fields(s)
    # tangent vs finite difference of N (N is quadratic: central FD is exact up to roundoff)
    h = 1e-3
    fd = (ops.N(s + h * d) - ops.N(s - h * d)) / (2 * h)
    tan = ops.N_tangent(flds, d)
    assert np.linalg.norm(fd - tan) < 1e-10 * np.linalg.norm(tan)
    # the oracle's production calcN agrees with ops.N
    N = np.zeros_like(s)
    oph.vorticity_calcN(N, s, 0, None, None, None, g)
    assert np.linalg.norm(N - ops.N(s)) < 1e-13 * np.linalg.norm(N)
    # adjoint identity
    lam = _rand_spec(g, rng)
    lhs = oa.sdot(tan, lam)
    rhs = oa.sdot(d, ops.N_adjoint(flds, lam))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def _vort_problem(n=32, nu=1e-2, dt=0.01):
    g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    L = oph.vorticity_linear(g, nu=nu)
    return g, oa.ETDRK4Adjoint(g, L, dt)


def test_etdrk4_adjoint_step_identity():
    g, adj = _vort_problem()
    rng = np.random.default_rng(2)
    u = oics.vorticity_ic(g, seed=4, k0=4.0)
    d = og.dealias(_rand_spec(g, rng), g)
    lam = _rand_spec(g, rng)
    _, stages = adj.step(u)
    h = 1e-6
    tan = (adj.step(u + h * d)[0] - adj.step(u - h * d)[0]) / (2 * h)
    lhs = oa.sdot(tan, lam)
    rhs = oa.sdot(d, adj.step_adjoint(stages, lam))
    assert abs(lhs - rhs) < 1e-7 * abs(lhs)


def test_vorticity_taylor_and_checkpointing():
    """Central-FD Taylor test in 4 seeded directions: error ~ h² and small."""
    g, adj = _vort_problem(n=32)
    z0 = g.ldiv(oics.vorticity_ic(g, seed=1, k0=4.0))
    zobs = g.ldiv(adj.forward(g.ldiv(oics.vorticity_ic(g, seed=2, k0=4.0)), 30))
    nsteps = 30
    J, grad = adj.gradient(z0, zobs, nsteps, ncheck=4)
    J2, grad2 = adj.gradient(z0, zobs, nsteps, ncheck=nsteps)
    assert abs(J - J2) < 1e-14 * J and np.allclose(grad, grad2, rtol=0, atol=1e-13 * np.abs(grad).max())
    rng = np.random.default_rng(5)
    f = lambda x: adj.misfit(x, zobs, nsteps)  # noqa: E731
    hs = [1e-1, 1e-2, 1e-3]
    for _ in range(4):
        d = rng.standard_normal(z0.shape)
        gd, errs = oa.taylor_errors(f, grad, z0, d, hs)
        assert errs[1] < 0.02 * errs[0] and errs[2] < 1e-6 * abs(gd) + 0.02 * errs[1], (gd, errs)


def test_tracer_adjoint_identity_and_taylor():
    g = og.ThreeDGrid(12, 2 * np.pi, x0=0, y0=0, z0=0)
    params = oph.TracerParams(kappa=1e-2)
    vel = oph.abc_velocity(g, params)
    adj = oa.RK4TracerAdjoint(g, vel, oph.tracer_linear(g, params.kappa), 0.05)
    rng = np.random.default_rng(3)
    lam = _rand_spec(g, rng)
    d = og.dealias(_rand_spec(g, rng), g)
    assert abs(oa.sdot(adj.ops.N(d), lam) - oa.sdot(d, adj.ops.N_adjoint(lam))) < 1e-11 * np.linalg.norm(lam) ** 2
    # linear problem: the Taylor remainder is pure roundoff
    c0 = oics.tracer_ic(g, seed=1, ngauss=3, sigma=0.7)
    cobs = oics.tracer_ic(g, seed=2, ngauss=3, sigma=0.7)
    J, grad = adj.gradient(c0, cobs, 10)
    for _ in range(3):
        dd = rng.standard_normal(c0.shape)
        gd, errs = oa.taylor_errors(lambda x: adj.misfit(x, cobs, 10), grad, c0, dd, [1e-1, 1e-2])
        assert max(errs) < 1e-8 * abs(gd) + 1e-12, (gd, errs)
