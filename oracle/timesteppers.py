"""ETDRK4 / RK4 time stepping and the Problem model of the reference.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The operation order of every pointwise update follows the reference's
broadcasts exactly, so that an FP64 implementation following the same order
rounds the same way.
"""

from __future__ import annotations

import numpy as np


def getexpLs(dt, L):
    """src/timesteppers.jl:673-678: ``exp(dt*L)`` and ``exp(dt*L/2)``."""
    L = np.asarray(L)
    return np.exp(dt * L), np.exp(dt * L / 2)


def getetdcoeffs(dt, L, ncirc=32, rcirc=1):
    """src/timesteppers.jl:689-718: Kassam-Trefethen contour means in complex128.

    ``z = dt*L + rcirc*exp(2*pi*i*(j-1/2)/ncirc)`` for j = 1..ncirc, then

    * zeta  = dt * mean((exp(z/2) - 1) / z)
    * alpha = dt * mean((-4 - z + exp(z)(4 - 3z + z^2)) / z^3)
    * beta  = dt * mean((2 + z + exp(z)(-2 + z)) / z^3)
    * Gamma = dt * mean((-4 - 3z - z^2 + exp(z)(4 - z)) / z^3)

    and the real parts are taken when L is real (:710-715).
    """
    L = np.asarray(L)
    circ = rcirc * np.exp(2j * np.pi / ncirc * (np.arange(ncirc) + 0.5))
    zc = (dt * L)[..., None].astype(np.complex128) + circ
    ez = np.exp(zc)
    zc3 = zc**3
    zeta = (np.exp(zc / 2) - 1) / zc
    alpha = (-4 - zc + ez * (4 - 3 * zc + zc**2)) / zc3
    beta = (2 + zc + ez * (-2 + zc)) / zc3
    gamma = (-4 - 3 * zc - zc**2 + ez * (4 - zc)) / zc3
    out = [dt * np.mean(c, axis=-1) for c in (zeta, alpha, beta, gamma)]
    if not np.iscomplexobj(L):
        out = [np.real(c) for c in out]
    return tuple(out)


class Clock:
    """src/problem.jl:43-50."""

    def __init__(self, dt, t=0.0, step=0):
        self.dt, self.t, self.step = dt, t, step


class Equation:
    """src/problem.jl:11-34: ``Equation(L, calcN!, grid; dims, T)``."""

    def __init__(self, L, calcN, grid, dims=None, T=None):
        self.L = L
        self.calcN = calcN
        self.grid = grid
        self.dims = np.shape(L) if dims is None else tuple(dims)
        self.T = np.result_type(grid.T, np.complex64) if T is None else np.dtype(T)


class ETDRK4TimeStepper:
    """src/timesteppers.jl:434-463 (construction) and :501-547 (stepping)."""

    def __init__(self, eqn, dt):
        L = np.broadcast_to(np.asarray(eqn.L), eqn.dims)
        self.expLdt, self.exp_half_Ldt = getexpLs(dt, L)
        self.zeta, self.alpha, self.beta, self.gamma = getetdcoeffs(dt, L)
        z = lambda: np.zeros(eqn.dims, dtype=eqn.T)  # noqa: E731
        self.sol1, self.sol2 = z(), z()
        self.N1, self.N2, self.N3, self.N4 = z(), z(), z(), z()

    def substeps(self, sol, clock, eqn, vars, params, grid):
        """ETDRK4substeps! (src/timesteppers.jl:518-537)."""
        eqn.calcN(self.N1, sol, clock.t, clock, vars, params, grid)
        # ETDRK4substep12! (:506-510)
        self.sol1[...] = self.exp_half_Ldt * sol + self.zeta * self.N1
        t2 = clock.t + clock.dt / 2
        eqn.calcN(self.N2, self.sol1, t2, clock, vars, params, grid)
        self.sol2[...] = self.exp_half_Ldt * sol + self.zeta * self.N2
        eqn.calcN(self.N3, self.sol2, t2, clock, vars, params, grid)
        # ETDRK4substep3! (:512-516)
        self.sol2[...] = self.exp_half_Ldt * self.sol1 + self.zeta * (2 * self.N3 - self.N1)
        t3 = clock.t + clock.dt
        eqn.calcN(self.N4, self.sol2, t3, clock, vars, params, grid)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        """src/timesteppers.jl:539-547 with ETDRK4update! (:501-504)."""
        self.substeps(sol, clock, eqn, vars, params, grid)
        sol[...] = (self.expLdt * sol + self.alpha * self.N1
                    + 2 * self.beta * (self.N2 + self.N3) + self.gamma * self.N4)
        clock.t += clock.dt
        clock.step += 1


class RK4TimeStepper:
    """src/timesteppers.jl:180-197 (construction) and :225-274 (stepping)."""

    def __init__(self, eqn, dt=None):
        z = lambda: np.zeros(eqn.dims, dtype=eqn.T)  # noqa: E731
        self.sol1 = z()
        self.RHS1, self.RHS2, self.RHS3, self.RHS4 = z(), z(), z(), z()
        self.L = eqn.L

    def substeps(self, sol, clock, eqn, vars, params, grid, t, dt):
        """RK4substeps! (:237-258) with addlinearterm! (:225-229) and substepsol! (:231-235)."""
        L = self.L
        eqn.calcN(self.RHS1, sol, t, clock, vars, params, grid)
        self.RHS1 += L * sol
        self.sol1[...] = sol + dt / 2 * self.RHS1
        eqn.calcN(self.RHS2, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.RHS2 += L * self.sol1
        self.sol1[...] = sol + dt / 2 * self.RHS2
        eqn.calcN(self.RHS3, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.RHS3 += L * self.sol1
        self.sol1[...] = sol + dt * self.RHS3
        eqn.calcN(self.RHS4, self.sol1, t + dt, clock, vars, params, grid)
        self.RHS4 += L * self.sol1

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        """src/timesteppers.jl:266-274 with RK4update! (:260-264)."""
        dt = clock.dt
        self.substeps(sol, clock, eqn, vars, params, grid, clock.t, dt)
        sol += dt / 6 * (self.RHS1 + 2 * self.RHS2 + 2 * self.RHS3 + self.RHS4)
        clock.t += clock.dt
        clock.step += 1


def _filter_for(eqn, filterkw):
    """makefilter(equation; kw...) (src/domains.jl:542): realvars iff the
    equation's leading Julia dimension is the grid's nkr."""
    from .grids import makefilter

    g = eqn.grid
    realvars = eqn.dims[-1] == g.nkr
    return makefilter(g, realvars=realvars, dims=eqn.dims, **filterkw)


class FilteredETDRK4TimeStepper(ETDRK4TimeStepper):
    """src/timesteppers.jl:471-499, 549-558: ETDRK4 then sol *= filter."""

    def __init__(self, eqn, dt, **filterkw):
        super().__init__(eqn, dt)
        self.filter = _filter_for(eqn, filterkw)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        self.substeps(sol, clock, eqn, vars, params, grid)
        sol[...] = (self.expLdt * sol + self.alpha * self.N1
                    + 2 * self.beta * (self.N2 + self.N3) + self.gamma * self.N4)
        sol *= self.filter
        clock.t += clock.dt
        clock.step += 1


class FilteredRK4TimeStepper(RK4TimeStepper):
    """src/timesteppers.jl:200-223, 276-285: RK4 then sol *= filter."""

    def __init__(self, eqn, dt=None, **filterkw):
        super().__init__(eqn, dt)
        self.filter = _filter_for(eqn, filterkw)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        dt = clock.dt
        self.substeps(sol, clock, eqn, vars, params, grid, clock.t, dt)
        sol += dt / 6 * (self.RHS1 + 2 * self.RHS2 + 2 * self.RHS3 + self.RHS4)
        sol *= self.filter
        clock.t += clock.dt
        clock.step += 1


class ForwardEulerTimeStepper:
    """src/timesteppers.jl:98-118: sol += dt (L sol + N)."""

    def __init__(self, eqn, dt=None):
        self.N = np.zeros(eqn.dims, dtype=eqn.T)
        self.filter = None

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        eqn.calcN(self.N, sol, clock.t, clock, vars, params, grid)
        if self.filter is None:
            sol += clock.dt * (eqn.L * sol + self.N)
        else:  # :142-149
            sol[...] = self.filter * (sol + clock.dt * (self.N + eqn.L * sol))
        clock.t += clock.dt
        clock.step += 1


class FilteredForwardEulerTimeStepper(ForwardEulerTimeStepper):
    """src/timesteppers.jl:126-149."""

    def __init__(self, eqn, dt=None, **filterkw):
        super().__init__(eqn, dt)
        self.filter = _filter_for(eqn, filterkw)


# src/timesteppers.jl:564-566
AB3H1, AB3H2, AB3H3 = 23 / 12, 16 / 12, 5 / 12


class AB3TimeStepper:
    """src/timesteppers.jl:588-650: forward Euler for the first two steps, then
    sol += dt (23/12 RHS - 16/12 RHS_-1 + 5/12 RHS_-2)."""

    def __init__(self, eqn, dt=None):
        z = lambda: np.zeros(eqn.dims, dtype=eqn.T)  # noqa: E731
        self.RHS, self.RHSm1, self.RHSm2 = z(), z(), z()
        self.filter = None

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        eqn.calcN(self.RHS, sol, clock.t, clock, vars, params, grid)
        self.RHS += eqn.L * sol
        if clock.step < 3:  # AB3update! (:630-638)
            sol += clock.dt * self.RHS
        else:
            sol += clock.dt * (AB3H1 * self.RHS - AB3H2 * self.RHSm1 + AB3H3 * self.RHSm2)
        if self.filter is not None:
            sol *= self.filter
        clock.t += clock.dt
        clock.step += 1
        self.RHSm2[...] = self.RHSm1
        self.RHSm1[...] = self.RHS


class FilteredAB3TimeStepper(AB3TimeStepper):
    """src/timesteppers.jl:609-628, 653-666."""

    def __init__(self, eqn, dt=None, **filterkw):
        super().__init__(eqn, dt)
        self.filter = _filter_for(eqn, filterkw)


# src/timesteppers.jl:330-345 (Carpenter & Kennedy 1994, 5-stage 4th order, 2N storage)
LSRK54_A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
            -3550918686646 / 2091501179385, -1275806237668 / 842570457699)
LSRK54_B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357, 1720146321549 / 2090206949498,
            3134564353537 / 4481467310338, 2277821191437 / 14882151754819)
LSRK54_C = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
            2006345519317 / 3224310063776, 2802321613138 / 2924317926251)


class LSRK54TimeStepper:
    """src/timesteppers.jl:318-414: for i = 1..5: RHS = N(sol, t + C_i dt) + L sol;
    S2 = A_i S2 + dt RHS; sol += B_i S2."""

    def __init__(self, eqn, dt=None):
        self.S2 = np.zeros(eqn.dims, dtype=eqn.T)
        self.RHS = np.zeros(eqn.dims, dtype=eqn.T)
        self.filter = None

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        t, dt = clock.t, clock.dt
        self.S2[...] = 0
        for i in range(5):
            eqn.calcN(self.RHS, sol, t + LSRK54_C[i] * dt, clock, vars, params, grid)
            self.RHS += eqn.L * sol
            self.S2[...] = LSRK54_A[i] * self.S2 + dt * self.RHS
            sol += LSRK54_B[i] * self.S2
        if self.filter is not None:
            sol *= self.filter
        clock.t += clock.dt
        clock.step += 1


class FilteredLSRK54TimeStepper(LSRK54TimeStepper):
    """src/timesteppers.jl:365-387, 405-414."""

    def __init__(self, eqn, dt=None, **filterkw):
        super().__init__(eqn, dt)
        self.filter = _filter_for(eqn, filterkw)


_STEPPERS = {"ETDRK4": ETDRK4TimeStepper, "RK4": RK4TimeStepper, "ForwardEuler": ForwardEulerTimeStepper,
             "AB3": AB3TimeStepper, "LSRK54": LSRK54TimeStepper, "FilteredETDRK4": FilteredETDRK4TimeStepper,
             "FilteredRK4": FilteredRK4TimeStepper, "FilteredForwardEuler": FilteredForwardEulerTimeStepper,
             "FilteredAB3": FilteredAB3TimeStepper, "FilteredLSRK54": FilteredLSRK54TimeStepper}

# src/timesteppers.jl:672-681
FULLY_EXPLICIT = ("ForwardEuler", "RK4", "AB3", "LSRK54", "FilteredForwardEuler", "FilteredRK4", "FilteredAB3",
                  "FilteredLSRK54")


def isexplicit(stepper):
    return stepper in FULLY_EXPLICIT


def TimeStepper(stepper, eqn, dt, **kw):
    """src/timesteppers.jl:690-702 (string -> stepper; keywords go to the filter)."""
    try:
        cls = _STEPPERS[stepper]
    except KeyError:
        raise ValueError(f"unknown time stepper {stepper!r}") from None
    return cls(eqn, dt, **kw)


class Problem:
    """src/problem.jl:99-111."""

    def __init__(self, eqn, stepper, dt, grid, vars=None, params=None, **stepperkw):
        self.eqn, self.grid, self.vars, self.params = eqn, grid, vars, params
        self.clock = Clock(dt)
        self.timestepper = TimeStepper(stepper, eqn, dt, **stepperkw)
        self.sol = np.zeros(eqn.dims, dtype=eqn.T)

    def stepforward(self, nsteps=1):
        """src/timesteppers.jl:6-20."""
        ts = self.timestepper
        for _ in range(nsteps):
            ts.stepforward(self.sol, self.clock, self.eqn, self.vars, self.params, self.grid)


def step_until(prob, stop_time):
    """src/timesteppers.jl:734-760 (rejects ETDRK4 at :736-737)."""
    if isinstance(prob.timestepper, (ETDRK4TimeStepper,)):
        raise RuntimeError("step_until! requires fully explicit time stepper; does not work with ETDRK4")
    if not stop_time > prob.clock.t:
        raise ValueError("stop_time must be greater than prob.clock.t")
    dt = prob.clock.dt
    interval = stop_time - prob.clock.t
    nsteps = int(np.floor(interval / dt))
    prob.stepforward(nsteps)
    # the reference subtracts the current time from the interval here (:754)
    t_remaining = interval - prob.clock.t
    prob.clock.dt = t_remaining
    prob.stepforward(1)
    prob.clock.dt = dt
