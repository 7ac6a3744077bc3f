"""Equation, Clock, Problem and the ETDRK4 / RK4 time steppers.

Mirrors src/problem.jl and src/timesteppers.jl.  The pointwise passes
(ETDRK4substep12!/substep3!/update!, addlinearterm!/substepsol!/RK4update!)
run in libffhip; a built-in operator as ``calcN!`` makes ``stepforward!``
run whole fused steps inside the library without returning to Python.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .core import HIP, cxtype, fltype, prec_of, ptr, stream_ptr, zeros


def _host_L(L, dims):
    """L as FP64 host arrays (re, im or None) broadcast to the Julia dims."""
    try:
        import torch

        if isinstance(L, torch.Tensor):
            L = L.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    a = np.broadcast_to(np.asarray(L), tuple(reversed(tuple(dims))))
    if np.iscomplexobj(a):
        return (np.ascontiguousarray(a.real, dtype=np.float64), np.ascontiguousarray(a.imag, dtype=np.float64))
    return np.ascontiguousarray(a, dtype=np.float64), None


def _supersize(L):
    return tuple(reversed(np.shape(L))) if not np.isscalar(L) else ()


class Equation:
    """``Equation(L, calcN!, grid; dims, T)`` (src/problem.jl:30-34).

    ``dims`` is in Julia order; ``T`` defaults to ``cxtype(eltype(grid))``.
    """

    def __init__(self, L, calcN, grid, dims=None, T=None):
        self.L = L
        self.calcN = calcN
        self.grid = grid
        if dims is None:
            try:
                import torch

                shp = tuple(L.shape) if isinstance(L, (torch.Tensor, np.ndarray)) else np.shape(L)
            except ImportError:  # pragma: no cover
                shp = np.shape(L)
            dims = tuple(reversed(shp))
        self.dims = tuple(int(d) for d in dims)
        self.T = cxtype(grid.T) if T is None else np.dtype(T)


class Clock:
    """src/problem.jl:43-50."""

    def __init__(self, dt, t=0.0, step=0):
        self.dt, self.t, self.step = dt, t, step

    def __repr__(self):
        return f"Clock(dt={self.dt}, step={self.step}, t={self.t})"


class _Handle:
    _destroy = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            getattr(_lib._lib, self._destroy)(h)
            self._h = None


class ETDRK4TimeStepper(_Handle):
    """src/timesteppers.jl:434-463.  Coefficients (zeta, alpha, beta, Gamma,
    expLdt, exp_half_Ldt) are computed in FP64 on the host by libffhip and
    kept in FP64 on the device; the scratch arrays are in the state type."""

    _destroy = "ffhip_etdrk4_destroy"

    def __init__(self, eqn, dt, dev=None):
        dev = eqn.grid.device if dev is None else dev
        self.T = eqn.T
        self.dims = eqn.dims
        Lre, Lim = _host_L(eqn.L, eqn.dims)
        h = C.c_void_p()
        _lib.call("ffhip_etdrk4_create", C.byref(h), _lib.dptr(Lre), _lib.dptr(Lim), len(self.dims),
                  _lib.i64arr(self.dims), float(dt), prec_of(eqn.T))
        self._h = h
        self.dt = float(np.dtype(fltype(eqn.T)).type(dt))  # src/timesteppers.jl:457
        z = lambda: zeros(dev, eqn.T, eqn.dims)  # noqa: E731
        self.sol1, self.sol2 = z(), z()
        self.N1, self.N2, self.N3, self.N4 = z(), z(), z(), z()

    @property
    def handle(self):
        return self._h

    def substep12(self, out, sol, N):
        _lib.call("ffhip_etdrk4_substep12", self._h, ptr(out), ptr(sol), ptr(N), stream_ptr())

    def substep3(self, out, sol1, N1, N3):
        _lib.call("ffhip_etdrk4_substep3", self._h, ptr(out), ptr(sol1), ptr(N1), ptr(N3), stream_ptr())

    def update(self, sol, N1, N2, N3, N4):
        _lib.call("ffhip_etdrk4_update", self._h, ptr(sol), ptr(N1), ptr(N2), ptr(N3), ptr(N4), stream_ptr())

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        """src/timesteppers.jl:518-547."""
        t = clock.t
        eqn.calcN(self.N1, sol, t, clock, vars, params, grid)
        self.substep12(self.sol1, sol, self.N1)
        t2 = clock.t + clock.dt / 2
        eqn.calcN(self.N2, self.sol1, t2, clock, vars, params, grid)
        self.substep12(self.sol2, sol, self.N2)
        eqn.calcN(self.N3, self.sol2, t2, clock, vars, params, grid)
        self.substep3(self.sol2, self.sol1, self.N1, self.N3)
        eqn.calcN(self.N4, self.sol2, clock.t + clock.dt, clock, vars, params, grid)
        self.update(sol, self.N1, self.N2, self.N3, self.N4)
        clock.t += clock.dt
        clock.step += 1

    def run_fused(self, op, sol, clock, nsteps):
        _lib.call("ffhip_op_etdrk4_run", op.handle, self._h, ptr(sol), int(nsteps), stream_ptr())
        clock.t += nsteps * clock.dt
        clock.step += nsteps


class RK4TimeStepper(_Handle):
    """src/timesteppers.jl:180-197 (L kept in FP64 on the device)."""

    _destroy = "ffhip_rk4_destroy"

    def __init__(self, eqn, dt=None, dev=None):
        dev = eqn.grid.device if dev is None else dev
        self.T = eqn.T
        self.dims = eqn.dims
        Lre, Lim = _host_L(eqn.L, eqn.dims)
        h = C.c_void_p()
        _lib.call("ffhip_rk4_create", C.byref(h), _lib.dptr(Lre), _lib.dptr(Lim), len(self.dims),
                  _lib.i64arr(self.dims), prec_of(eqn.T))
        self._h = h
        z = lambda: zeros(dev, eqn.T, eqn.dims)  # noqa: E731
        self.sol1 = z()
        self.RHS1, self.RHS2, self.RHS3, self.RHS4 = z(), z(), z(), z()

    @property
    def handle(self):
        return self._h

    def addlinearterm(self, RHS, sol):
        _lib.call("ffhip_rk4_addlinearterm", self._h, ptr(RHS), ptr(sol), stream_ptr())

    def substepsol(self, newsol, sol, RHS, dt):
        _lib.call("ffhip_rk4_substepsol", self._h, ptr(newsol), ptr(sol), ptr(RHS), float(dt), stream_ptr())

    def update(self, sol, dt):
        _lib.call("ffhip_rk4_update", self._h, ptr(sol), ptr(self.RHS1), ptr(self.RHS2), ptr(self.RHS3),
                  ptr(self.RHS4), float(dt), stream_ptr())

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        """src/timesteppers.jl:237-274."""
        t, dt = clock.t, clock.dt
        eqn.calcN(self.RHS1, sol, t, clock, vars, params, grid)
        self.addlinearterm(self.RHS1, sol)
        self.substepsol(self.sol1, sol, self.RHS1, dt / 2)
        eqn.calcN(self.RHS2, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.addlinearterm(self.RHS2, self.sol1)
        self.substepsol(self.sol1, sol, self.RHS2, dt / 2)
        eqn.calcN(self.RHS3, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.addlinearterm(self.RHS3, self.sol1)
        self.substepsol(self.sol1, sol, self.RHS3, dt)
        eqn.calcN(self.RHS4, self.sol1, t + dt, clock, vars, params, grid)
        self.addlinearterm(self.RHS4, self.sol1)
        self.update(sol, dt)
        clock.t += clock.dt
        clock.step += 1

    def run_fused(self, op, sol, clock, nsteps):
        _lib.call("ffhip_op_rk4_run", op.handle, self._h, ptr(sol), float(clock.dt), int(nsteps), stream_ptr())
        clock.t += nsteps * clock.dt
        clock.step += nsteps


_STEPPERS = {"ETDRK4": ETDRK4TimeStepper, "RK4": RK4TimeStepper}
fullyexplicitsteppers = ["RK4"]


def isexplicit(stepper):
    return stepper in fullyexplicitsteppers


def TimeStepper(stepper, equation, dt=None, dev=None):
    """src/timesteppers.jl:57-69 (string -> stepper)."""
    try:
        cls = _STEPPERS[stepper]
    except KeyError:
        raise ValueError(f"time stepper {stepper!r} is not part of this build "
                         f"(available: {sorted(_STEPPERS)})") from None
    return cls(equation, dt, dev)


class EmptyParams:
    pass


class EmptyVars:
    pass


class Problem:
    """``Problem(eqn, stepper, dt, grid, vars, params)`` (src/problem.jl:99-111)."""

    def __init__(self, eqn, stepper, dt, grid, vars=None, params=None):
        dev = grid.device
        if not isinstance(dev, HIP):
            raise RuntimeError("libffhip problems run on HIP devices")
        self.eqn, self.grid = eqn, grid
        self.vars = EmptyVars() if vars is None else vars
        self.params = EmptyParams() if params is None else params
        self.clock = Clock(float(np.dtype(fltype(grid.T)).type(dt)))
        self.timestepper = TimeStepper(stepper, eqn, dt, dev)
        self.stepper = stepper
        self.sol = zeros(dev, eqn.T, eqn.dims)


def stepforward(prob, nsteps=1):
    """``stepforward!(prob, nsteps)`` (src/timesteppers.jl:6-20).

    With a built-in operator as ``calcN!`` all ``nsteps`` run inside
    libffhip (fused kernels, no host round trips); otherwise each step calls
    the Python ``calcN!`` between library pointwise passes.
    """
    op = getattr(prob.eqn.calcN, "builtin_operator", None)
    if op is not None and nsteps > 0:
        prob.timestepper.run_fused(op, prob.sol, prob.clock, nsteps)
        return None
    ts = prob.timestepper
    for _ in range(nsteps):
        ts.stepforward(prob.sol, prob.clock, prob.eqn, prob.vars, prob.params, prob.grid)
    return None


def step_until(prob, stop_time):
    """src/timesteppers.jl:734-760 (rejects ETDRK4, :736-737)."""
    if isinstance(prob.timestepper, ETDRK4TimeStepper):
        raise RuntimeError("step_until! requires fully explicit time stepper; does not work with ETDRK4")
    if not stop_time > prob.clock.t:
        raise ValueError("stop_time must be greater than prob.clock.t")
    dt = prob.clock.dt
    interval = stop_time - prob.clock.t
    nsteps = int(np.floor(interval / dt))
    stepforward(prob, nsteps)
    t_remaining = interval - prob.clock.t  # as in the reference (:754)
    prob.clock.dt = t_remaining
    stepforward(prob)
    prob.clock.dt = dt
