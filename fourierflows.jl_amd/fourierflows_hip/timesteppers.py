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
    """src/problem.jl:43-50: ``Clock{T}`` -- ``dt`` and ``t`` are scalars of the
    grid's float type ``T`` and every step adds ``dt`` to ``t`` once in that
    type (src/timesteppers.jl:542-545), also when the library takes many steps
    in one call."""

    def __init__(self, dt, t=0.0, step=0, T=np.float64):
        T = np.dtype(T).type
        self.dt, self.t, self.step = T(dt), T(t), step

    def advance(self, nsteps):
        """``clock.t += clock.dt; clock.step += 1``, ``nsteps`` times."""
        t, dt = self.t, self.dt
        for _ in range(int(nsteps)):
            t = t + dt
        self.t = t
        self.step += int(nsteps)

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
        clock.advance(1)

    def run_fused(self, op, sol, clock, nsteps):
        _lib.call("ffhip_op_etdrk4_run", op.handle, self._h, ptr(sol), int(nsteps), stream_ptr())
        clock.advance(nsteps)


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
        clock.advance(1)

    def run_fused(self, op, sol, clock, nsteps):
        _lib.call("ffhip_op_rk4_run", op.handle, self._h, ptr(sol), float(clock.dt), int(nsteps), stream_ptr())
        clock.advance(nsteps)


def _attach_filter(ts, eqn, setter, filterkw):
    """makefilter(equation; kw...) (src/domains.jl:542) uploaded to the stepper."""
    from .domains import makefilter

    g = eqn.grid
    f = makefilter(g, realvars=eqn.dims[0] == g.nkr, dims=eqn.dims, **filterkw)
    ts.filter_host = f
    _lib.call(setter, ts.handle, _lib.dptr(np.ascontiguousarray(f, dtype=np.float64)))


class FilteredETDRK4TimeStepper(ETDRK4TimeStepper):
    """src/timesteppers.jl:471-499, 549-558: ETDRK4, then ``sol .*= filter``
    (inside the fused update kernel when the operator is built in)."""

    def __init__(self, eqn, dt, dev=None, **filterkw):
        super().__init__(eqn, dt, dev)
        _attach_filter(self, eqn, "ffhip_etdrk4_set_filter", filterkw)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
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
        _lib.call("ffhip_etdrk4_apply_filter", self._h, ptr(sol), stream_ptr())
        clock.advance(1)


class FilteredRK4TimeStepper(RK4TimeStepper):
    """src/timesteppers.jl:200-223, 276-285: RK4, then ``sol .*= filter``."""

    def __init__(self, eqn, dt=None, dev=None, **filterkw):
        super().__init__(eqn, dt, dev)
        _attach_filter(self, eqn, "ffhip_rk4_set_filter", filterkw)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        super().stepforward(sol, clock, eqn, vars, params, grid)
        _lib.call("ffhip_rk4_apply_filter", self._h, ptr(sol), stream_ptr())


class _ExplicitBase(RK4TimeStepper):
    """ForwardEuler / AB3 / LSRK54 share the RK4 object of libffhip (L in FP64,
    optional filter); the scratch arrays are the library's own."""

    kind = -1

    def __init__(self, eqn, dt=None, dev=None, **filterkw):
        dev = eqn.grid.device if dev is None else dev
        self.T = eqn.T
        self.dims = eqn.dims
        Lre, Lim = _host_L(eqn.L, eqn.dims)
        h = C.c_void_p()
        _lib.call("ffhip_rk4_create", C.byref(h), _lib.dptr(Lre), _lib.dptr(Lim), len(self.dims),
                  _lib.i64arr(self.dims), prec_of(eqn.T))
        self._h = h
        self.filtered = False
        if filterkw.pop("_filtered", False):
            _attach_filter(self, eqn, "ffhip_rk4_set_filter", filterkw)
            self.filtered = True

    def _history(self):
        return None, None

    def run_fused(self, op, sol, clock, nsteps):
        m1, m2 = self._history()
        _lib.call("ffhip_op_explicit_run", op.handle, self._h, self.kind, ptr(sol), float(clock.dt), int(nsteps),
                  int(clock.step), ptr(m1) if m1 is not None else None, ptr(m2) if m2 is not None else None,
                  stream_ptr())
        clock.advance(nsteps)


class ForwardEulerTimeStepper(_ExplicitBase):
    """src/timesteppers.jl:98-118: ``sol += dt (L sol + N)``."""

    kind = 0

    def __init__(self, eqn, dt=None, dev=None, **kw):
        super().__init__(eqn, dt, dev, **kw)
        self.N = zeros(eqn.grid.device if dev is None else dev, eqn.T, eqn.dims)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        eqn.calcN(self.N, sol, clock.t, clock, vars, params, grid)
        _lib.call("ffhip_euler_update", self._h, ptr(sol), ptr(self.N), float(clock.dt), stream_ptr())
        clock.advance(1)


class AB3TimeStepper(_ExplicitBase):
    """src/timesteppers.jl:588-650: forward Euler while ``clock.step < 3``, then
    third-order Adams-Bashforth; RHS, RHS_-1, RHS_-2 kept on the device."""

    kind = 1

    def __init__(self, eqn, dt=None, dev=None, **kw):
        super().__init__(eqn, dt, dev, **kw)
        dev = eqn.grid.device if dev is None else dev
        z = lambda: zeros(dev, eqn.T, eqn.dims)  # noqa: E731
        self.RHS, self.RHSm1, self.RHSm2 = z(), z(), z()

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        eqn.calcN(self.RHS, sol, clock.t, clock, vars, params, grid)
        self.addlinearterm(self.RHS, sol)
        _lib.call("ffhip_ab3_update", self._h, ptr(sol), ptr(self.RHS), ptr(self.RHSm1), ptr(self.RHSm2),
                  float(clock.dt), int(clock.step), stream_ptr())
        clock.advance(1)
        self.RHSm2.copy_(self.RHSm1)
        self.RHSm1.copy_(self.RHS)

    def _history(self):
        return self.RHSm1, self.RHSm2


# src/timesteppers.jl:330-345
LSRK54_A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238, -3550918686646 / 2091501179385,
            -1275806237668 / 842570457699)
LSRK54_B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357, 1720146321549 / 2090206949498,
            3134564353537 / 4481467310338, 2277821191437 / 14882151754819)
LSRK54_C = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896, 2006345519317 / 3224310063776,
            2802321613138 / 2924317926251)


class LSRK54TimeStepper(_ExplicitBase):
    """src/timesteppers.jl:318-414: five stages of S2 = A_i S2 + dt RHS(sol, t + C_i dt);
    sol += B_i S2 (coefficients rounded to the state's float type as ``T[...]`` does)."""

    kind = 2

    def __init__(self, eqn, dt=None, dev=None, **kw):
        super().__init__(eqn, dt, dev, **kw)
        dev = eqn.grid.device if dev is None else dev
        self.S2 = zeros(dev, eqn.T, eqn.dims)
        self.RHS = zeros(dev, eqn.T, eqn.dims)
        ft = np.dtype(fltype(eqn.T)).type
        self.A = tuple(float(ft(a)) for a in LSRK54_A)
        self.B = tuple(float(ft(b)) for b in LSRK54_B)
        self.C = tuple(float(ft(c)) for c in LSRK54_C)

    def stepforward(self, sol, clock, eqn, vars, params, grid):
        t, dt = clock.t, clock.dt
        for i in range(5):
            eqn.calcN(self.RHS, sol, t + self.C[i] * dt, clock, vars, params, grid)
            self.addlinearterm(self.RHS, sol)
            _lib.call("ffhip_lsrk54_stage", self._h, ptr(sol), ptr(self.S2), ptr(self.RHS), self.A[i], self.B[i],
                      float(dt), stream_ptr())
        if self.filtered:
            _lib.call("ffhip_rk4_apply_filter", self._h, ptr(sol), stream_ptr())
        clock.advance(1)


def _filtered(cls, name):
    def init(self, eqn, dt=None, dev=None, **filterkw):
        cls.__init__(self, eqn, dt, dev, _filtered=True, **filterkw)

    return type(name, (cls,), {"__init__": init, "__doc__": f"Filtered {cls.__name__} (src/timesteppers.jl)."})


FilteredForwardEulerTimeStepper = _filtered(ForwardEulerTimeStepper, "FilteredForwardEulerTimeStepper")
FilteredAB3TimeStepper = _filtered(AB3TimeStepper, "FilteredAB3TimeStepper")
FilteredLSRK54TimeStepper = _filtered(LSRK54TimeStepper, "FilteredLSRK54TimeStepper")


_STEPPERS = {"ETDRK4": ETDRK4TimeStepper, "RK4": RK4TimeStepper, "ForwardEuler": ForwardEulerTimeStepper,
             "AB3": AB3TimeStepper, "LSRK54": LSRK54TimeStepper, "FilteredETDRK4": FilteredETDRK4TimeStepper,
             "FilteredRK4": FilteredRK4TimeStepper, "FilteredForwardEuler": FilteredForwardEulerTimeStepper,
             "FilteredAB3": FilteredAB3TimeStepper, "FilteredLSRK54": FilteredLSRK54TimeStepper}
# src/timesteppers.jl:672-681
fullyexplicitsteppers = ["ForwardEuler", "RK4", "AB3", "LSRK54", "FilteredForwardEuler", "FilteredRK4",
                         "FilteredAB3", "FilteredLSRK54"]


def isexplicit(stepper):
    return stepper in fullyexplicitsteppers


def TimeStepper(stepper, equation, dt=None, dev=None, **kw):
    """src/timesteppers.jl:690-702 (string -> stepper; keywords go to the
    filtered steppers' makefilter)."""
    try:
        cls = _STEPPERS[stepper]
    except KeyError:
        raise ValueError(f"unknown time stepper {stepper!r} (available: {sorted(_STEPPERS)})") from None
    return cls(equation, dt, dev, **kw)


class EmptyParams:
    pass


class EmptyVars:
    pass


class Problem:
    """``Problem(eqn, stepper, dt, grid, vars, params)`` (src/problem.jl:99-111)."""

    def __init__(self, eqn, stepper, dt, grid, vars=None, params=None, **stepperkw):
        dev = grid.device
        if not isinstance(dev, HIP):
            raise RuntimeError("libffhip problems run on HIP devices")
        self.eqn, self.grid = eqn, grid
        self.vars = EmptyVars() if vars is None else vars
        self.params = EmptyParams() if params is None else params
        self.clock = Clock(dt, T=fltype(grid.T))
        self.timestepper = TimeStepper(stepper, eqn, dt, dev, **stepperkw)
        self.stepper = stepper
        self.sol = zeros(dev, eqn.T, eqn.dims)


def stepforward(prob, *args):
    """``stepforward!(prob[, diags], nsteps)`` (src/timesteppers.jl:6-35).

    With a built-in operator as ``calcN!`` the steps run inside libffhip
    (fused kernels, no host round trips); otherwise each step calls the
    Python ``calcN!`` between library pointwise passes.  With ``diags`` (a
    ``Diagnostic`` or a list of them) the diagnostics are incremented after
    every step as the reference does; with a built-in operator the library
    runs the steps between two diagnostic updates as one call.
    """
    if len(args) == 0:
        diags, nsteps = None, 1
    elif len(args) == 1:
        diags, nsteps = None, args[0]
    else:
        diags, nsteps = args
    nsteps = int(nsteps)
    if diags is not None:
        from .diagnostics import _as_list, increment

        dl = _as_list(diags)
        op = getattr(prob.eqn.calcN, "builtin_operator", None)
        if op is None:
            for _ in range(nsteps):
                stepforward(prob, 1)
                increment(dl)
            return None
        done = 0
        while done < nsteps:
            # steps until the next update of any diagnostic (or the end)
            step = prob.clock.step
            chunk = min(min(d.freq - (step % d.freq) for d in dl), nsteps - done)
            stepforward(prob, chunk)
            done += chunk
            increment(dl)
        return None
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
    # a Julia Float64 stop_time promotes the Float32 clock (np.float64 does too)
    interval = np.float64(stop_time) - prob.clock.t
    nsteps = int(np.floor(interval / dt))
    stepforward(prob, nsteps)
    t_remaining = interval - prob.clock.t  # as in the reference (:752)
    prob.clock.dt = type(dt)(t_remaining)  # stored in the clock's type, as Clock{T}'s field
    stepforward(prob)
    prob.clock.dt = dt
