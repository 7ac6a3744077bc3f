"""Reverse-mode differentiation of the pseudospectral step (no counterpart in
the reference, which has no adjoint).

These are the analogues of ChainRulesCore ``rrule``s, as torch autograd
Functions whose backward passes call libffhip's hand-written transposes:

* ``ldiv(grid, X)``, ``mul(grid, phi)``: ``ldiv!``/``mul!`` on
  ``grid.rfftplan`` with ``ldiv!^T``/``mul!^T`` as pullbacks;
* ``dealias(grid, X)``: ``dealias!`` on a copy (self-adjoint);
* ``etdrk4_solve(prob, sol0, nsteps, ncheck)``: ``nsteps`` fused ETDRK4
  steps of a built-in 2D vorticity problem, whose pullback is the discrete
  adjoint with ``ncheck`` checkpoints and per-segment recompute;
* ``rk4_solve(prob, sol0, nsteps)``: RK4 steps of the built-in tracer
  problem, whose pullback is the discrete RK4 adjoint (the equation is
  linear in c, so no checkpoints are needed).

Complex tensors follow torch's convention: the gradient of a real loss with
respect to ``z`` is ``dL/dRe z + i dL/dIm z``, which is exactly the
transpose for the real inner product ``Re sum conj(a) b`` used by libffhip.

``misfit_gradient`` is the same computation without autograd: the value of
``J = 1/2 ||ldiv(S^n(P mul(zeta0))) - zeta_obs||^2`` and its gradient with
respect to the physical initial field ``zeta0`` (``P`` is dealiasing).
"""

from __future__ import annotations

from . import _lib
from .core import ptr, stream_ptr, zeros
from .domains import dealias as _dealias
from .timesteppers import ETDRK4TimeStepper, RK4TimeStepper


def _torch():
    import torch

    return torch


def _plan_apply(fn, plan, out, inp):
    _lib.call(fn, plan.handle, ptr(inp), ptr(out), stream_ptr())
    return out


class _LDiv:
    @staticmethod
    def forward(ctx, grid, X):
        ctx.grid = grid
        out = zeros(grid.device, grid.T, grid.phys_dims)
        return _plan_apply("ffhip_plan_inverse", grid.rfftplan, out, X.contiguous())

    @staticmethod
    def backward(ctx, g):
        grid = ctx.grid
        out = zeros(grid.device, _spec_T(grid), grid.spec_dims)
        return None, _plan_apply("ffhip_plan_inverse_adjoint", grid.rfftplan, out, g.contiguous())


class _Mul:
    @staticmethod
    def forward(ctx, grid, phi):
        ctx.grid = grid
        out = zeros(grid.device, _spec_T(grid), grid.spec_dims)
        return _plan_apply("ffhip_plan_forward", grid.rfftplan, out, phi.contiguous())

    @staticmethod
    def backward(ctx, g):
        grid = ctx.grid
        out = zeros(grid.device, grid.T, grid.phys_dims)
        return None, _plan_apply("ffhip_plan_forward_adjoint", grid.rfftplan, out, g.contiguous())


class _Dealias:
    @staticmethod
    def forward(ctx, grid, X):
        ctx.grid = grid
        out = X.detach().clone().contiguous()
        _dealias(out, grid)
        return out

    @staticmethod
    def backward(ctx, g):
        out = g.detach().clone().contiguous()
        _dealias(out, ctx.grid)
        return None, out


def _checkpoint_id(op):
    import ctypes

    v = ctypes.c_int64(0)
    _lib.call("ffhip_op_checkpoint_id", op.handle, ctypes.byref(v))
    return v.value


class _ETDRK4Solve:
    @staticmethod
    def forward(ctx, prob, sol0, nsteps, ncheck):
        op = _fused_operator(prob)
        ts = prob.timestepper
        sol0 = sol0.detach().contiguous()
        solT = sol0.clone()
        _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, ptr(solT), ptr(solT), int(nsteps),
                  int(ncheck), stream_ptr())
        ctx.prob, ctx.nsteps, ctx.ncheck = prob, int(nsteps), int(ncheck)
        # the operator holds one set of checkpoints: remember which run made
        # them and keep sol0, so that backward can rebuild them if a later
        # solve on the same operator has replaced them
        ctx.sol0 = sol0.clone()
        ctx.ck_id = _checkpoint_id(op)
        return solT

    @staticmethod
    def backward(ctx, g):
        prob = ctx.prob
        op = _fused_operator(prob)
        ts = prob.timestepper
        if _checkpoint_id(op) != ctx.ck_id:
            scratch = ctx.sol0.clone()
            _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, ptr(scratch), ptr(scratch), ctx.nsteps,
                      ctx.ncheck, stream_ptr())
            ctx.ck_id = _checkpoint_id(op)
        lam = g.detach().clone().contiguous()
        _lib.call("ffhip_op_etdrk4_adjoint", op.handle, ts.handle, ptr(lam), ptr(lam), ctx.nsteps,
                  ctx.ncheck, stream_ptr())
        return None, lam, None, None


class _RK4Solve:
    @staticmethod
    def forward(ctx, prob, sol0, nsteps):
        op = _tracer_operator(prob)
        solT = sol0.detach().clone().contiguous()
        _lib.call("ffhip_op_rk4_run", op.handle, prob.timestepper.handle, ptr(solT), float(prob.clock.dt),
                  int(nsteps), stream_ptr())
        ctx.prob, ctx.nsteps = prob, int(nsteps)
        return solT

    @staticmethod
    def backward(ctx, g):
        prob = ctx.prob
        op = _tracer_operator(prob)
        lam = g.detach().clone().contiguous()
        _lib.call("ffhip_op_rk4_adjoint", op.handle, prob.timestepper.handle, ptr(lam), ptr(lam),
                  float(prob.clock.dt), ctx.nsteps, stream_ptr())
        return None, lam, None


def _tracer_operator(prob):
    op = getattr(prob.eqn.calcN, "builtin_operator", None)
    if op is None or op.kind != _lib.OP_TRACER3D:
        raise _lib.FFHIPError(_lib.ERR_UNSUPPORTED, "the RK4 adjoint needs the built-in tracer operator")
    if not isinstance(prob.timestepper, RK4TimeStepper):
        raise TypeError("the tracer adjoint is implemented for the RK4 stepper")
    return op


def _spec_T(grid):
    import numpy as np

    return np.result_type(grid.T, np.complex64)


def _fused_operator(prob):
    op = getattr(prob.eqn.calcN, "builtin_operator", None)
    if op is None or not getattr(op, "fused", False):
        raise _lib.FFHIPError(_lib.ERR_UNSUPPORTED,
                              "the ETDRK4 adjoint needs a built-in 2D vorticity operator on a power-of-two grid")
    if not isinstance(prob.timestepper, ETDRK4TimeStepper):
        raise TypeError("the adjoint is implemented for the ETDRK4 stepper")
    return op


_FUNCS = {}


def _function(cls):
    """torch.autograd.Function subclass built on first use (torch is imported lazily)."""
    f = _FUNCS.get(cls)
    if f is None:
        torch = _torch()
        f = type(cls.__name__.lstrip("_"), (torch.autograd.Function,),
                 {"forward": cls.forward, "backward": cls.backward})
        _FUNCS[cls] = f
    return f


def ldiv(grid, X):
    """Differentiable ``ldiv!(x, grid.rfftplan, X)``; returns ``x``."""
    return _function(_LDiv).apply(grid, X)


def mul(grid, phi):
    """Differentiable ``mul!(X, grid.rfftplan, phi)``; returns ``X``."""
    return _function(_Mul).apply(grid, phi)


def dealias(grid, X):
    """Differentiable out-of-place ``dealias!``."""
    return _function(_Dealias).apply(grid, X)


def etdrk4_solve(prob, sol0, nsteps, ncheck=None):
    """``nsteps`` ETDRK4 steps from ``sol0`` (not modified), differentiable in
    ``sol0``; ``ncheck`` uniformly spaced checkpoints (default ~sqrt(nsteps)).
    The problem's ``clock`` and ``sol`` are left untouched."""
    if ncheck is None:
        ncheck = max(1, int(round(nsteps**0.5)))
    return _function(_ETDRK4Solve).apply(prob, sol0, int(nsteps), int(ncheck))


def rk4_solve(prob, sol0, nsteps):
    """``nsteps`` RK4 steps of a built-in tracer problem from ``sol0`` (not
    modified), differentiable in ``sol0``.  The tracer equation is linear in
    c, so the pullback needs no stored states."""
    return _function(_RK4Solve).apply(prob, sol0, int(nsteps))


def tracer_misfit_gradient(prob, c0, c_obs, nsteps):
    """(J, dJ/dc0) for ``J = 1/2 sum (ldiv(S^n(P mul(c0))) - c_obs)^2`` with
    RK4 steps of the built-in tracer operator (explicit library calls)."""
    torch = _torch()
    grid = prob.grid
    op = _tracer_operator(prob)
    ts = prob.timestepper
    plan = grid.rfftplan
    s = zeros(grid.device, _spec_T(grid), grid.spec_dims)
    _plan_apply("ffhip_plan_forward", plan, s, c0.contiguous())
    _dealias(s, grid)
    _lib.call("ffhip_op_rk4_run", op.handle, ts.handle, ptr(s), float(prob.clock.dt), int(nsteps), stream_ptr())
    cT = zeros(grid.device, grid.T, grid.phys_dims)
    _plan_apply("ffhip_plan_inverse", plan, cT, s)
    r = cT - c_obs
    J = 0.5 * float(torch.sum(r.double() ** 2))
    lam = s
    _plan_apply("ffhip_plan_inverse_adjoint", plan, lam, r.contiguous())
    _lib.call("ffhip_op_rk4_adjoint", op.handle, ts.handle, ptr(lam), ptr(lam), float(prob.clock.dt), int(nsteps),
              stream_ptr())
    _dealias(lam, grid)
    grad = zeros(grid.device, grid.T, grid.phys_dims)
    _plan_apply("ffhip_plan_forward_adjoint", plan, grad, lam)
    return J, grad


def misfit_gradient(prob, zeta0, zeta_obs, nsteps, ncheck):
    """(J, dJ/dzeta0) for ``J = 1/2 sum (ldiv(S^n(P mul(zeta0))) - zeta_obs)^2``
    over the physical grid points, with explicit library calls (no autograd
    graph).  ``zeta0``/``zeta_obs`` are physical fields (C order (ny, nx))."""
    torch = _torch()
    grid = prob.grid
    op = _fused_operator(prob)
    ts = prob.timestepper
    plan = grid.rfftplan
    s = zeros(grid.device, _spec_T(grid), grid.spec_dims)
    _plan_apply("ffhip_plan_forward", plan, s, zeta0.contiguous())
    _dealias(s, grid)
    _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, ptr(s), ptr(s), int(nsteps), int(ncheck),
              stream_ptr())
    zT = zeros(grid.device, grid.T, grid.phys_dims)
    _plan_apply("ffhip_plan_inverse", plan, zT, s)
    r = zT - zeta_obs
    J = 0.5 * float(torch.sum(r.double() ** 2))
    lam = s  # reuse
    _plan_apply("ffhip_plan_inverse_adjoint", plan, lam, r.contiguous())
    _lib.call("ffhip_op_etdrk4_adjoint", op.handle, ts.handle, ptr(lam), ptr(lam), int(nsteps), int(ncheck),
              stream_ptr())
    _dealias(lam, grid)
    grad = zeros(grid.device, grid.T, grid.phys_dims)
    _plan_apply("ffhip_plan_forward_adjoint", plan, grad, lam)
    return J, grad
