"""Nonlinear terms: 2D vorticity, 3D passive tracer, 1D diffusion testbed.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The 2D vorticity and 3D tracer ``calcN!`` bodies are not in the reference
tree (they live in its child packages); they are written here in the
reference's own idiom: spectral multiplies, ``ldiv!`` to physical space,
physical products, ``mul!`` back, ``dealias!`` (pattern
src/diffusion.jl:135-143; Jacobian formula src/utils.jl:190-205).

Convention shared with the HIP operators: the operator evaluates N on the
dealiased state (modes that ``dealias!`` zeroes are treated as zero on
input) and returns a dealiased N.
"""

from __future__ import annotations

import numpy as np

from .grids import dealias


# ---------------------------------------------------------------------------
# 2D incompressible vorticity:  dζ/dt = L ζ - J(ψ, ζ),  ψ̂ = -ζ̂/|k|²
# ---------------------------------------------------------------------------

class VorticityVars:
    def __init__(self, grid):
        self.grid = grid


def vorticity_calcN(N, sol, t, clock, vars, params, grid):
    """N̂ = -i kr (uζ)^ - i l (vζ)^, dealiased; u = -ψ_y, v = ψ_x.

    û = i l invKrsq ζ̂ and v̂ = -i kr invKrsq ζ̂ (so ψ̂ = -invKrsq ζ̂).
    Equals -J(ψ, ζ) with J from src/utils.jl:190-205.
    """
    zh = dealias(sol.copy(), grid)
    uh = 1j * grid.l * grid.invKrsq * zh
    vh = -1j * grid.kr * grid.invKrsq * zh
    u = grid.ldiv(uh)
    v = grid.ldiv(vh)
    z = grid.ldiv(zh)
    uzh = grid.mul(u * z)
    vzh = grid.mul(v * z)
    N[...] = -1j * grid.kr * uzh - 1j * grid.l * vzh
    dealias(N, grid)


def vorticity_linear(grid, nu=0.0, nu_n=0.0, n_hyper=1, mu=0.0):
    """L = -nu |k|^2 - nu_n |k|^(2 n_hyper) - mu on the rfft grid."""
    K2 = grid.Krsq
    L = -nu * K2 - mu
    if nu_n != 0.0:
        L = L - nu_n * K2**n_hyper
    return L


def hyperviscosity_for(n, dt, n_hyper=4, aliased_fraction=1 / 3, Lx=2 * np.pi):
    """nu_n such that nu_n * (kmax * (1 - aliased_fraction))^(2 n_hyper) = 1/dt.

    kmax = n/2 * 2π/Lx, so at n = 2048 the damping at 2/3 of kmax is 1/dt.
    """
    kcut = (n / 2) * (2 * np.pi / Lx) * (1 - aliased_fraction)
    return 1.0 / (dt * kcut ** (2 * n_hyper))


# ---------------------------------------------------------------------------
# 3D passive tracer in a steady ABC flow:  dc/dt = κ∇²c - ∇·(u c)
# ---------------------------------------------------------------------------

class TracerParams:
    def __init__(self, A=1.0, B=1.0, C=1.0, kappa=1e-3):
        self.A, self.B, self.C, self.kappa = A, B, C, kappa


def abc_velocity(grid, params):
    """u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x) on the grid."""
    x = grid.x[None, None, :]
    y = grid.y[None, :, None]
    z = grid.z[:, None, None]
    A, B, C = params.A, params.B, params.C
    shape = grid.phys_shape
    u = np.broadcast_to(A * np.sin(z) + C * np.cos(y), shape)
    v = np.broadcast_to(B * np.sin(x) + A * np.cos(z), shape)
    w = np.broadcast_to(C * np.sin(y) + B * np.cos(x), shape)
    return u, v, w


class TracerVars:
    def __init__(self, grid, params):
        self.u, self.v, self.w = abc_velocity(grid, params)


def tracer_calcN(N, sol, t, clock, vars, params, grid):
    """N̂ = -i (kr (uc)^ + l (vc)^ + m (wc)^), dealiased (conservative form)."""
    ch = dealias(sol.copy(), grid)
    c = grid.ldiv(ch)
    ucH = grid.mul(vars.u * c)
    vcH = grid.mul(vars.v * c)
    wcH = grid.mul(vars.w * c)
    N[...] = -1j * (grid.kr * ucH + grid.l * vcH + grid.m * wcH)
    dealias(N, grid)


def tracer_linear(grid, kappa):
    return -kappa * grid.Krsq


# ---------------------------------------------------------------------------
# 1D diffusion testbed (src/diffusion.jl)
# ---------------------------------------------------------------------------

class DiffusionParams:
    """src/diffusion.jl:63-71: κ is a number or an array on the x grid."""

    def __init__(self, kappa):
        self.kappa = kappa


class DiffusionVars:
    """src/diffusion.jl:97-124."""

    def __init__(self, grid):
        self.c = np.zeros(grid.nx)
        self.cx = np.zeros(grid.nx)
        self.ch = np.zeros(grid.nkr, dtype=np.complex128)
        self.cxh = np.zeros(grid.nkr, dtype=np.complex128)


def diffusion_calcN(N, sol, t, clock, vars, params, grid):
    """src/diffusion.jl:129-143."""
    if np.ndim(params.kappa) == 0:
        N[...] = 0
        return
    vars.cxh[...] = 1j * grid.kr * sol
    vars.cx[...] = grid.ldiv(vars.cxh)
    vars.cx *= params.kappa
    vars.cxh[...] = grid.mul(vars.cx)
    N[...] = 1j * grid.kr * vars.cxh


def diffusion_equation(grid, params):
    """src/diffusion.jl:82-90: L = -κ kr² for constant κ, scalar L = 0 otherwise."""
    from .timesteppers import Equation

    if np.ndim(params.kappa) == 0:
        L = -params.kappa * grid.kr**2
        return Equation(L, diffusion_calcN, grid)
    return Equation(0.0, diffusion_calcN, grid, dims=(grid.nkr,), T=np.complex128)


def diffusion_problem(nx=128, Lx=2 * np.pi, kappa=0.0, dt=0.01, stepper="RK4",
                      aliased_fraction=0):
    """src/diffusion.jl:44-57."""
    from .grids import OneDGrid
    from .timesteppers import Problem

    grid = OneDGrid(nx, Lx, aliased_fraction=aliased_fraction)
    params = DiffusionParams(kappa)
    vars = DiffusionVars(grid)
    eqn = diffusion_equation(grid, params)
    return Problem(eqn, stepper, dt, grid, vars, params)


def diffusion_set_c(prob, c):
    """src/diffusion.jl:166-176."""
    prob.vars.c[...] = c
    prob.sol[...] = prob.grid.mul(prob.vars.c)
    diffusion_updatevars(prob)


def diffusion_updatevars(prob):
    """src/diffusion.jl:150-161."""
    v, g = prob.vars, prob.grid
    v.ch[...] = prob.sol
    v.cxh[...] = 1j * g.kr * prob.sol
    v.c[...] = g.ldiv(v.ch.copy())
    v.cx[...] = g.ldiv(v.cxh.copy())
