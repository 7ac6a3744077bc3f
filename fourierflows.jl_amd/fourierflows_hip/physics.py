"""Built-in nonlinear operators and the 1D diffusion testbed.

The reference's ``calcN!`` is arbitrary user Julia; a C ABI cannot compile
it, so the physics the hot path needs is built into libffhip and exposed as
callables with the ``calcN!(N, sol, t, clock, vars, params, grid)``
signature.  Passing one as ``calcN!`` to ``Equation`` lets ``stepforward``
run whole fused steps in the library.

* ``VorticityOperator``: 2D incompressible vorticity, N = -J(ψ, ζ),
  ψ̂ = -ζ̂/|k|² (the physics of GeophysicalFlows' TwoDNavierStokes module,
  written in the reference's idiom: src/diffusion.jl:135-143 pattern,
  src/utils.jl:190-205 Jacobian).
* ``TracerOperator``: 3D passive tracer in a steady ABC flow, N = -∇·(u c).
* ``Diffusion``: the reference's own testbed (src/diffusion.jl), whose
  ``calcN!`` is plain Python over ``mul``/``ldiv`` and device arrays.
"""

from __future__ import annotations

import ctypes as C
import types

import numpy as np

from . import _lib
from .core import HIP, device_array, prec_of, ptr, stream_ptr, zeros
from .domains import OneDGrid, ThreeDGrid, TwoDGrid, ldiv, mul
from .timesteppers import Equation, Problem


class BuiltinOperator:
    kind = 0

    def _create(self, grid, params=()):
        self.grid = grid
        h = C.c_void_p()
        p = np.asarray(params, dtype=np.float64)
        _lib.call("ffhip_op_create", C.byref(h), self.kind, prec_of(grid.T), grid.ndim,
                  _lib.i64arr(grid.ns), _lib.dptr(np.asarray(grid.Ls, dtype=np.float64)),
                  float(grid.aliased_fraction), _lib.dptr(p) if p.size else None, int(p.size))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def builtin_operator(self):
        return self

    def __call__(self, N, sol, t, clock=None, vars=None, params=None, grid=None):
        _lib.call("ffhip_op_calcN", self._h, ptr(N), ptr(sol), float(t), stream_ptr())

    @property
    def fused(self):
        """True when libffhip steps this operator's grid with fused kernels."""
        v = C.c_int()
        _lib.call("ffhip_op_is_fused", self._h, C.byref(v))
        return bool(v.value)

    def diagnostics(self, sol):
        out = (C.c_double * 2)()
        _lib.call("ffhip_op_diagnostics", self._h, ptr(sol), out, stream_ptr())
        return out[0], out[1]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.ffhip_op_destroy(h)
            self._h = None


class VorticityOperator(BuiltinOperator):
    """``calcN!`` of 2D vorticity: N̂ = -i kr (uζ)^ - i l (vζ)^, dealiased,
    with û = i l ζ̂/|k|², v̂ = -i kr ζ̂/|k|².  Evaluated on the kept band of
    ``sol`` (modes ``dealias!`` would zero are read as zero)."""

    kind = _lib.OP_VORTICITY2D

    def __init__(self, grid):
        if not isinstance(grid, TwoDGrid):
            raise TypeError("VorticityOperator needs a TwoDGrid")
        self._create(grid)

    def energy_enstrophy(self, sol):
        """(½⟨|u|²⟩, ½⟨ζ²⟩) by compensated FP64 Parseval sums."""
        return self.diagnostics(sol)


class TracerOperator(BuiltinOperator):
    """``calcN!`` of a passive tracer in the steady ABC flow
    u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x):
    N̂ = -i (kr (uc)^ + l (vc)^ + m (wc)^), dealiased."""

    kind = _lib.OP_TRACER3D

    def __init__(self, grid, A=1.0, B=1.0, C_=1.0):
        if not isinstance(grid, ThreeDGrid):
            raise TypeError("TracerOperator needs a ThreeDGrid")
        self.A, self.B, self.C = A, B, C_
        self._create(grid, (A, B, C_, grid.x0s[0], grid.x0s[1], grid.x0s[2]))

    def variance_mean(self, sol):
        """(½⟨c²⟩, ⟨c⟩)."""
        return self.diagnostics(sol)


# ------------------------------------------------------------------ problem builders
def vorticity_linear(grid, nu=0.0, nu_n=0.0, n_hyper=1, mu=0.0):
    """L = -ν|k|² - ν_n|k|^(2 n_hyper) - μ (FP64 host array, shape (ny, nkr))."""
    K2 = grid.Krsq_host
    L = -nu * K2 - mu
    if nu_n != 0.0:
        L = L - nu_n * K2**n_hyper
    return L


def TwoDVorticityProblem(dev=None, *, nx=256, Lx=2 * np.pi, ny=None, Ly=None, nu=0.0, nu_n=0.0, n_hyper=1,
                         mu=0.0, dt=0.01, stepper="ETDRK4", T=np.float64, aliased_fraction=1 / 3, x0=0.0,
                         y0=0.0):
    """A 2D vorticity problem on [x0, x0+Lx) x [y0, y0+Ly) with a built-in fused operator."""
    dev = HIP() if dev is None else dev
    grid = TwoDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, x0=x0, y0=y0, T=T, aliased_fraction=aliased_fraction)
    op = VorticityOperator(grid)
    L = vorticity_linear(grid, nu, nu_n, n_hyper, mu)
    eqn = Equation(L, op, grid)
    return Problem(eqn, stepper, dt, grid, None, types.SimpleNamespace(nu=nu, nu_n=nu_n, n_hyper=n_hyper, mu=mu))


def TracerProblem(dev=None, *, n=256, L=2 * np.pi, kappa=1e-3, A=1.0, B=1.0, C=1.0, dt=0.01, stepper="RK4",
                  T=np.float32, aliased_fraction=1 / 3):
    """3D passive tracer on [0, L)^3 in the steady ABC flow with L = -κ|k|²."""
    dev = HIP() if dev is None else dev
    grid = ThreeDGrid(dev, nx=n, Lx=L, x0=0.0, y0=0.0, z0=0.0, T=T, aliased_fraction=aliased_fraction)
    op = TracerOperator(grid, A, B, C)
    Lin = -kappa * grid.Krsq_host
    eqn = Equation(Lin, op, grid)
    return Problem(eqn, stepper, dt, grid, None, types.SimpleNamespace(kappa=kappa, A=A, B=B, C=C))


# ------------------------------------------------------------------ diffusion testbed
class Diffusion:
    """src/diffusion.jl: 1D diffusion with constant κ (L = -κ k², N = 0) or
    array κ (scalar L = 0 and N = ∂x(κ ∂x c) through mul!/ldiv!)."""

    class Params:
        def __init__(self, grid, kappa):
            if np.ndim(kappa) == 0:
                self.kappa = float(kappa)
            else:
                self.kappa = device_array(grid.device)(np.asarray(kappa, dtype=grid.T))

    class Vars:
        def __init__(self, grid):
            dev = grid.device
            self.c = zeros(dev, grid.T, (grid.nx,))
            self.cx = zeros(dev, grid.T, (grid.nx,))
            self.ch = zeros(dev, np.result_type(grid.T, np.complex64), (grid.nkr,))
            self.cxh = zeros(dev, np.result_type(grid.T, np.complex64), (grid.nkr,))

    @staticmethod
    def calcN(N, sol, t, clock, vars, params, grid):
        """src/diffusion.jl:129-143."""
        if isinstance(params.kappa, float):
            N.zero_()
            return
        torch = __import__("torch")
        ikr = 1j * grid.kr.to(sol.dtype)
        torch.mul(ikr, sol, out=vars.cxh)
        ldiv(vars.cx, grid.rfftplan, vars.cxh)
        vars.cx.mul_(params.kappa)
        mul(vars.cxh, grid.rfftplan, vars.cx)
        torch.mul(ikr, vars.cxh, out=N)

    @staticmethod
    def Equation(grid, params):
        """src/diffusion.jl:82-90."""
        if isinstance(params.kappa, float):
            L = -params.kappa * np.asarray(grid.kr_host) ** 2
            return Equation(L, Diffusion.calcN, grid)
        return Equation(0.0, Diffusion.calcN, grid, dims=(grid.nkr,),
                        T=np.result_type(grid.T, np.complex64))

    @staticmethod
    def Problem(dev=None, *, nx=128, Lx=2 * np.pi, kappa=0.0, dt=0.01, stepper="RK4", aliased_fraction=0,
                T=np.float64):
        """src/diffusion.jl:44-57."""
        dev = HIP() if dev is None else dev
        grid = OneDGrid(dev, nx=nx, Lx=Lx, aliased_fraction=aliased_fraction, T=T)
        params = Diffusion.Params(grid, kappa)
        vars = Diffusion.Vars(grid)
        eqn = Diffusion.Equation(grid, params)
        return Problem(eqn, stepper, dt, grid, vars, params)

    @staticmethod
    def updatevars(prob):
        """src/diffusion.jl:150-161 (ldiv! keeps its input here, so no deepcopy)."""
        v, g = prob.vars, prob.grid
        v.ch.copy_(prob.sol)
        torch = __import__("torch")
        torch.mul(1j * g.kr.to(prob.sol.dtype), prob.sol, out=v.cxh)
        ldiv(v.c, g.rfftplan, v.ch)
        ldiv(v.cx, g.rfftplan, v.cxh)

    @staticmethod
    def set_c(prob, c):
        """src/diffusion.jl:166-176."""
        v = prob.vars
        v.c.copy_(device_array(prob.grid.device)(np.asarray(c, dtype=prob.grid.T)))
        mul(prob.sol, prob.grid.rfftplan, v.c)
        Diffusion.updatevars(prob)


# ------------------------------------------------------------------ ensembles
class VorticityEnsemble:
    """A batch of independent 2D vorticity problems on mixed power-of-two
    grids (64..1024), stepped together by libffhip: one kernel launch per
    pass covers every member and steps replay from a captured HIP graph.
    No counterpart in the reference, which steps one ``Problem`` at a time.

    ``members``: list of dicts with keys ``nx`` (and optionally ``ny``,
    ``Lx``, ``Ly``, ``nu``, ``nu_n``, ``n_hyper``, ``mu``); ``dt`` is common.
    ``sols[m]`` is member m's spectral state (torch tensor, reference
    layout) and ``grids[m]`` its ``TwoDGrid``.
    """

    def __init__(self, members, dt, T=np.float32, aliased_fraction=1 / 3, autotune=False, dev=None):
        dev = HIP() if dev is None else dev
        self.T = T
        self.dt = dt
        self.grids, self.sols, Ls = [], [], []
        n, box = [], []
        for m in members:
            nx = int(m["nx"])
            ny = int(m.get("ny", nx))
            Lx = float(m.get("Lx", 2 * np.pi))
            Ly = float(m.get("Ly", Lx))
            g = TwoDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, x0=0.0, y0=0.0, T=T, aliased_fraction=aliased_fraction)
            L = vorticity_linear(g, m.get("nu", 0.0), m.get("nu_n", 0.0), m.get("n_hyper", 1), m.get("mu", 0.0))
            Ls.append(np.ascontiguousarray(np.broadcast_to(L, (ny, g.nkr)), dtype=np.float64))
            self.grids.append(g)
            self.sols.append(zeros(dev, np.result_type(T, np.complex64), (g.nkr, ny)))
            n += [nx, ny]
            box += [Lx, Ly]
        self._Ls = Ls  # keep the host arrays alive for the call
        Lp = (C.POINTER(C.c_double) * len(Ls))(*[_lib.dptr(L) for L in Ls])
        h = C.c_void_p()
        _lib.call("ffhip_ensemble_create", C.byref(h), prec_of(T), len(members), _lib.i64arr(n),
                  _lib.dptr(np.asarray(box, dtype=np.float64)), float(aliased_fraction), Lp, float(dt),
                  1 if autotune else 0)
        self._h = h
        self.step = 0

    def __len__(self):
        return len(self.sols)

    def run(self, nsteps):
        """``nsteps`` ETDRK4 steps of every member (states updated in place)."""
        ptrs = (C.c_void_p * len(self.sols))(*[ptr(s) for s in self.sols])
        _lib.call("ffhip_ensemble_run", self._h, ptrs, int(nsteps), stream_ptr())
        self.step += int(nsteps)

    def info(self):
        """(threads per workgroup, points-per-thread masks, tuned ms per step or 0).

        Mask bit s (grid size 64 << s) set: 16 points per thread in that
        size's column passes, else 8; bit 8 + s the same for its row passes."""
        t, p, ms = C.c_int(), C.c_int(), C.c_double()
        _lib.call("ffhip_ensemble_info", self._h, C.byref(t), C.byref(p), C.byref(ms))
        return t.value, p.value, ms.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.ffhip_ensemble_destroy(h)
            self._h = None
