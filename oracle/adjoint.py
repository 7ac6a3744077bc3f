"""Discrete adjoints of the ETDRK4 vorticity step and the RK4 tracer step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  No counterpart in
the reference: this is the FP64 CPU adjoint the GPU adjoint is checked
against.  It is pinned by Taylor tests (central finite differences, tests
in tests/test_oracle_adjoint.py) and by dot-product identities of every
transposed operator.

Inner products.  Physical fields: <a, b> = sum a*b.  Spectral (rfft)
arrays: <a, b> = Re sum conj(a) b over the stored entries (a real inner
product on the real vector space of the stored complex entries).  With
these, for the reference's transforms (src/domains.jl:3,5; ldiv! scaled by
1/N, c2r ignoring Im at kx = 0 and kx = nx/2):

    ldiv!^T(phi) = W * mul!(phi) / N          (W = 1 at kx = 0, nx/2; 2 elsewhere)
    mul!^T(Y)    = N * ldiv!(Y / W)
"""

from __future__ import annotations

import math

import numpy as np

from .grids import kept_mask
from .timesteppers import getetdcoeffs, getexpLs


def rfft_weights(grid):
    """W: 1 on the kx = 0 and kx = nx/2 columns, 2 elsewhere (spectral shape)."""
    W = np.full(grid.spec_shape, 2.0)
    W[..., 0] = 1.0
    W[..., grid.nkr - 1] = 1.0
    return W


def ldiv_T(phi, grid, W=None):
    W = rfft_weights(grid) if W is None else W
    return W * grid.mul(phi) / float(np.prod(grid.ns))


def mul_T(Y, grid, W=None):
    W = rfft_weights(grid) if W is None else W
    return float(np.prod(grid.ns)) * grid.ldiv(Y / W)


def sdot(a, b):
    """Real inner product of two spectral arrays."""
    return float(np.sum(a.real * b.real + a.imag * b.imag))


# ---------------------------------------------------------------------------
# 2D vorticity: N(s) = P(Dx mul(u z) + Dy mul(v z)), u = ldiv(Du P s), v = ldiv(Dv P s), z = ldiv(P s)
# ---------------------------------------------------------------------------

class VorticityOps:
    def __init__(self, grid):
        self.g = grid
        self.mask = kept_mask(grid).astype(float)
        self.W = rfft_weights(grid)
        self.Du = 1j * grid.l * grid.invKrsq
        self.Dv = -1j * grid.kr * grid.invKrsq
        self.Dx = -1j * grid.kr
        self.Dy = -1j * grid.l

    def fields(self, s):
        m = self.mask * s
        g = self.g
        return g.ldiv(self.Du * m), g.ldiv(self.Dv * m), g.ldiv(m)

    def N(self, s, flds=None):
        u, v, z = self.fields(s) if flds is None else flds
        g = self.g
        return self.mask * (self.Dx * g.mul(u * z) + self.Dy * g.mul(v * z))

    def N_tangent(self, flds, d):
        """dN[s](d) with flds = fields(s)."""
        u, v, z = flds
        du, dv, dz = self.fields(d)
        g = self.g
        return self.mask * (self.Dx * g.mul(du * z + u * dz) + self.Dy * g.mul(dv * z + v * dz))

    def N_adjoint(self, flds, lam):
        """dN[s]^T(lam) with flds = fields(s)."""
        u, v, z = flds
        g, W = self.g, self.W
        l1 = self.mask * lam
        pbar = mul_T(np.conj(self.Dx) * l1, g, W)
        qbar = mul_T(np.conj(self.Dy) * l1, g, W)
        mbar = (np.conj(self.Du) * ldiv_T(pbar * z, g, W) + np.conj(self.Dv) * ldiv_T(qbar * z, g, W)
                + ldiv_T(pbar * u + qbar * v, g, W))
        return self.mask * mbar


class ETDRK4Adjoint:
    """ETDRK4 (src/timesteppers.jl:501-547) for dζ/dt = Lζ + N(ζ), its
    discrete adjoint, and the gradient of J = ½‖ζ(T) - ζ_obs‖² (physical grid
    points) with respect to the physical initial field ζ₀, with
    checkpoint-and-recompute over the time window."""

    def __init__(self, grid, L, dt):
        self.g = grid
        self.ops = VorticityOps(grid)
        L = np.broadcast_to(L, grid.spec_shape)
        self.E, self.Eh = getexpLs(dt, L)
        self.Z, self.Al, self.Be, self.Ga = getetdcoeffs(dt, L)

    def step(self, u):
        ops = self.ops
        N1 = ops.N(u)
        a = self.Eh * u + self.Z * N1
        N2 = ops.N(a)
        b = self.Eh * u + self.Z * N2
        N3 = ops.N(b)
        c = self.Eh * a + self.Z * (2 * N3 - N1)
        N4 = ops.N(c)
        return self.E * u + self.Al * N1 + 2 * self.Be * (N2 + N3) + self.Ga * N4, (u, a, b, c)

    def step_adjoint(self, stages, lam):
        """Transpose of the step's linearisation at the stored stage states."""
        ops = self.ops
        u, a, b, c = stages
        fu, fa, fb, fc = (ops.fields(s) for s in stages)
        ubar = np.conj(self.E) * lam
        N1b = np.conj(self.Al) * lam
        N2b = np.conj(2 * self.Be) * lam
        N3b = np.conj(2 * self.Be) * lam
        N4b = np.conj(self.Ga) * lam
        cbar = ops.N_adjoint(fc, N4b)
        abar = np.conj(self.Eh) * cbar
        N3b = N3b + 2 * np.conj(self.Z) * cbar
        N1b = N1b - np.conj(self.Z) * cbar
        bbar = ops.N_adjoint(fb, N3b)
        ubar = ubar + np.conj(self.Eh) * bbar
        N2b = N2b + np.conj(self.Z) * bbar
        abar = abar + ops.N_adjoint(fa, N2b)
        ubar = ubar + np.conj(self.Eh) * abar
        N1b = N1b + np.conj(self.Z) * abar
        ubar = ubar + ops.N_adjoint(fu, N1b)
        return ubar

    def initial_spectrum(self, zeta0):
        return self.ops.mask * self.g.mul(zeta0)

    def forward(self, zeta0, nsteps):
        u = self.initial_spectrum(zeta0)
        for _ in range(nsteps):
            u, _ = self.step(u)
        return u

    def misfit(self, zeta0, zeta_obs, nsteps):
        zT = self.g.ldiv(self.forward(zeta0, nsteps))
        return 0.5 * float(np.sum((zT - zeta_obs) ** 2))

    def gradient(self, zeta0, zeta_obs, nsteps, ncheck):
        """(J, dJ/dζ₀) with ``ncheck`` uniformly spaced checkpoints."""
        seg = math.ceil(nsteps / ncheck)
        u = self.initial_spectrum(zeta0)
        checkpoints = []
        for k in range(nsteps):
            if k % seg == 0:
                checkpoints.append((k, u.copy()))
            u, _ = self.step(u)
        r = self.g.ldiv(u) - zeta_obs
        J = 0.5 * float(np.sum(r**2))
        lam = ldiv_T(r, self.g, self.ops.W)
        for (k0, uc) in reversed(checkpoints):
            k1 = min(k0 + seg, nsteps)
            stages = []
            v = uc
            for _ in range(k0, k1):
                v, st = self.step(v)
                stages.append(st)
            for st in reversed(stages):
                lam = self.step_adjoint(st, lam)
        grad = mul_T(self.ops.mask * lam, self.g, self.ops.W)
        return J, grad


# ---------------------------------------------------------------------------
# 3D tracer: linear N(c) = P(-i (kr mul(u c) + l mul(v c) + m mul(w c))), c = ldiv(P ĉ)
# ---------------------------------------------------------------------------

class TracerOps:
    def __init__(self, grid, vel):
        self.g = grid
        self.u, self.v, self.w = vel
        self.mask = kept_mask(grid).astype(float)
        self.W = rfft_weights(grid)

    def N(self, s):
        g = self.g
        c = g.ldiv(self.mask * s)
        return self.mask * (-1j * (g.kr * g.mul(self.u * c) + g.l * g.mul(self.v * c) + g.m * g.mul(self.w * c)))

    def N_adjoint(self, lam):
        g, W = self.g, self.W
        l1 = self.mask * lam
        cbar = (self.u * mul_T(1j * g.kr * l1, g, W) + self.v * mul_T(1j * g.l * l1, g, W)
                + self.w * mul_T(1j * g.m * l1, g, W))
        return self.mask * ldiv_T(cbar, g, W)


class RK4TracerAdjoint:
    """RK4 (src/timesteppers.jl:237-274) with RHS(c) = N(c) + L c, linear in c,
    so the adjoint step needs no stored forward states."""

    def __init__(self, grid, vel, L, dt):
        self.g = grid
        self.ops = TracerOps(grid, vel)
        self.L = np.broadcast_to(L, grid.spec_shape)
        self.dt = dt

    def rhs(self, s):
        return self.ops.N(s) + self.L * s

    def rhs_adjoint(self, lam):
        return self.ops.N_adjoint(lam) + np.conj(self.L) * lam

    def step(self, u):
        h = self.dt
        k1 = self.rhs(u)
        k2 = self.rhs(u + h / 2 * k1)
        k3 = self.rhs(u + h / 2 * k2)
        k4 = self.rhs(u + h * k3)
        return u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)

    def step_adjoint(self, lam):
        h = self.dt
        ubar = lam.copy()
        k1b, k2b, k3b, k4b = h / 6 * lam, h / 3 * lam, h / 3 * lam, h / 6 * lam
        u3b = self.rhs_adjoint(k4b)
        ubar += u3b
        k3b = k3b + h * u3b
        u2b = self.rhs_adjoint(k3b)
        ubar += u2b
        k2b = k2b + h / 2 * u2b
        u1b = self.rhs_adjoint(k2b)
        ubar += u1b
        k1b = k1b + h / 2 * u1b
        ubar += self.rhs_adjoint(k1b)
        return ubar

    def initial_spectrum(self, c0):
        return self.ops.mask * self.g.mul(c0)

    def forward(self, c0, nsteps):
        u = self.initial_spectrum(c0)
        for _ in range(nsteps):
            u = self.step(u)
        return u

    def misfit(self, c0, c_obs, nsteps):
        cT = self.g.ldiv(self.forward(c0, nsteps))
        return 0.5 * float(np.sum((cT - c_obs) ** 2))

    def gradient(self, c0, c_obs, nsteps):
        u = self.forward(c0, nsteps)
        r = self.g.ldiv(u) - c_obs
        J = 0.5 * float(np.sum(r**2))
        lam = ldiv_T(r, self.g, self.ops.W)
        for _ in range(nsteps):
            lam = self.step_adjoint(lam)
        return J, mul_T(self.ops.mask * lam, self.g, self.ops.W)


def taylor_errors(J, grad, x, d, hs):
    """|(J(x+hd) - J(x-hd))/2h - <grad, d>| for each h (central differences)."""
    gd = float(np.sum(grad * d))
    out = []
    for h in hs:
        fd = (J(x + h * d) - J(x - h * d)) / (2 * h)
        out.append(abs(fd - gd))
    return gd, out
