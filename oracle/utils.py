"""Parseval sums and the Jacobian helper of the reference.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

from .grids import WORKERS


def _weights_rfft(shape, nkr):
    """Weight 1 on the kx = 0 and kx = nx/2 entries, 2 elsewhere (src/utils.jl:114-117)."""
    w = np.full(shape[-1], 2.0)
    w[0] = 1.0
    w[nkr - 1] = 1.0
    return w


def _norm(grid):
    if grid.ndim == 1:
        return grid.Lx / grid.nx**2
    if grid.ndim == 2:
        return grid.Lx * grid.Ly / (grid.nx**2 * grid.ny**2)
    return grid.Lx * grid.Ly * grid.Lz / (grid.nx**2 * grid.ny**2 * grid.nz**2)


def parsevalsum2(uh, grid):
    """src/utils.jl:113-139 (2D and 1D forms; 3D follows the same pattern).

    Sum of |uh|^2 with the r2c double counting, times L/(n^2) per dimension.
    Summed with math.fsum (exactly rounded) so it serves as a reference for
    compensated GPU reductions.
    """
    a2 = np.abs(uh) ** 2
    if uh.shape[-1] == grid.nkr:
        a2 = a2 * _weights_rfft(uh.shape, grid.nkr)
    return math.fsum(a2.ravel().tolist()) * _norm(grid)


def parsevalsum(uh, grid):
    """src/utils.jl:157-183: real part of the weighted sum of uh."""
    if uh.shape[-1] == grid.nkr:
        s = uh * _weights_rfft(uh.shape, grid.nkr)
    else:
        s = uh
    re = math.fsum(np.real(s).ravel().tolist())
    return re * _norm(grid)


def jacobianh(a, b, grid):
    """src/utils.jl:190-205."""
    if not np.iscomplexobj(a):
        bh = sfft.rfftn(b, workers=WORKERS)
        bx = sfft.irfftn(1j * grid.kr * bh, s=grid.phys_shape, workers=WORKERS)
        by = sfft.irfftn(1j * grid.l * bh, s=grid.phys_shape, workers=WORKERS)
        return (1j * grid.kr * sfft.rfftn(a * by, workers=WORKERS)
                - 1j * grid.l * sfft.rfftn(a * bx, workers=WORKERS))
    bh = sfft.fftn(b, workers=WORKERS)
    bx = sfft.ifftn(1j * grid.k * bh, workers=WORKERS)
    by = sfft.ifftn(1j * grid.l * bh, workers=WORKERS)
    return (1j * grid.k * sfft.fftn(a * by, workers=WORKERS)
            - 1j * grid.l * sfft.fftn(a * bx, workers=WORKERS))


def jacobian(a, b, grid):
    """src/utils.jl:212-218."""
    if not np.iscomplexobj(a):
        return sfft.irfftn(jacobianh(a, b, grid), s=grid.phys_shape, workers=WORKERS)
    return sfft.ifftn(jacobianh(a, b, grid), workers=WORKERS)


def enstrophy(zh, grid):
    """½ ⟨ζ²⟩ = ½ parsevalsum2(ζ̂)/(Lx Ly)."""
    return 0.5 * parsevalsum2(zh, grid) / (grid.Lx * grid.Ly)


def energy(zh, grid):
    """½ ⟨|u|²⟩ = ½ parsevalsum2(|k| ψ̂)/(Lx Ly), ψ̂ = -invKrsq ζ̂."""
    psih = -grid.invKrsq * zh
    return 0.5 * parsevalsum2(np.sqrt(grid.Krsq) * psih, grid) / (grid.Lx * grid.Ly)


def rel_l2(a, b):
    """Norm-wise relative difference ‖a-b‖/max(‖a‖,‖b‖) (Julia isapprox semantics)."""
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    m = max(na, nb)
    return 0.0 if m == 0 else d / m


def isapprox(a, b, rtol=None, atol=0.0):
    """Julia ``isapprox`` on arrays: ‖a-b‖ ≤ max(atol, rtol·max(‖a‖,‖b‖))."""
    a, b = np.asarray(a), np.asarray(b)
    if rtol is None:
        ft = np.finfo(np.result_type(a.real.dtype, b.real.dtype, np.float32))
        rtol = 0.0 if atol > 0 else math.sqrt(ft.eps)
    d = np.linalg.norm(a - b)
    return d <= max(atol, rtol * max(np.linalg.norm(a), np.linalg.norm(b)))
