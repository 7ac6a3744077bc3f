"""Seeded initial conditions built on the device.

Random phases come from SplitMix64 (Steele, Lea & Flood, OOPSLA 2014)
computed on the host in unsigned 64-bit integer arithmetic, so any language
can regenerate the exact stream.  The spectral shaping, inverse transform,
rms scaling, forward transform and dealiasing run on the GPU through
libffhip.
"""

from __future__ import annotations

import numpy as np

from .core import device_array, zeros
from .domains import dealias, ldiv, mul

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """Outputs ``offset+1 .. offset+n`` of SplitMix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, offset=0):
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0**-53


def vorticity_spectrum_amplitude(grid, seed, k0=6.0, spectrum="k4exp1"):
    """Host FP64 random-phase ζ̂ before normalisation: |ζ̂| ∝ sqrt(K E(K))."""
    K = np.sqrt(grid.Krsq_host)
    if spectrum == "k4exp1":
        E = K**4 * np.exp(-((K / k0) ** 2))
    elif spectrum == "k4exp2":
        E = K**4 * np.exp(-2 * (K / k0) ** 2)
    else:
        raise ValueError(spectrum)
    amp = np.sqrt(K * E)
    theta = 2 * np.pi * uniform01(seed, amp.size).reshape(amp.shape)
    zh = amp * np.exp(1j * theta)
    zh.flat[0] = 0
    return zh


def vorticity_ic(grid, seed, k0=6.0, rms=1.0, spectrum="k4exp1"):
    """Random-phase isotropic ζ̂ scaled to rms ζ and dealiased, on the device
    (complex tensor of the grid's type, shape (ny, nkr))."""
    torch = __import__("torch")
    T = grid.T
    CT = np.result_type(T, np.complex64)
    zh = device_array(grid.device)(vorticity_spectrum_amplitude(grid, seed, k0, spectrum).astype(CT))
    z = zeros(grid.device, T, grid.ns)
    ldiv(z, grid.rfftplan, zh)
    scale = rms / torch.sqrt(torch.mean(z.double() ** 2))
    z.mul_(scale.to(z.dtype))
    mul(zh, grid.rfftplan, z)
    zh.view(-1)[0] = 0
    dealias(zh, grid)
    return zh


def tracer_ic(grid, seed, ngauss=16, sigma=0.3):
    """Sum of ``ngauss`` periodic Gaussians of width ``sigma`` at centres drawn
    uniformly in the box (SplitMix64, 3 draws per centre: x, y, z), with
    minimum-image distances.  Returns the physical field on the device in the
    grid's float type (C order (nz, ny, nx)); built in FP64 as separable
    products."""
    import torch

    u = uniform01(seed, 3 * ngauss).reshape(ngauss, 3)
    dev = f"cuda:{grid.device.index}"
    axes = []
    for d in range(3):
        n, L, x0 = grid.ns[d], grid.Ls[d], grid.x0s[d]
        axes.append(torch.as_tensor(x0 + (L / n) * np.arange(n), dtype=torch.float64, device=dev))
    c = torch.zeros(tuple(reversed(grid.ns)), dtype=torch.float64, device=dev)
    for g in range(ngauss):
        f = []
        for d in range(3):
            L, x0 = grid.Ls[d], grid.x0s[d]
            cen = x0 + u[g, d] * L
            dx = torch.remainder(axes[d] - cen + L / 2, L) - L / 2
            f.append(torch.exp(-(dx**2) / (2 * sigma**2)))
        c += f[2][:, None, None] * f[1][None, :, None] * f[0][None, None, :]
    return c.to(torch.float32 if np.dtype(grid.T) == np.float32 else torch.float64)


def ensemble_members(master_seed=2026, counts=((128, 32), (256, 24), (512, 8)), nu_range=(1e-4, 1e-2),
                     k0_range=(4.0, 10.0)):
    """Per-member (n, seed, nu, k0) of a mixed-grid ensemble: nu log-uniform
    on ``nu_range``, k0 uniform on ``k0_range``, seeds and draws from one
    SplitMix64 stream (reproducible in any language)."""
    total = sum(c for _, c in counts)
    u = uniform01(master_seed, 3 * total).reshape(total, 3)
    out = []
    i = 0
    lo, hi = np.log(nu_range[0]), np.log(nu_range[1])
    for n, c in counts:
        for _ in range(c):
            seed = int(splitmix64(master_seed + 7919, 1, offset=i)[0] & np.uint64(0x7FFFFFFF))
            nu = float(np.exp(lo + (hi - lo) * u[i, 0]))
            k0 = float(k0_range[0] + (k0_range[1] - k0_range[0]) * u[i, 1])
            out.append((n, seed, nu, k0))
            i += 1
    return out
