"""Seeded initial conditions shared bit-for-bit by the oracle and the HIP path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The product package
carries its own copy of the same generator (it must not import the oracle);
a test checks that both produce identical bytes.

The random stream is SplitMix64 (Steele, Lea & Flood, OOPSLA 2014), which is
trivial to restate in any language, so a Julia caller can rebuild the same
fields.
"""

from __future__ import annotations

import numpy as np

from .grids import dealias

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
    """Doubles in [0, 1): the top 53 bits of each SplitMix64 output."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0**-53


def vorticity_spectrum_amplitude(grid, seed, k0=6.0, spectrum="k4exp1"):
    """Host FP64 random-phase ζ̂ before normalisation: |ζ̂| ∝ sqrt(K E(K))."""
    K = np.sqrt(grid.Krsq)
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
    """Random-phase ζ̂ with isotropic spectrum, scaled to rms ζ, dealiased.

    spectrum "k4exp1": E(K) ∝ K⁴ exp(-(K/k0)²)      (256² and 2048² cases)
    spectrum "k4exp2": E(K) ∝ K⁴ exp(-2 (K/k0)²)    (peak at K = k0; ensembles)

    |ζ̂| ∝ sqrt(K E(K)); phases uniform from SplitMix64 over the rfft
    half-plane in C order; then irfft -> scale to rms -> rfft -> dealias.
    Returns complex128 spectral field of shape ``grid.spec_shape``.
    """
    zh = vorticity_spectrum_amplitude(grid, seed, k0, spectrum)
    z = grid.ldiv(zh)
    z *= rms / np.sqrt(np.mean(z**2))
    zh = grid.mul(z)
    zh.flat[0] = 0  # exact zero mean (the FFT leaves ~1e-14 of roundoff there)
    return dealias(zh, grid)


def ensemble_members(master_seed=2026, counts=((128, 32), (256, 24), (512, 8)),
                     nu_range=(1e-4, 1e-2), k0_range=(4.0, 10.0)):
    """Per-member (n, seed, nu, k0) for a mixed-grid ensemble.

    nu is log-uniform on nu_range, k0 uniform on k0_range; all draws come
    from one SplitMix64 stream so the list is reproducible.
    """
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


def tracer_ic(grid, seed, ngauss=16, sigma=0.3):
    """Sum of ``ngauss`` periodic Gaussians (width sigma) at uniform random centres.

    Minimum-image distance on the periodic box; returns the physical field
    (shape ``grid.phys_shape``, float64).
    """
    u = uniform01(seed, 3 * ngauss).reshape(ngauss, 3)
    Ls = np.array([grid.Lx, grid.Ly, grid.Lz])
    x0s = np.array(grid.x0s)
    centres = x0s + u * Ls
    c = np.zeros(grid.phys_shape)
    x = grid.x[None, None, :]
    y = grid.y[None, :, None]
    z = grid.z[:, None, None]
    for cx, cy, cz in centres:
        dx = (x - cx + grid.Lx / 2) % grid.Lx - grid.Lx / 2
        dy = (y - cy + grid.Ly / 2) % grid.Ly - grid.Ly / 2
        dz = (z - cz + grid.Lz / 2) % grid.Lz - grid.Lz / 2
        c += np.exp(-(dx**2) / (2 * sigma**2)) * np.exp(-(dy**2) / (2 * sigma**2)) \
            * np.exp(-(dz**2) / (2 * sigma**2))
    return c
