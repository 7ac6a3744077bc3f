"""Grids, wavenumbers, dealiasing and the FFT conventions of the reference.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Array layout.  Julia arrays are column-major with x fastest; the same bytes
are a C-order NumPy array with the axes reversed.  So a 2D physical field
``(nx, ny)`` in Julia is ``[ny][nx]`` here and a spectral field
``(nkr, ny)`` is ``[ny][nkr]``; 3D is ``[nz][ny][nx]`` / ``[nz][ny][nkr]``.
"""

from __future__ import annotations

import math
import os

import numpy as np
import scipy.fft as sfft

WORKERS = int(os.environ.get("FFORACLE_WORKERS", os.cpu_count() or 1))


def fftfreq(n, scale):
    """Julia ``fftfreq(n, fs)``: ``fs/n * [0..n/2-1, -n/2..-1]`` (src/domains.jl:193)."""
    return np.fft.fftfreq(n, 1.0 / scale)


def rfftfreq(n, scale):
    """Julia ``rfftfreq(n, fs)``: ``fs/n * [0..n/2]``, Nyquist positive (src/domains.jl:195)."""
    return np.fft.rfftfreq(n, 1.0 / scale)


def getaliasedwavenumbers(nk, nkr, aliased_fraction):
    """src/domains.jl:408-421.  Returns 1-based inclusive (lo, hi) pairs, or None.

    ``iL = floor((1-f)/2 * nk) + 1``, ``iR = ceil((1+f)/2 * nk)``;
    ``kalias = iL:iR`` and ``kralias = iL:nkr``.
    """
    if not aliased_fraction < 1:
        raise ValueError("`aliased_fraction` must be less than 1")
    L = (1 - aliased_fraction) / 2
    R = (1 + aliased_fraction) / 2
    iL = math.floor(L * nk) + 1
    iR = math.ceil(R * nk)
    if aliased_fraction > 0:
        return (iL, iR), (iL, nkr)
    return None, None


def _zero_range(a, axis, rng):
    """Zero the 1-based inclusive index range ``rng`` of ``a`` along ``axis``."""
    if rng is None:
        return
    lo, hi = rng
    if hi < lo:
        return
    sl = [slice(None)] * a.ndim
    sl[axis] = slice(lo - 1, hi)
    a[tuple(sl)] = 0


class _Grid:
    ndim = 0

    def __init__(self, ns, Ls, x0s, T, aliased_fraction):
        for n in ns:
            if n % 2 != 0:
                raise DomainError("grid sizes must be even")
        self.T = np.dtype(T)
        self.aliased_fraction = aliased_fraction
        self.ns = tuple(int(n) for n in ns)
        self.Ls = tuple(float(L) for L in Ls)
        self.ds = tuple(L / n for n, L in zip(self.ns, self.Ls))
        self.x0s = tuple(float(x0) for x0 in x0s)

    # shapes in NumPy (C) order
    @property
    def phys_shape(self):
        return tuple(reversed(self.ns))

    @property
    def spec_shape(self):
        s = list(reversed(self.ns))
        s[-1] = self.ns[0] // 2 + 1
        return tuple(s)

    @property
    def cspec_shape(self):
        return self.phys_shape

    def points(self, axis):
        """``range(x0, step=dx, length=n)`` (src/domains.jl:190-191)."""
        n, d, x0 = self.ns[axis], self.ds[axis], self.x0s[axis]
        return x0 + d * np.arange(n)

    # FFTs with the reference's conventions ------------------------------
    def mul(self, f):
        """``mul!(fh, grid.rfftplan, f)``: unnormalised r2c (src/domains.jl:3,208)."""
        return sfft.rfftn(f, workers=WORKERS)

    def ldiv(self, fh):
        """``ldiv!(f, grid.rfftplan, fh)``: c2r with 1/N (AbstractFFTs scaled inverse).

        pocketfft's irfftn does the c2c transforms over the leading axes and
        then a c2r along x that ignores the imaginary part of the x-DC and
        x-Nyquist entries -- the same order FFTW uses.
        """
        return sfft.irfftn(fh, s=self.phys_shape, workers=WORKERS)

    def fft(self, f):
        """``grid.fftplan * f``: unnormalised c2c (src/domains.jl:2,207)."""
        return sfft.fftn(f, workers=WORKERS)

    def ifft(self, fh):
        return sfft.ifftn(fh, workers=WORKERS)


class DomainError(ValueError):
    """Odd grid size (src/domains.jl:66,179,316)."""


class OneDGrid(_Grid):
    """src/domains.jl:61-101."""

    ndim = 1

    def __init__(self, nx, Lx, x0=None, T=np.float64, aliased_fraction=1 / 3):
        x0 = -Lx / 2 if x0 is None else x0
        super().__init__((nx,), (Lx,), (x0,), T, aliased_fraction)
        self.nx, self.Lx, self.dx = nx, float(Lx), Lx / nx
        self.nk, self.nkr = nx, nx // 2 + 1
        self.x = self.points(0)
        self.k = fftfreq(nx, 2 * np.pi / Lx * nx)
        self.kr = rfftfreq(nx, 2 * np.pi / Lx * nx)
        with np.errstate(divide="ignore"):
            self.invksq = 1 / self.k**2
            self.invkrsq = 1 / self.kr**2
        self.invksq[0] = 0
        self.invkrsq[0] = 0
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)


class TwoDGrid(_Grid):
    """src/domains.jl:175-223."""

    ndim = 2

    def __init__(self, nx, Lx, ny=None, Ly=None, x0=None, y0=None, T=np.float64,
                 aliased_fraction=1 / 3):
        ny = nx if ny is None else ny
        Ly = Lx if Ly is None else Ly
        x0 = -Lx / 2 if x0 is None else x0
        y0 = -Ly / 2 if y0 is None else y0
        super().__init__((nx, ny), (Lx, Ly), (x0, y0), T, aliased_fraction)
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, float(Lx), float(Ly)
        self.dx, self.dy = Lx / nx, Ly / ny
        self.nk, self.nl, self.nkr = nx, ny, nx // 2 + 1
        self.x = self.points(0)
        self.y = self.points(1)
        # NumPy order: k varies along the last axis, l along the first
        self.k = fftfreq(nx, 2 * np.pi / Lx * nx)[None, :]
        self.l = fftfreq(ny, 2 * np.pi / Ly * ny)[:, None]
        self.kr = rfftfreq(nx, 2 * np.pi / Lx * nx)[None, :]
        self.Ksq = self.k**2 + self.l**2
        self.Krsq = self.kr**2 + self.l**2
        with np.errstate(divide="ignore"):
            self.invKsq = 1 / self.Ksq
            self.invKrsq = 1 / self.Krsq
        self.invKsq[0, 0] = 0
        self.invKrsq[0, 0] = 0
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)
        self.lalias, _ = getaliasedwavenumbers(self.nl, self.nl, aliased_fraction)


class ThreeDGrid(_Grid):
    """src/domains.jl:311-366."""

    ndim = 3

    def __init__(self, nx, Lx, ny=None, Ly=None, nz=None, Lz=None, x0=None, y0=None,
                 z0=None, T=np.float64, aliased_fraction=1 / 3):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        x0 = -Lx / 2 if x0 is None else x0
        y0 = -Ly / 2 if y0 is None else y0
        z0 = -Lz / 2 if z0 is None else z0
        super().__init__((nx, ny, nz), (Lx, Ly, Lz), (x0, y0, z0), T, aliased_fraction)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.dx, self.dy, self.dz = Lx / nx, Ly / ny, Lz / nz
        self.nk, self.nl, self.nm, self.nkr = nx, ny, nz, nx // 2 + 1
        self.x, self.y, self.z = self.points(0), self.points(1), self.points(2)
        self.k = fftfreq(nx, 2 * np.pi / Lx * nx)[None, None, :]
        self.l = fftfreq(ny, 2 * np.pi / Ly * ny)[None, :, None]
        self.m = fftfreq(nz, 2 * np.pi / Lz * nz)[:, None, None]
        self.kr = rfftfreq(nx, 2 * np.pi / Lx * nx)[None, None, :]
        self.Ksq = self.k**2 + self.l**2 + self.m**2
        self.Krsq = self.kr**2 + self.l**2 + self.m**2
        with np.errstate(divide="ignore"):
            self.invKsq = 1 / self.Ksq
            self.invKrsq = 1 / self.Krsq
        self.invKsq[0, 0, 0] = 0
        self.invKrsq[0, 0, 0] = 0
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)
        self.lalias, _ = getaliasedwavenumbers(self.nl, self.nl // 2 + 1, aliased_fraction)
        self.malias, _ = getaliasedwavenumbers(self.nm, self.nm // 2 + 1, aliased_fraction)


def dealias(fh, grid):
    """``dealias!(fh, grid)`` in place (src/domains.jl:428-476).

    The x range is ``kralias`` when the leading Julia dimension is ``nkr``
    and ``kalias`` otherwise (:437, :451, :465).  No-op when
    ``aliased_fraction == 0`` (:434).
    """
    if grid.aliased_fraction == 0:
        return fh
    xlen = fh.shape[-1]
    kal = grid.kralias if xlen == grid.nkr else grid.kalias
    _zero_range(fh, fh.ndim - 1, kal)
    if grid.ndim >= 2:
        _zero_range(fh, fh.ndim - 2, grid.lalias)
    if grid.ndim >= 3:
        _zero_range(fh, fh.ndim - 3, grid.malias)
    return fh


def kept_mask(grid, spectral_shape=None):
    """Boolean mask of the modes ``dealias!`` keeps (True = kept)."""
    shape = grid.spec_shape if spectral_shape is None else spectral_shape
    m = np.ones(shape, dtype=bool)
    dealias(m, grid)
    return m


def makefilter(grid, realvars=True, order=4, innerK=2 / 3, outerK=1, tol=1e-15, dims=None):
    """src/domains.jl:506-543: exp(-decay (K - innerK)^order) with
    decay = -log(tol)/(outerK - innerK)^order and 1 where K < innerK, K the
    wavenumber magnitude scaled by dx/pi per axis (realvars: rfft layout).
    ``dims`` (C order) broadcasts it to an equation's shape
    (``makefilter(g, T, sz)``, :541)."""
    ks = [(grid.kr if realvars else grid.k) * grid.dx / np.pi]
    if grid.ndim > 1:
        ks.append(grid.l * grid.dy / np.pi)
    if grid.ndim > 2:
        ks.append(grid.m * grid.dz / np.pi)
    K = np.sqrt(sum(k**2 for k in ks)) if grid.ndim > 1 else np.abs(ks[0])
    decay = -np.log(tol) / (outerK - innerK) ** order
    filt = np.exp(-decay * (K - innerK) ** order)
    filt = np.where(K < innerK, 1.0, filt)
    if dims is not None:
        filt = np.ones(dims) * filt
    return filt

