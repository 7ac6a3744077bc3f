"""Grids, FFT plans and dealiasing (src/domains.jl).

``grid.rfftplan`` / ``grid.fftplan`` are libffhip plans; ``mul`` and
``ldiv`` are ``mul!`` and ``ldiv!`` on them.  Wavenumber arrays follow the
reference exactly (src/domains.jl:193-203) and live on the device in ``T``.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .core import HIP, CPU, device_array, fltype, init, prec_of, ptr, stream_ptr

DomainError = _lib.DomainError


class Plan:
    """An FFT plan (``plan_flows_rfft`` / ``plan_flows_fft``, src/domains.jl:2-5).

    ``ns`` are the transform sizes in Julia order (x first); ``kind`` is
    "rfft" or "fft"; ``batch`` transforms are stored back to back.
    """

    def __init__(self, ns, T, kind="rfft", batch=1, device=0):
        init(device)
        self.ns = tuple(int(n) for n in ns)
        self.kind = kind
        self.batch = int(batch)
        self.T = fltype(T)
        self.prec = prec_of(T)
        h = C.c_void_p()
        fn = "ffhip_plan_rfft" if kind == "rfft" else "ffhip_plan_fft"
        _lib.call(fn, C.byref(h), len(self.ns), _lib.i64arr(self.ns), self.batch, self.prec)
        self._h = h

    @property
    def handle(self):
        return self._h

    def forward(self, out, inp):
        _lib.call("ffhip_plan_forward", self._h, ptr(inp), ptr(out), stream_ptr())
        return out

    def inverse(self, out, inp):
        _lib.call("ffhip_plan_inverse", self._h, ptr(inp), ptr(out), stream_ptr())
        return out

    def forward_adjoint(self, out, inp):
        """``mul!^T``: spectral ``inp`` -> physical ``out`` = prod(n) ldiv!(inp / W)."""
        _lib.call("ffhip_plan_forward_adjoint", self._h, ptr(inp), ptr(out), stream_ptr())
        return out

    def inverse_adjoint(self, out, inp):
        """``ldiv!^T``: physical ``inp`` -> spectral ``out`` = W mul!(inp) / prod(n)."""
        _lib.call("ffhip_plan_inverse_adjoint", self._h, ptr(inp), ptr(out), stream_ptr())
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.ffhip_plan_destroy(h)
            self._h = None


def mul(out, plan, inp):
    """``mul!(out, plan, inp)``: unnormalised forward transform."""
    return plan.forward(out, inp)


def ldiv(out, plan, inp):
    """``ldiv!(out, plan, inp)``: inverse transform scaled by 1/N (input kept)."""
    return plan.inverse(out, inp)


def getaliasedwavenumbers(nk, nkr, aliased_fraction):
    """src/domains.jl:408-421 (computed by libffhip's host code).

    Returns ``(kalias, kralias)`` as 1-based inclusive ``(lo, hi)`` pairs, or
    ``(None, None)`` when ``aliased_fraction == 0``.
    """
    v = [C.c_int64() for _ in range(4)]
    _lib.call("ffhip_aliased_wavenumbers", nk, nkr, float(aliased_fraction), *[C.byref(x) for x in v])
    klo, khi, krlo, krhi = (x.value for x in v)
    if aliased_fraction == 0:
        return None, None
    return (klo, khi), (krlo, krhi)


def _fftfreq(n, L):
    return np.fft.fftfreq(n, 1.0 / (2 * np.pi / L * n))


def _rfftfreq(n, L):
    return np.fft.rfftfreq(n, 1.0 / (2 * np.pi / L * n))


class AbstractGrid:
    ndim = 0

    def _setup(self, dev, ns, Ls, x0s, T, aliased_fraction):
        if any(n % 2 != 0 for n in ns):
            raise DomainError(_lib.ERR_DOMAIN, "grid sizes must be even")
        self.device = dev
        self.T = fltype(T)
        self.ns = tuple(int(n) for n in ns)
        self.Ls = tuple(float(L) for L in Ls)
        self.x0s = tuple(float(x) for x in x0s)
        self.aliased_fraction = float(aliased_fraction)
        self.nkr = self.ns[0] // 2 + 1
        self._rfftplan = None
        self._fftplan = None
        self._conv = device_array(dev)

    def _dev(self, a):
        return self._conv(np.asarray(a, dtype=self.T))

    def points(self, d):
        n, L, x0 = self.ns[d], self.Ls[d], self.x0s[d]
        return (x0 + (L / n) * np.arange(n)).astype(self.T)

    @property
    def phys_dims(self):
        return self.ns

    @property
    def spec_dims(self):
        return (self.nkr,) + self.ns[1:]

    @property
    def rfftplan(self):
        if self._rfftplan is None:
            self._need_hip()
            self._rfftplan = Plan(self.ns, self.T, "rfft", device=self.device.index)
        return self._rfftplan

    @property
    def fftplan(self):
        if self._fftplan is None:
            self._need_hip()
            self._fftplan = Plan(self.ns, self.T, "fft", device=self.device.index)
        return self._fftplan

    def _need_hip(self):
        if not isinstance(self.device, HIP):
            raise RuntimeError("FFT plans run on HIP devices only (construct the grid with HIP())")

    def alias_ranges(self, xlen):
        """Per-dimension 1-based (lo, hi) ranges used by ``dealias!``."""
        kal = self.kralias if xlen == self.nkr else self.kalias
        rs = [kal]
        if self.ndim >= 2:
            rs.append(self.lalias)
        if self.ndim >= 3:
            rs.append(self.malias)
        return [(1, 0) if r is None else r for r in rs]


class OneDGrid(AbstractGrid):
    """src/domains.jl:61-101."""

    ndim = 1

    def __init__(self, dev=None, *, nx, Lx, x0=None, nthreads=None, effort=None, T=np.float64,
                 aliased_fraction=1 / 3):
        dev = HIP() if dev is None else dev
        x0 = -Lx / 2 if x0 is None else x0
        self._setup(dev, (nx,), (Lx,), (x0,), T, aliased_fraction)
        self.nx, self.Lx, self.dx, self.nk = nx, float(Lx), Lx / nx, nx
        self.x = self.points(0)
        k = _fftfreq(nx, Lx)
        kr = _rfftfreq(nx, Lx)
        self.k_host, self.kr_host = k, kr
        self.k, self.kr = self._dev(k), self._dev(kr)
        with np.errstate(divide="ignore"):
            invksq, invkrsq = 1 / k**2, 1 / kr**2
        invksq[0] = invkrsq[0] = 0
        self.invksq, self.invkrsq = self._dev(invksq), self._dev(invkrsq)
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)


class TwoDGrid(AbstractGrid):
    """src/domains.jl:175-223.  Device arrays have NumPy/torch shape (ny, nkr)
    for the Julia (nkr, ny) layout."""

    ndim = 2

    def __init__(self, dev=None, *, nx, Lx, ny=None, Ly=None, x0=None, y0=None, nthreads=None, effort=None,
                 T=np.float64, aliased_fraction=1 / 3):
        dev = HIP() if dev is None else dev
        ny = nx if ny is None else ny
        Ly = Lx if Ly is None else Ly
        x0 = -Lx / 2 if x0 is None else x0
        y0 = -Ly / 2 if y0 is None else y0
        self._setup(dev, (nx, ny), (Lx, Ly), (x0, y0), T, aliased_fraction)
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, float(Lx), float(Ly)
        self.dx, self.dy = Lx / nx, Ly / ny
        self.nk, self.nl = nx, ny
        self.x, self.y = self.points(0), self.points(1)
        k = _fftfreq(nx, Lx)[None, :]
        l = _fftfreq(ny, Ly)[:, None]
        kr = _rfftfreq(nx, Lx)[None, :]
        self.k_host, self.l_host, self.kr_host = k, l, kr
        Ksq = k**2 + l**2
        Krsq = kr**2 + l**2
        with np.errstate(divide="ignore"):
            invKsq, invKrsq = 1 / Ksq, 1 / Krsq
        invKsq[0, 0] = invKrsq[0, 0] = 0
        self.Krsq_host, self.invKrsq_host = Krsq, invKrsq
        self.k, self.l, self.kr = self._dev(k), self._dev(l), self._dev(kr)
        self.Ksq, self.invKsq = self._dev(Ksq), self._dev(invKsq)
        self.Krsq, self.invKrsq = self._dev(Krsq), self._dev(invKrsq)
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)
        self.lalias, _ = getaliasedwavenumbers(self.nl, self.nl, aliased_fraction)


class ThreeDGrid(AbstractGrid):
    """src/domains.jl:311-366.  Device arrays have shape (nz, ny, nkr)."""

    ndim = 3

    def __init__(self, dev=None, *, nx, Lx, ny=None, Ly=None, nz=None, Lz=None, x0=None, y0=None, z0=None,
                 nthreads=None, effort=None, T=np.float64, aliased_fraction=1 / 3):
        dev = HIP() if dev is None else dev
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        x0 = -Lx / 2 if x0 is None else x0
        y0 = -Ly / 2 if y0 is None else y0
        z0 = -Lz / 2 if z0 is None else z0
        self._setup(dev, (nx, ny, nz), (Lx, Ly, Lz), (x0, y0, z0), T, aliased_fraction)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.dx, self.dy, self.dz = Lx / nx, Ly / ny, Lz / nz
        self.nk, self.nl, self.nm = nx, ny, nz
        self.x, self.y, self.z = self.points(0), self.points(1), self.points(2)
        k = _fftfreq(nx, Lx)[None, None, :]
        l = _fftfreq(ny, Ly)[None, :, None]
        m = _fftfreq(nz, Lz)[:, None, None]
        kr = _rfftfreq(nx, Lx)[None, None, :]
        self.k_host, self.l_host, self.m_host, self.kr_host = k, l, m, kr
        Ksq = k**2 + l**2 + m**2
        Krsq = kr**2 + l**2 + m**2
        with np.errstate(divide="ignore"):
            invKsq, invKrsq = 1 / Ksq, 1 / Krsq
        invKsq[0, 0, 0] = invKrsq[0, 0, 0] = 0
        self.Krsq_host, self.invKrsq_host = Krsq, invKrsq
        self.k, self.l, self.m, self.kr = self._dev(k), self._dev(l), self._dev(m), self._dev(kr)
        self.Ksq, self.invKsq = self._dev(Ksq), self._dev(invKsq)
        self.Krsq, self.invKrsq = self._dev(Krsq), self._dev(invKrsq)
        self.kalias, self.kralias = getaliasedwavenumbers(self.nk, self.nkr, aliased_fraction)
        self.lalias, _ = getaliasedwavenumbers(self.nl, self.nl // 2 + 1, aliased_fraction)
        self.malias, _ = getaliasedwavenumbers(self.nm, self.nm // 2 + 1, aliased_fraction)


def gridpoints(grid):
    """src/domains.jl:380-398 (NumPy arrays in C order)."""
    if grid.ndim == 1:
        return grid.x.copy()
    if grid.ndim == 2:
        X, Y = np.meshgrid(grid.x, grid.y)
        return X, Y
    Z, Y, X = np.meshgrid(grid.z, grid.y, grid.x, indexing="ij")
    return X, Y, Z


def makefilter(grid, realvars=True, order=4, innerK=2 / 3, outerK=1, tol=1e-15, dims=None):
    """src/domains.jl:506-543 (host FP64 array, C order).

    ``K`` is the wavenumber magnitude with each axis scaled by d/pi, computed
    in the grid's float type as the reference does; the filter is
    ``exp(-decay (K - innerK)^order)`` with ``decay = -log(tol)/(outerK -
    innerK)^order``, and 1 where ``K < innerK``.  ``dims`` (Julia order)
    broadcasts it to an equation's shape (``makefilter(g, T, sz)``, :541).
    """
    T = np.dtype(grid.T).type
    pi = T(np.pi)
    ks = [np.asarray(grid.kr_host if realvars else grid.k_host, dtype=T) * T(grid.dx) / pi]
    if grid.ndim > 1:
        ks.append(np.asarray(grid.l_host, dtype=T) * T(grid.dy) / pi)
    if grid.ndim > 2:
        ks.append(np.asarray(grid.m_host, dtype=T) * T(grid.dz) / pi)
    if grid.ndim == 1:
        K = np.abs(ks[0])
    else:
        acc = ks[0] ** 2
        for k in ks[1:]:
            acc = acc + k**2
        K = np.sqrt(acc)
    K = K.astype(np.float64)
    decay = -np.log(tol) / (outerK - innerK) ** order
    f = np.exp(-decay * (K - innerK) ** order)
    f = np.where(K < innerK, 1.0, f)
    if dims is not None:
        f = np.ones(tuple(reversed(tuple(dims)))) * f
    return np.ascontiguousarray(f)


def dealias(fh, grid):
    """``dealias!(fh, grid)`` on a device array (src/domains.jl:428-476).

    The x range is ``kralias`` when the leading Julia dimension equals
    ``nkr`` and ``kalias`` otherwise; a trailing extra dimension is treated
    as a batch.  No-op when ``aliased_fraction == 0`` (:434).
    """
    if grid.aliased_fraction == 0:
        return None
    shape = tuple(fh.shape)  # C order
    dims = tuple(reversed(shape))
    nd = grid.ndim
    batch = int(np.prod(dims[nd:])) if len(dims) > nd else 1
    rs = grid.alias_ranges(dims[0])
    lo = _lib.i64arr(r[0] for r in rs)
    hi = _lib.i64arr(r[1] for r in rs)
    prec = _lib.F32 if str(fh.dtype) == "torch.complex64" else _lib.F64
    _lib.call("ffhip_dealias", ptr(fh), prec, nd, _lib.i64arr(dims[:nd]), batch, lo, hi, stream_ptr())
    return None


