"""Parseval sums (src/utils.jl:113-183), compensated FP64 in libffhip."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .core import ptr, stream_ptr


def _norm(grid):
    n2 = 1.0
    vol = 1.0
    for n, L in zip(grid.ns, grid.Ls):
        n2 *= float(n) ** 2
        vol *= L
    return vol / n2


def _layout(uh, grid):
    dims = tuple(reversed(tuple(uh.shape)))
    prec = _lib.F32 if str(uh.dtype) == "torch.complex64" else _lib.F64
    if str(uh.dtype) not in ("torch.complex64", "torch.complex128"):
        raise TypeError("parseval sums take complex device arrays")
    rfft = 1 if dims[0] == grid.nkr and dims[0] != grid.ns[0] else 0
    return dims, prec, rfft


def parsevalsum2(uh, grid):
    """Σ|û|² L/(n²) per dimension, r2c modes 0 < kx < nx/2 counted twice."""
    dims, prec, rfft = _layout(uh, grid)
    out = C.c_double()
    _lib.call("ffhip_parsevalsum2", ptr(uh), prec, len(dims), _lib.i64arr(dims), rfft, grid.ns[0],
              _norm(grid), C.byref(out), stream_ptr())
    return out.value


def parsevalsum(uh, grid):
    """Re Σ û · L/(n²) per dimension, with the same r2c double counting."""
    dims, prec, rfft = _layout(uh, grid)
    out = C.c_double()
    _lib.call("ffhip_parsevalsum", ptr(uh), prec, len(dims), _lib.i64arr(dims), rfft, grid.ns[0],
              _norm(grid), C.byref(out), stream_ptr())
    return out.value


def etdcoeffs(dt, L, ncirc=32, rcirc=1.0):
    """getetdcoeffs (src/timesteppers.jl:689-718) computed by libffhip's host code.

    Host-only (no GPU needed).  Returns (zeta, alpha, beta, Gamma) as NumPy
    arrays, real when L is real.
    """
    L = np.asarray(L)
    shape = L.shape
    cplx = np.iscomplexobj(L)
    Lr = np.ascontiguousarray(L.real if cplx else L, dtype=np.float64).ravel()
    Li = np.ascontiguousarray(L.imag, dtype=np.float64).ravel() if cplx else None
    n = Lr.size
    outs = [np.empty(n) for _ in range(8 if cplx else 4)]
    ims = outs[4:] if cplx else [None] * 4
    _lib.call("ffhip_etdcoeffs", float(dt), _lib.dptr(Lr), _lib.dptr(Li), n, int(ncirc), float(rcirc),
              *[_lib.dptr(o) for o in outs[:4]], *[_lib.dptr(o) for o in ims])
    if cplx:
        return tuple((outs[i] + 1j * outs[i + 4]).reshape(shape) for i in range(4))
    return tuple(o.reshape(shape) for o in outs[:4])
