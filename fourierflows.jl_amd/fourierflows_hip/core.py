"""Device tag, array helpers and type utilities.

Mirrors the reference's device layer (src/FourierFlows.jl:94-101,
src/utils.jl:1-94, :329-340): ``HIP()`` plays the role of ``GPU()``, device
arrays are ``torch`` tensors on the ROCm device (the analogue of
``CuArray``), and every transform or pass on them runs in libffhip.
"""

from __future__ import annotations

import numpy as np

from . import _lib


class Device:
    pass


class CPU(Device):
    """Host device tag.  Accepted for signature compatibility; the compute
    entry points of this package run on the GPU only."""

    def __repr__(self):
        return "CPU()"


class HIP(Device):
    """MI355X device tag (the analogue of FourierFlows' ``GPU()``)."""

    def __init__(self, index=0):
        self.index = index

    def __repr__(self):
        return f"HIP({self.index})"


GPU = HIP

_initialised = set()


def init(device=0):
    """Select the HIP device used by libffhip (idempotent)."""
    if device in _initialised:
        return
    _lib.call("ffhip_init", device)
    _initialised.add(device)


def torch():
    import torch as _t

    return _t


def _tdtype(T):
    t = torch()
    T = np.dtype(T) if not isinstance(T, t.dtype) else T
    if isinstance(T, t.dtype):
        return T
    return {
        np.dtype(np.float32): t.float32,
        np.dtype(np.float64): t.float64,
        np.dtype(np.complex64): t.complex64,
        np.dtype(np.complex128): t.complex128,
    }[T]


def cxtype(T):
    """src/utils.jl:14-21."""
    T = np.dtype(T)
    return T if T.kind == "c" else np.result_type(T, np.complex64)


def fltype(T):
    """src/utils.jl:23-30."""
    if isinstance(T, tuple):
        return fltype(T[0])
    T = np.dtype(T)
    return np.dtype(np.float32) if T in (np.float32, np.complex64) else np.dtype(np.float64)


def prec_of(T):
    return _lib.F32 if fltype(T) == np.float32 else _lib.F64


def zeros(dev, T, dims):
    """``zeros(dev, T, dims)`` (src/utils.jl:79-80).  ``dims`` in Julia order
    (x first); the tensor has the reversed (C-order) shape, same bytes."""
    t = torch()
    shape = tuple(reversed(tuple(dims))) if isinstance(dims, (tuple, list)) else (dims,)
    if isinstance(dev, HIP):
        init(dev.index)
        return t.zeros(shape, dtype=_tdtype(T), device=f"cuda:{dev.index}")
    return t.zeros(shape, dtype=_tdtype(T))


def device_array(dev):
    """src/utils.jl:329-332: a converter from host arrays to ``dev`` arrays."""
    t = torch()

    def conv(a, dtype=None):
        x = t.as_tensor(np.ascontiguousarray(a))
        if dtype is not None:
            x = x.to(_tdtype(dtype))
        return x.to(f"cuda:{dev.index}") if isinstance(dev, HIP) else x

    return conv


def stream_ptr():
    """The current torch stream as a hipStream_t, so library work is ordered
    with the caller's torch work."""
    t = torch()
    return t.cuda.current_stream().cuda_stream


def ptr(x):
    """Device pointer of a contiguous GPU tensor; raises on anything else."""
    t = torch()
    if not isinstance(x, t.Tensor) or not x.is_cuda:
        raise TypeError("expected a GPU tensor (HIP device array)")
    if not x.is_contiguous():
        raise ValueError("device arrays passed to libffhip must be contiguous")
    return x.data_ptr()


def to_host(x):
    return x.detach().cpu().numpy()
