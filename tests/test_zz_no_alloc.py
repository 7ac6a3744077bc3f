"""The allocation-free reduction entry point (ffhip_parseval_noalloc).

It is called through the C ABI on caller-allocated tensors and must give
bit-for-bit the double that ffhip_parsevalsum2 / ffhip_parsevalsum return
for the same input, on two calls in a row (the second with a smaller input,
so stale partials of the first call would show), in FP32 and FP64.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ff, mode, uh, dims, rfft, nx, norm, out):
    from fourierflows_hip import _lib
    from fourierflows_hip.core import stream_ptr

    prec = _lib.F32 if str(uh.dtype) == "torch.complex64" else _lib.F64
    _lib.call("ffhip_parseval_noalloc", mode, uh.data_ptr(), prec, len(dims), _lib.i64arr(dims), rfft, nx, norm,
              out.data_ptr(), stream_ptr())


def _with_alloc(ff, mode, uh, dims, rfft, nx, norm):
    from fourierflows_hip import _lib
    from fourierflows_hip.core import stream_ptr

    prec = _lib.F32 if str(uh.dtype) == "torch.complex64" else _lib.F64
    res = C.c_double()
    name = "ffhip_parsevalsum2" if mode == 0 else "ffhip_parsevalsum"
    _lib.call(name, uh.data_ptr(), prec, len(dims), _lib.i64arr(dims), rfft, nx, norm, C.byref(res), stream_ptr())
    return res.value


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_matches_allocating_entry(ff, dtype, mode):
    import torch

    rng = np.random.default_rng(5 + mode)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")  # one result slot per call
    # (ny, nx): a large 2D rfft field (more partials than the second call
    # makes), then a small one; both r2c layouts weight kx = 0 and nx/2 once
    for call, (ny, nx, rfft) in enumerate([(1536, 2048, 1), (24, 40, 1)]):
        nkr = nx // 2 + 1 if rfft else nx
        a = rng.standard_normal((ny, nkr)) + 1j * rng.standard_normal((ny, nkr))
        uh = torch.as_tensor(a.astype(dtype)).cuda()
        dims = (nkr, ny)
        norm = (2 * np.pi) ** 2 / (nx**2 * ny**2)
        want = _with_alloc(ff, mode, uh, dims, rfft, nx, norm)
        _noalloc(ff, mode, uh, dims, rfft, nx, norm, out[call:])
        torch.cuda.synchronize()
        got = float(out[call].item())
        assert got == want, (call, got, want)


def test_noalloc_1d_and_3d_and_errors(ff):
    import torch

    from fourierflows_hip import _lib

    rng = np.random.default_rng(9)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for shape, nx, rfft in [((1, 1, 33), 64, 1), ((10, 12, 9), 16, 1), ((3, 40), 40, 0)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        uh = torch.as_tensor(a).cuda()
        dims = tuple(reversed(shape))
        want = _with_alloc(ff, 0, uh, dims, rfft, nx, 0.5)
        _noalloc(ff, 0, uh, dims, rfft, nx, 0.5, out)
        torch.cuda.synchronize()
        assert float(out[0].item()) == want
    uh = torch.zeros((4, 5), dtype=torch.complex128, device="cuda")
    with pytest.raises(_lib.FFHIPError):  # bad mode
        _noalloc(ff, 2, uh, (5, 4), 1, 8, 1.0, out)
    with pytest.raises(_lib.FFHIPError):  # r2c layout that does not match nx
        _noalloc(ff, 0, uh, (5, 4), 1, 10, 1.0, out)
