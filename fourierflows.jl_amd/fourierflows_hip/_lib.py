"""ctypes binding of libffhip.so (the C ABI declared in include/ffhip.h).

The library is the product: every transform, pointwise pass, reduction and
fused step runs in it.  There is no fallback -- if the shared object is
missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FFHIP_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libffhip.so"))

OK = 0
ERR_DOMAIN = -1
ERR_ARG = -2
ERR_OOM = -3
ERR_HIP = -4
ERR_UNSUPPORTED = -5
ERR_NONFINITE = -6

F32 = 32
F64 = 64

OP_VORTICITY2D = 1
OP_TRACER3D = 2


class FFHIPError(RuntimeError):
    """A libffhip call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libffhip error {code}: {msg}")
        self.code = code


class DomainError(FFHIPError, ValueError):
    """Odd grid size (Julia's DomainError, src/domains.jl:66,179,316)."""


class NonFiniteError(FFHIPError, FloatingPointError):
    """A checked reduction met NaN or Inf."""


_vp = C.c_void_p
_i64 = C.c_int64
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_int = C.c_int
_dbl = C.c_double

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "ffhip_init": [_int],
    "ffhip_last_error": [],
    "ffhip_version": [C.POINTER(_int), C.POINTER(_int), C.POINTER(_int)],
    "ffhip_malloc": [C.POINTER(_vp), C.c_size_t],
    "ffhip_free": [_vp],
    "ffhip_memcpy_htod": [_vp, _vp, C.c_size_t],
    "ffhip_memcpy_dtoh": [_vp, _vp, C.c_size_t],
    "ffhip_memcpy_dtod": [_vp, _vp, C.c_size_t, _vp],
    "ffhip_memset_zero": [_vp, C.c_size_t, _vp],
    "ffhip_synchronize": [],
    "ffhip_plan_rfft": [C.POINTER(_vp), _int, _i64p, _i64, _int],
    "ffhip_plan_fft": [C.POINTER(_vp), _int, _i64p, _i64, _int],
    "ffhip_plan_destroy": [_vp],
    "ffhip_plan_forward": [_vp, _vp, _vp, _vp],
    "ffhip_plan_inverse": [_vp, _vp, _vp, _vp],
    "ffhip_plan_forward_adjoint": [_vp, _vp, _vp, _vp],
    "ffhip_plan_inverse_adjoint": [_vp, _vp, _vp, _vp],
    "ffhip_plan_info": [_vp, C.POINTER(_int), _i64p, _i64p, C.POINTER(_int), C.POINTER(_int)],
    "ffhip_aliased_wavenumbers": [_i64, _i64, _dbl, _i64p, _i64p, _i64p, _i64p],
    "ffhip_dealias": [_vp, _int, _int, _i64p, _i64, _i64p, _i64p, _vp],
    "ffhip_etdcoeffs": [_dbl, _dp, _dp, _i64, _int, _dbl, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp],
    "ffhip_etdrk4_create": [C.POINTER(_vp), _dp, _dp, _int, _i64p, _dbl, _int],
    "ffhip_etdrk4_destroy": [_vp],
    "ffhip_etdrk4_scratch": [_vp, _int, C.POINTER(_vp)],
    "ffhip_etdrk4_coeff": [_vp, _int, C.POINTER(_vp), C.POINTER(_int)],
    "ffhip_etdrk4_substep12": [_vp, _vp, _vp, _vp, _vp],
    "ffhip_etdrk4_substep3": [_vp, _vp, _vp, _vp, _vp, _vp],
    "ffhip_etdrk4_update": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ffhip_rk4_create": [C.POINTER(_vp), _dp, _dp, _int, _i64p, _int],
    "ffhip_rk4_destroy": [_vp],
    "ffhip_rk4_scratch": [_vp, _int, C.POINTER(_vp)],
    "ffhip_rk4_addlinearterm": [_vp, _vp, _vp, _vp],
    "ffhip_rk4_substepsol": [_vp, _vp, _vp, _vp, _dbl, _vp],
    "ffhip_rk4_update": [_vp, _vp, _vp, _vp, _vp, _vp, _dbl, _vp],
    "ffhip_parsevalsum2": [_vp, _int, _int, _i64p, _int, _i64, _dbl, _dp, _vp],
    "ffhip_parsevalsum": [_vp, _int, _int, _i64p, _int, _i64, _dbl, _dp, _vp],
    "ffhip_parseval_noalloc": [_int, _vp, _int, _int, _i64p, _int, _i64, _dbl, _vp, _vp],
    "ffhip_op_create": [C.POINTER(_vp), _int, _int, _int, _i64p, _dp, _dbl, _dp, _int],
    "ffhip_op_destroy": [_vp],
    "ffhip_op_calcN": [_vp, _vp, _vp, _dbl, _vp],
    "ffhip_op_etdrk4_run": [_vp, _vp, _vp, _i64, _vp],
    "ffhip_op_rk4_run": [_vp, _vp, _vp, _dbl, _i64, _vp],
    "ffhip_op_diagnostics": [_vp, _vp, _dp, _vp],
    "ffhip_op_is_fused": [_vp, C.POINTER(_int)],
    "ffhip_op_etdrk4_forward_ckpt": [_vp, _vp, _vp, _vp, _i64, _i64, _vp],
    "ffhip_op_etdrk4_adjoint": [_vp, _vp, _vp, _vp, _i64, _i64, _vp],
    "ffhip_op_checkpoint_id": [_vp, _vp],
    "ffhip_op_rk4_adjoint": [_vp, _vp, _vp, _vp, _dbl, _i64, _vp],
    "ffhip_op_explicit_run": [_vp, _vp, _int, _vp, _dbl, _i64, _i64, _vp, _vp, _vp],
    "ffhip_etdrk4_set_filter": [_vp, _dp],
    "ffhip_rk4_set_filter": [_vp, _dp],
    "ffhip_etdrk4_apply_filter": [_vp, _vp, _vp],
    "ffhip_rk4_apply_filter": [_vp, _vp, _vp],
    "ffhip_euler_update": [_vp, _vp, _vp, _dbl, _vp],
    "ffhip_ab3_update": [_vp, _vp, _vp, _vp, _vp, _dbl, _i64, _vp],
    "ffhip_lsrk54_stage": [_vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _vp],
    "ffhip_ensemble_create": [C.POINTER(_vp), _int, _int, _i64p, _dp, _dbl, C.POINTER(_dp), _dbl, _int],
    "ffhip_ensemble_destroy": [_vp],
    "ffhip_ensemble_run": [_vp, C.POINTER(_vp), _i64, _vp],
    "ffhip_ensemble_info": [_vp, C.POINTER(_int), C.POINTER(_int), _dp],
    "ffhip_profile_start": [],
    "ffhip_profile_stop": [C.c_char_p, C.c_size_t],
}
_RESTYPE = {"ffhip_last_error": C.c_char_p}

_lib = None


def load():
    """Load libffhip.so once; raise if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FFHIPError(ERR_HIP, f"{LIB_PATH} not found: build it with "
                                  "`python fourierflows.jl_amd/build.py`")
    # torch bundles its own HIP and HSA runtimes (same sonames as ROCm's).
    # Loading torch first makes libffhip bind to that one runtime instead of
    # pulling a second HSA runtime into the process.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - plain C users
        pass
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def last_error():
    return load().ffhip_last_error().decode(errors="replace")


def check(rc):
    if rc == OK:
        return
    msg = last_error()
    if rc == ERR_DOMAIN:
        raise DomainError(rc, msg)
    if rc == ERR_NONFINITE:
        raise NonFiniteError(rc, msg)
    raise FFHIPError(rc, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def i64arr(vals):
    vals = list(vals)
    return (C.c_int64 * max(1, len(vals)))(*vals)


def profile_start():
    call("ffhip_profile_start")


def profile_stop():
    """Per-kernel event timings since profile_start() (dict from the JSON report)."""
    import json

    buf = C.create_string_buffer(1 << 16)
    call("ffhip_profile_stop", buf, len(buf))
    return json.loads(buf.value.decode())


def dptr(a):
    """ctypes double* for a C-contiguous float64 NumPy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)
