"""ctypes binding of libisoext_amd.so (the C ABI in include/isoext_amd.h).

The shared library is built in-tree by ``isoext_amd/build.py``.  There is no
fallback: if the library is missing or a GPU is absent, the product path
raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libisoext_amd.so"

OK = 0
E_ARG = -1
E_DEVICE = -2
E_CAPACITY = -3
E_WORKSPACE = -4
E_TOO_LARGE = -5

_lib = None

c_i64 = ctypes.c_int64
c_f = ctypes.c_float
c_int = ctypes.c_int
vp = ctypes.c_void_p
I64x3 = c_i64 * 3
F32x3 = c_f * 3
I64x4 = c_i64 * 4
I64x6 = c_i64 * 6


class IsoextError(RuntimeError):
    pass


def _declare(L):
    L.isoext_last_error.restype = ctypes.c_char_p
    L.isoext_last_error.argtypes = []
    L.isoext_version.restype = c_int
    L.isoext_mc_config_count.restype = c_int
    L.isoext_mc_auto_config.restype = c_int
    L.isoext_mc_auto_config.argtypes = [I64x3]
    L.isoext_set_timing.restype = None
    L.isoext_set_timing.argtypes = [c_int]
    L.isoext_last_kernel_ms.restype = c_f
    L.isoext_last_kernel_ms.argtypes = []
    L.isoext_mc_last_stage_ms.restype = None
    L.isoext_mc_last_stage_ms.argtypes = [c_f * 5]
    L.isoext_mc_dense_workspace_bytes.restype = c_i64
    L.isoext_mc_dense_workspace_bytes.argtypes = [I64x3, c_int, c_i64]
    L.isoext_mc_dense.restype = c_int
    L.isoext_mc_dense.argtypes = [vp, I64x3, F32x3, F32x3, c_f, c_int, c_int, vp, c_i64, c_i64,
                                  vp, c_i64, vp, c_i64, I64x6, vp]
    L.isoext_mc_spheres.restype = c_int
    L.isoext_mc_spheres.argtypes = [vp, c_int, I64x3, F32x3, F32x3, c_f, c_int, c_int, vp, c_i64, c_i64,
                                    vp, c_i64, vp, c_i64, I64x6, vp]
    L.isoext_mc_dense_keyed.restype = c_int
    L.isoext_mc_dense_keyed.argtypes = [vp, I64x3, F32x3, F32x3, c_f, c_int, c_int, vp, c_i64, c_i64,
                                        vp, c_i64, vp, c_i64, vp, I64x6, vp]
    L.isoext_uniform_points.restype = c_int
    L.isoext_uniform_points.argtypes = [I64x3, F32x3, F32x3, vp, vp]
    L.isoext_fill_f32.restype = c_int
    L.isoext_fill_f32.argtypes = [vp, c_i64, c_f, vp]
    L.isoext_write_obj.restype = c_int
    L.isoext_write_obj.argtypes = [ctypes.c_char_p, vp, c_i64, vp, c_i64, c_int]
    L.isoext_format_float.restype = c_i64
    L.isoext_format_float.argtypes = [ctypes.c_double, ctypes.c_char_p, c_i64]
    for name, restype, argtypes in EXTRA_DECLS:
        fn = getattr(L, name, None)
        if fn is not None:
            fn.restype = restype
            fn.argtypes = argtypes


# Declarations added by other modules (name, restype, argtypes).
EXTRA_DECLS: list = []


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -m isoext_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)")
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def declare(name, restype, argtypes):
    EXTRA_DECLS.append((name, restype, argtypes))
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def last_error() -> str:
    return lib().isoext_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc == OK:
        return
    msg = last_error()
    if rc == E_ARG:
        raise ValueError(msg)
    if rc == E_TOO_LARGE:
        raise MemoryError(msg)
    raise IsoextError(f"{what}: {msg}" if what else msg)


def on_device(dev):
    """Context that makes `dev` current (a no-op when it already is)."""
    import contextlib
    if dev.index is None or dev.index == torch.cuda.current_device():
        return contextlib.nullcontext()
    return torch.cuda.device(dev)


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def i64x3(v) -> I64x3:
    return I64x3(*[int(x) for x in v])


def f32x3(v) -> F32x3:
    return F32x3(*[float(x) for x in v])
