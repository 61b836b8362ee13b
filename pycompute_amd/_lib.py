"""ctypes binding of libpycompute_amd.so (include/pycompute_amd.h).

torch is imported first on purpose: torch ships its own libamdhip64.so.7 and
libhiprtc.so.7, and loading this library afterwards makes the dynamic loader
bind it to those same copies, so torch's device pointers and streams are valid
here (one HIP runtime per process).

There is no fallback: a missing library raises, on every path.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must be loaded before the HIP library, see above)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpycompute_amd.so")
ABI_VERSION = 4

PCA_UC_MAX = 256

_P = ctypes.c_void_p


class LMArgs(ctypes.Structure):
    _fields_ = [("data", _P), ("consts", _P), ("lb", _P), ("ub", _P), ("params", _P),
                ("lam", _P), ("step_type", _P), ("f", _P), ("first_f", _P), ("last_f", _P),
                ("n_iter", _P), ("g_out", _P), ("h_out", _P), ("hl_out", _P), ("step_out", _P),
                ("N", ctypes.c_int64), ("iters", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("tol", ctypes.c_double), ("mu", ctypes.c_double), ("eta", ctypes.c_double),
                ("acc", ctypes.c_double), ("dec", ctypes.c_double)]


class SSRArgs(ctypes.Structure):
    _fields_ = [("params", _P), ("data", _P), ("consts", _P), ("step_type", _P), ("f", _P),
                ("N", ctypes.c_int64)]


class EvalArgs(ctypes.Structure):
    _fields_ = [("params", _P), ("consts", _P), ("m", _P), ("J", _P), ("d2", _P),
                ("N", ctypes.c_int64)]


class WS(ctypes.Structure):
    _fields_ = [("y", _P), ("cv", _P), ("lo", _P), ("hi", _P), ("p", _P), ("lam", _P),
                ("f", _P), ("f_first", _P), ("idx", _P), ("cnt", _P), ("st", _P),
                ("cap", ctypes.c_int64)]


class ModelJob(ctypes.Structure):
    _fields_ = [("a", LMArgs), ("uc", _P), ("ws", WS * 2)]


class PcaError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PcaError(f"{LIB_PATH} is not built; run `python -m pycompute_amd.build_ext` "
                       "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.pca_last_error.restype = ctypes.c_char_p
    L.pca_last_error.argtypes = []
    L.pca_abi_version.restype = ctypes.c_int32
    L.pca_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                              ctypes.c_int32, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_size_t)]
    L.pca_free.argtypes = [_P]
    L.pca_free.restype = None
    L.pca_module_load.argtypes = [_P, ctypes.c_size_t, ctypes.POINTER(_P)]
    L.pca_module_unload.argtypes = [_P]
    L.pca_module_kernel.argtypes = [_P, ctypes.c_char_p, ctypes.POINTER(_P)]
    for name, st in (("pca_lm_launch", LMArgs), ("pca_ssr_launch", SSRArgs),
                     ("pca_eval_launch", EvalArgs)):
        fn = getattr(L, name)
        fn.argtypes = [_P, ctypes.POINTER(st), _P, ctypes.c_size_t, ctypes.c_uint32, _P]
    L.pca_launch.argtypes = [_P] + [ctypes.c_uint32] * 7 + [_P, ctypes.POINTER(_P)]
    L.pca_device_arch.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]
    L.pca_lm_chunked_launch.argtypes = [_P, _P, ctypes.POINTER(ModelJob), ctypes.c_int32, _P,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _P]
    L.pca_static_scratch_bytes.restype = ctypes.c_int64
    L.pca_static_scratch_bytes.argtypes = []
    L.pca_lm_chunked_launch_static.argtypes = [
        _P, ctypes.POINTER(LMArgs), ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int32),
        ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _P]
    L.pca_graph_begin.argtypes = [_P]
    L.pca_graph_end.argtypes = [_P, ctypes.POINTER(_P)]
    L.pca_graph_launch.argtypes = [_P, _P]
    L.pca_graph_destroy.argtypes = [_P]
    for name in ("pca_compile", "pca_module_load", "pca_module_unload", "pca_module_kernel",
                 "pca_lm_launch", "pca_ssr_launch", "pca_eval_launch", "pca_launch",
                 "pca_device_arch", "pca_lm_chunked_launch", "pca_lm_chunked_launch_static",
                 "pca_graph_begin",
                 "pca_graph_end", "pca_graph_launch", "pca_graph_destroy"):
        getattr(L, name).restype = ctypes.c_int
    if L.pca_abi_version() != ABI_VERSION:
        raise PcaError(f"{LIB_PATH} has ABI {L.pca_abi_version()}, expected {ABI_VERSION}; rebuild")
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().pca_last_error().decode(errors="replace")
        raise PcaError(f"{what} failed ({rc}): {msg}")


def compile_source(source: str, name: str, options) -> bytes:
    L = lib()
    opts = [o.encode() for o in options]
    arr = (ctypes.c_char_p * len(opts))(*opts)
    code, size = _P(), ctypes.c_size_t()
    check(L.pca_compile(source.encode(), name.encode(), arr, len(opts), ctypes.byref(code),
                        ctypes.byref(size)), "hiprtc compile of " + name)
    try:
        return ctypes.string_at(code, size.value)
    finally:
        L.pca_free(code)


def device_arch(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(lib().pca_device_arch(device, buf, 256), "pca_device_arch")
    return buf.value.decode()
