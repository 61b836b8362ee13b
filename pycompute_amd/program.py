"""Code assembly, hiprtc compilation with an on-disk code-object cache, and
per-device module loading.

The reference builds each kernel by walking a dependency tree of
``CudaFunction`` objects (pycompute/cuda/cuda_program.py:16-36, 184-203):
post-order over ``get_deps()``, de-duplicated by function id, device code
concatenated, then the kernel wrapper, and the result handed to
``cp.RawModule`` (NVRTC). ``HipFunction`` and ``code_gen_walking`` keep that
protocol so user plugins written against it link the same way; the source is
compiled by hiprtc for gfx950 through libpycompute_amd.so.
"""
from __future__ import annotations

import hashlib
import os
import threading

from . import _lib

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
ARCH = "gfx950"
# Packed-FP32 VALU forms are kept out of every kernel compiled here.
BASE_OPTIONS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17",
                "-fhip-fp32-correctly-rounded-divide-sqrt",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def read_csrc(name: str) -> str:
    with open(os.path.join(CSRC, name)) as fh:
        return fh.read()


class HipFunction:
    """The reference's CudaFunction protocol (cuda_program.py:16-36)."""

    def get_device_funcid(self) -> str:
        raise NotImplementedError

    def get_kernel_funcid(self) -> str:
        return "k_" + self.get_device_funcid()

    def get_device_code(self) -> str:
        raise NotImplementedError

    def get_kernel_code(self) -> str:
        return ""

    def get_deps(self) -> list:
        return []


CudaFunction = HipFunction


def gen_deps_dict(func: HipFunction, deps: dict) -> dict:
    """Post-order walk of get_deps(), keyed (de-duplicated) by device funcid
    (cuda_program.py:184-191)."""
    for d in func.get_deps():
        gen_deps_dict(d, deps)
        deps.setdefault(d.get_device_funcid(), d.get_device_code())
    return deps


def code_gen_walking(func: HipFunction, code: str = "") -> str:
    """Dependencies' device code in post order, then the function's own device
    and kernel code (cuda_program.py:193-203)."""
    deps = gen_deps_dict(func, {})
    for c in deps.values():
        code += c + "\n"
    code += func.get_device_code() + "\n"
    code += func.get_kernel_code() + "\n"
    return code


# ---------------------------------------------------------------------------
# code-object cache
# ---------------------------------------------------------------------------

def cache_dir() -> str:
    d = os.environ.get("PYCOMPUTE_AMD_KCACHE")
    if not d:
        d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_kcache")
    try:
        os.makedirs(d, exist_ok=True)
        if os.access(d, os.W_OK):
            return d
    except OSError:
        pass
    d = os.path.join(os.path.expanduser("~"), ".cache", "pycompute_amd")
    os.makedirs(d, exist_ok=True)
    return d


def _compiler_id() -> str:
    import torch
    return f"hip={torch.version.hip};abi={_lib.ABI_VERSION}"


def cache_key(source: str, options) -> str:
    h = hashlib.sha256()
    h.update(source.encode())
    h.update(b"\0" + "\0".join(options).encode())
    h.update(b"\0" + _compiler_id().encode())
    return h.hexdigest()[:32]


_USED_CACHE_FILES: set = set()


def prune_cache() -> int:
    """Delete cached code objects this process neither compiled nor loaded
    (stale sources); returns the number removed. Called at the end of a build
    that has compiled every shipped workload."""
    d = cache_dir()
    removed = 0
    for fn in os.listdir(d):
        path = os.path.join(d, fn)
        if fn.endswith(".hsaco") and path not in _USED_CACHE_FILES:
            os.remove(path)
            removed += 1
    return removed


def compile_cached(source: str, name: str, options=None) -> bytes:
    """hiprtc-compile ``source`` (gfx950), reusing a cached code object."""
    options = list(BASE_OPTIONS if options is None else options)
    path = os.path.join(cache_dir(), f"{name}-{cache_key(source, options)}.hsaco")
    _USED_CACHE_FILES.add(path)
    if os.path.exists(path):
        with open(path, "rb") as fh:
            return fh.read()
    code = _lib.compile_source(source, name, options)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "wb") as fh:
        fh.write(code)
    os.replace(tmp, path)
    return code


class Module:
    """A compiled code object, loaded lazily once per device."""

    def __init__(self, source: str, name: str, options=None):
        self.source = source
        self.name = name
        self.code = compile_cached(source, name, options)
        self._mods: dict = {}
        self._kernels: dict = {}
        self._lock = threading.Lock()

    def kernel(self, kname: str, device: int | None = None):
        import ctypes
        import torch
        if not torch.cuda.is_available():
            raise _lib.PcaError("no HIP device is visible; the fitter runs on the GPU only")
        if device is None:
            device = torch.cuda.current_device()
        key = (device, kname)
        k = self._kernels.get(key)
        if k is not None:
            return k
        with self._lock:
            L = _lib.lib()
            with torch.cuda.device(device):
                torch.cuda.current_stream(device)  # make sure the context exists
                mod = self._mods.get(device)
                if mod is None:
                    h = ctypes.c_void_p()
                    buf = ctypes.create_string_buffer(self.code, len(self.code))
                    _lib.check(L.pca_module_load(buf, len(self.code), ctypes.byref(h)),
                               f"load module {self.name}")
                    mod = h
                    self._mods[device] = mod
                kh = ctypes.c_void_p()
                _lib.check(L.pca_module_kernel(mod, kname.encode(), ctypes.byref(kh)),
                           f"kernel {kname} of {self.name}")
                self._kernels[key] = kh
                return kh


# ---------------------------------------------------------------------------
# host helpers for packed lower-triangle storage (l = i(i+1)/2 + j, j <= i),
# cuda_program.py:208-235
# ---------------------------------------------------------------------------

def compact_to_full(mat):
    """Packed (n(n+1)/2,) lower triangle -> full symmetric (n, n)."""
    import math
    import numpy as np
    m = np.asarray(mat.cpu() if hasattr(mat, "cpu") else mat)
    n = int(math.floor(math.sqrt(2 * m.shape[0])))
    out = np.empty((n, n), dtype=np.float64)
    k = 0
    for i in range(n):
        for j in range(i + 1):
            out[i, j] = out[j, i] = m[k]
            k += 1
    return out


def compact_to_LD(mat):
    """Packed L\\D factor (unit-lower L below the diagonal, D on it, as GMW81
    leaves it) -> (L, D) as full matrices."""
    import math
    import numpy as np
    m = np.asarray(mat.cpu() if hasattr(mat, "cpu") else mat)
    n = int(math.floor(math.sqrt(2 * m.shape[0])))
    L = np.zeros((n, n))
    D = np.zeros((n, n))
    k = 0
    for i in range(n):
        for j in range(i + 1):
            if i != j:
                L[i, j] = m[k]
            else:
                L[i, j] = 1.0
                D[i, j] = m[k]
            k += 1
    return L, D
