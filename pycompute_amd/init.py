"""Default initial guesses ("reference default") on the GPU.

The reference's IVIM initial guess is caller code (test/ivim/test_ivim.ipynb
cell 6, lines 146-165): log-linear fits over the low-b measurements [0:10)
and the high-b measurements [15:) of its 21-point protocol, then
S0 = exp(c0_lo), f = 1 - exp(c0_hi)/exp(c0_lo), D_1 = c1_lo, D_2 = c1_hi,
nan_to_num and clip to the bounds. For M = 21 the same index split is used;
for other protocols the split is by b-value at the same physical thresholds,
low: b <= 140 s/mm^2 and high: b >= 500 s/mm^2.

The fits use the pseudo-inverse of the shared design, computed once on the
host in double precision, and one kernel applies it per voxel
(pca_loglinear_init in csrc/init_lib.hip).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .models import IVIM, T2, SE, DKI, ModelSpec, dki_design

KIND = {"dki": 0, "t2": 1, "se": 2, "ivim": 3}


def ivim_split(b):
    b = np.asarray(b, np.float64).ravel()
    if b.size == 21:
        return np.arange(0, 10), np.arange(15, 21)
    lo = np.nonzero(b <= 140.0)[0]
    hi = np.nonzero(b >= 500.0)[0]
    return lo, hi


def _pinv_line(x):
    """Pseudo-inverse of [1, x] (m x 2) -> (2, m)."""
    A = np.stack([np.ones_like(x), x], axis=1)
    return np.linalg.pinv(A)


def init_matrix(model: ModelSpec, consts: np.ndarray) -> np.ndarray:
    """(K, M) matrix P with T(P log y) the default initial guess."""
    consts = np.asarray(consts, np.float64)
    if consts.ndim == 1:
        consts = consts[None, :]
    M = consts.shape[1]
    if model.name == "ivim":
        b = consts[0]
        lo, hi = ivim_split(b)
        if lo.size < 2 or hi.size < 2:
            raise ValueError("IVIM init needs at least two low-b and two high-b measurements")
        P = np.zeros((4, M))
        P[0:2, lo] = _pinv_line(-b[lo])
        P[2:4, hi] = _pinv_line(-b[hi])
        return P
    if model.name == "t2":
        return _pinv_line(-consts[0])
    if model.name == "se":
        return _pinv_line(-consts[0])
    if model.name == "dki":
        return np.linalg.pinv(dki_design(consts))
    raise ValueError(f"no default init for model {model.name!r}")


_P_CACHE: dict = {}


def loglinear_spec(model: ModelSpec, consts, device):
    """(P, kind) of the model's default guess: P (K, M) float64 on ``device``
    (cached per model, protocol and device), kind as pca_loglinear_init."""
    if model.name not in KIND:
        raise ValueError(f"no default init for model {model.name!r}")
    consts = np.asarray(consts, np.float64)
    key = (model.name, consts.tobytes(), str(device))
    P = _P_CACHE.get(key)
    if P is None:
        P = torch.tensor(init_matrix(model, consts), dtype=torch.float64, device=device)
        _P_CACHE[key] = P
    return P, KIND[model.name]


def default_init(model: ModelSpec, consts, data_t, lb_t, ub_t, out=None, stream=None):
    """Initial parameters (n, N) for ``data_t`` (M, N) on its device."""
    P, _ = loglinear_spec(model, consts, data_t.device)
    K, M = P.shape
    n = model.nparam
    N = int(data_t.shape[1])
    if data_t.shape[0] != M:
        raise ValueError(f"data has {data_t.shape[0]} measurements, consts describe {M}")
    for t in (data_t, lb_t, ub_t):
        if not t.is_contiguous():
            raise ValueError("init inputs must be contiguous (rows, N)")
    if out is None:
        out = torch.empty((n, N), dtype=data_t.dtype, device=data_t.device)
    is_double = 1 if data_t.dtype == torch.float64 else 0
    s = torch.cuda.current_stream(data_t.device).cuda_stream if stream is None else stream
    L = _lib.lib()
    L.pca_loglinear_init.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
    _lib.check(L.pca_loglinear_init(P.data_ptr(), K, M, data_t.data_ptr(), KIND[model.name], n,
                                    lb_t.data_ptr(), ub_t.data_ptr(), out.data_ptr(), N, is_double,
                                    s), "pca_loglinear_init")
    return out


__all__ = ["default_init", "loglinear_spec", "init_matrix", "ivim_split", "IVIM", "T2", "SE",
           "DKI"]


class PickArgs(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("N", ctypes.c_int64)]


def keep_better(model: ModelSpec, consts, data_t, params_t, other_t, stream=None):
    """Per voxel, replace params_t (n, N) by other_t where other_t has the
    lower SSR on data_t (or params_t's SSR is not finite): kernel
    pca_pick_init of the model's module (voxel-uniform consts)."""
    from . import lm
    consts = np.asarray(consts, np.float64)
    if consts.ndim == 1:
        consts = consts[None, :]
    M = int(data_t.shape[0])
    dt = lm.dtype_name(data_t.dtype)
    prog = lm.get_program(lm.LMSpec(model.expr, tuple(model.pars), tuple(model.consts), M, dt))
    for t in (data_t, params_t, other_t):
        if not t.is_contiguous():
            raise ValueError("keep_better inputs must be contiguous (rows, N)")
    N = int(data_t.shape[1])
    ucbuf, ucbytes = lm.uconsts_buffer(consts.reshape(-1), dt)
    a = PickArgs(params_t.data_ptr(), other_t.data_ptr(), data_t.data_ptr(), N)
    params = (ctypes.c_void_p * 2)(ctypes.cast(ctypes.pointer(a), ctypes.c_void_p),
                                   ctypes.cast(ucbuf, ctypes.c_void_p))
    s = torch.cuda.current_stream(data_t.device).cuda_stream if stream is None else stream
    block = lm.DEFAULT_BLOCK
    _lib.check(_lib.lib().pca_launch(prog.kernel("pca_pick_init"), (N + block - 1) // block, 1,
                                     1, block, 1, 1, 0, s, params), "pca_pick_init")
    return params_t
