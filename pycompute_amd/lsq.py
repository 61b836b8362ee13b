"""bpolyfit (pycompute/cuda/lsq.py:22-30) on the GPU.

    bpolyfit(x, y, deg, dtype) -> (deg+1, N)

x: (1, m, N), (m, N) per voxel, or (m,) shared; y: (m, N). Each voxel's
coefficients c minimise ||[1, x, .., x^deg] c - y||; the reference computes
pinv(A) @ y with a batched SVD, which is the same solution for full-rank A.
The kernel (pca_bpolyfit, csrc/init_lib.hip) solves the normal equations in
double precision.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .lm import dtype_name, torch_dtype


def bpolyfit(x, y, deg: int, dtype=torch.float32):
    td = torch_dtype(dtype_name(dtype))
    if not (isinstance(y, torch.Tensor) and y.is_cuda):
        raise TypeError("y must be a torch tensor on the GPU")
    y = y.to(td).contiguous()
    m, N = y.shape
    x = torch.as_tensor(x, device=y.device).to(td)
    if x.dim() == 3:
        x = x.reshape(x.shape[-2], x.shape[-1])
    if x.dim() == 2 and x.shape[1] == 1:
        x = x[:, 0]
    per_voxel = x.dim() == 2
    if per_voxel and tuple(x.shape) != (m, N):
        raise ValueError(f"x has shape {tuple(x.shape)}, expected ({m}, {N})")
    if not per_voxel and x.numel() != m:
        raise ValueError(f"x has {x.numel()} values, expected {m}")
    x = x.contiguous()
    out = torch.empty((deg + 1, N), dtype=td, device=y.device)
    L = _lib.lib()
    L.pca_bpolyfit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                               ctypes.c_void_p]
    _lib.check(L.pca_bpolyfit(x.data_ptr(), int(per_voxel), y.data_ptr(), m, deg,
                              out.data_ptr(), N, int(td == torch.float64),
                              torch.cuda.current_stream(y.device).cuda_stream), "pca_bpolyfit")
    return out
