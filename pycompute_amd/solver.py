"""Standalone batched GMW81 solver (pycompute/cuda/solver.py:425-595).

    GMW81Solver(ndim, dtype, write_to_file=False).run(mat, rhs, sol)

mat: (ndim(ndim+1)/2, N) packed lower triangles (l = i(i+1)/2 + j), rhs and
sol: (ndim, N), torch tensors on the GPU. Each system is solved as the LM step
solves its damped Hessian: the reference's pivot (the fixed rotation its
diagonal pivot produces), diagonal scaling, the GMW81 modified Cholesky and
the LDL^T substitutions (kernel pca_gmw81_batch, csrc/gmw81_kernel.hip). mat
is left unchanged (the reference factors it in place); program.compact_to_full
and compact_to_LD expand packed storage on the host.
strict=True builds without FMA contraction, so results replay bit for bit on
the CPU.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib, lm
from .program import BASE_OPTIONS, Module, read_csrc

_modules: dict = {}


class GMW81Args(ctypes.Structure):
    _fields_ = [("mat", ctypes.c_void_p), ("rhs", ctypes.c_void_p), ("sol", ctypes.c_void_p),
                ("N", ctypes.c_int64)]


def source(ndim: int, dtype: str, block: int) -> str:
    head = [f"// generated: gmw81_solver_{ndim}_{'f' if dtype == 'float32' else 'd'}",
            f"#define PCA_NP {ndim}", f"#define PCA_BLOCK {block}"]
    if dtype == "float64":
        head.append("#define PCA_REAL_DOUBLE 1")
    return "\n".join(["\n".join(head), read_csrc("strict_math.h"), read_csrc("prelude.hip"),
                      read_csrc("lm_solver.hip"), read_csrc("gmw81_kernel.hip")])


class GMW81Solver:
    def __init__(self, ndim: int, dtype=torch.float32, write_to_file: bool = False,
                 block: int = 256, strict: bool = False):
        if not 1 <= int(ndim) <= 32:
            raise ValueError("ndim must be in 1..32")
        self.ndim = int(ndim)
        self.nmat = self.ndim * (self.ndim + 1) // 2
        self.dtype_name = lm.dtype_name(dtype)
        self.dtype = lm.torch_dtype(self.dtype_name)
        self.block = int(block)
        self.funcid = f"gmw81_solver_{self.ndim}_{'f' if self.dtype_name == 'float32' else 'd'}"
        self.code = source(self.ndim, self.dtype_name, self.block)
        if write_to_file:
            with open(self.funcid + ".hip", "w") as fh:
                fh.write(self.code)
        opts = list(BASE_OPTIONS) + (["-ffp-contract=off"] if strict else [])
        key = (self.code, tuple(opts))
        if key not in _modules:
            _modules[key] = Module(self.code, self.funcid + ("_strict" if strict else ""), opts)
        self.module = _modules[key]

    def run(self, mat, rhs, sol):
        for t, nm, rows in ((mat, "mat", self.nmat), (rhs, "rhs", self.ndim),
                            (sol, "sol", self.ndim)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda):
                raise ValueError(f"{nm} must be a torch tensor on the GPU")
            if t.dtype != self.dtype or t.dim() != 2 or t.shape[0] != rows:
                raise ValueError(f"{nm}: expected ({rows}, N) {self.dtype}, got "
                                 f"{tuple(t.shape)} {t.dtype}")
            if not t.is_contiguous():
                raise ValueError(f"{nm} must be contiguous")
        N = int(mat.shape[1])
        if rhs.shape[1] != N or sol.shape[1] != N:
            raise ValueError("mat, rhs and sol must have the same number of systems")
        if N == 0:
            return
        a = GMW81Args(mat.data_ptr(), rhs.data_ptr(), sol.data_ptr(), N)
        params = (ctypes.c_void_p * 1)(ctypes.cast(ctypes.pointer(a), ctypes.c_void_p))
        grid = (N + self.block - 1) // self.block
        _lib.check(_lib.lib().pca_launch(self.module.kernel("pca_gmw81_batch"), grid, 1, 1,
                                         self.block, 1, 1, 0, lm.stream_handle(), params),
                   "pca_gmw81_batch")
