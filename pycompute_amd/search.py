"""Random-search initialisers (pycompute/cuda/lsqnonlin.py:1076-1291).

    rs = RandomSearch(expr, pars_str, consts_str, ndata, dtype)
    rs.setup(consts_t, data_t, lower_bound_t, upper_bound_t)
    rs.run(iters)                      # best_pars_t, best_f_t

    srs = SecondOrderRandomSearch(expr, pars_str, consts_str, ndata, dtype)
    srs.setup(consts_t, data_t, lower_bound_t, upper_bound_t)
    srs.run(iters, lm_iters=20, tol=1e-8)

Per voxel, candidates are drawn uniformly in [lower_bound, upper_bound] from a
seeded counter-based stream (Philox4x32-10, candidate k of voxel v; a run
continues where the previous one stopped). RandomSearch scores every
candidate by its SSR inside one kernel launch (pca_random_search: all `iters`
candidates, best kept in registers) instead of one launch per candidate;
strict=True uses the reproducible build throughout (bit-exact host replay);
SecondOrderRandomSearch runs the fused LM (chunked, compacting) from each
random start and keeps the result with the lowest final SSR
(pca_random_params, then pca_keep_best).

Differences from the reference, by design: a candidate's score is its own
SSR -- the reference's RandomSearch accumulates every candidate's SSR into an
uninitialised, never-cleared f (D11), and its SecondOrderRandomSearch scores
the LM's last f plus the SSR at the final point. Random numbers come from the
device stream above, not CuPy's generator.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, lm
from .lsqnonlin import (SecondOrderLevenbergMarquardt, _as_rows, _bind, _check_device_tensor,
                        split_consts)

_P = ctypes.c_void_p


class RSearchArgs(ctypes.Structure):
    _fields_ = [("data", _P), ("consts", _P), ("lb", _P), ("ub", _P), ("best_p", _P),
                ("best_f", _P), ("seed", ctypes.c_uint64), ("N", ctypes.c_int64),
                ("draw0", ctypes.c_int64), ("ncand", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class RParamsArgs(ctypes.Structure):
    _fields_ = [("lb", _P), ("ub", _P), ("p", _P), ("seed", ctypes.c_uint64),
                ("N", ctypes.c_int64), ("draw", ctypes.c_int64)]


class KeepArgs(ctypes.Structure):
    _fields_ = [("cand", _P), ("data", _P), ("consts", _P), ("best_p", _P), ("best_f", _P),
                ("N", ctypes.c_int64)]


def _model_args(args, kwargs):
    """The reference's constructor (expr, pars_str, consts_str, ndata, dtype,
    write_to_file), lsqnonlin.py:1196, 1246."""
    a = _bind(args, kwargs, ("expr", "pars_str", "consts_str", "ndata", "dtype",
                             "write_to_file"), {"dtype": torch.float32, "write_to_file": False})
    return a["expr"], list(a["pars_str"]), list(a["consts_str"]), a


class _SearchBase:
    def __init__(self, *args, seed: int = 0, block: int = lm.DEFAULT_BLOCK,
                 strict: bool = False, **kwargs):
        expr, pars, consts, a = _model_args(args, kwargs)
        self.strict = bool(strict)
        self.expr, self.pars_str, self.consts_str = expr, pars, consts
        self.nparam, self.nconst = len(pars), len(consts)
        self.ndata = int(a["ndata"])
        self.dtype_name = lm.dtype_name(a["dtype"])
        self.dtype = lm.torch_dtype(self.dtype_name)
        self.write_to_file = a["write_to_file"]
        self.seed = int(seed) & ((1 << 64) - 1)
        self.block = int(block)
        self.draws = 0
        self.Nprobs = 1
        self._prog = {}

    def _program(self, per_voxel: bool):
        if per_voxel not in self._prog:
            spec = lm.LMSpec(self.expr, tuple(self.pars_str), tuple(self.consts_str), self.ndata,
                             self.dtype_name, self.block, per_voxel, strict=self.strict)
            self._prog[per_voxel] = lm.get_program(spec)
            if self.write_to_file:
                with open(spec.funcid("random_search") + ".hip", "w") as fh:
                    fh.write(self._prog[per_voxel].source)
        return self._prog[per_voxel]

    def setup(self, consts_t, data_t, lower_bound_t, upper_bound_t):
        """Bind the voxels; best_pars_t / best_f_t start empty (best_f = max/5,
        as lsqnonlin.py:1233)."""
        for t, nm in ((data_t, "data_t"), (lower_bound_t, "lower_bound_t"),
                      (upper_bound_t, "upper_bound_t"), (consts_t, "consts_t")):
            _check_device_tensor(t, nm, self.dtype)
        d = data_t[0] if data_t.dim() == 3 and data_t.shape[0] == 1 else data_t
        if d.dim() != 2 or d.shape[0] != self.ndata:
            raise ValueError(f"data_t has shape {tuple(data_t.shape)}, expected ({self.ndata}, N)")
        N = int(d.shape[1])
        self.Nprobs = N
        self.data_t = _as_rows(d, self.ndata, N, "data_t")
        self.lower_bound_t = _as_rows(lower_bound_t, self.nparam, N, "lower_bound_t")
        self.upper_bound_t = _as_rows(upper_bound_t, self.nparam, N, "upper_bound_t")
        self.consts_t = consts_t
        self._uc, self._cpv = split_consts(consts_t, self.nconst, self.ndata, N)
        # the kernels take the full consts block by value (unused, zero, for
        # per-voxel consts)
        self._ucbuf, _ = lm.uconsts_buffer(self._uc if self._uc is not None else [],
                                           self.dtype_name)
        kw = dict(dtype=self.dtype, device=self.data_t.device)
        self.best_pars_t = torch.empty((self.nparam, N), **kw)
        fmax = float(np.finfo(np.float32 if self.dtype == torch.float32 else np.float64).max)
        self.best_f_t = torch.full((1, N), fmax / 5, **kw)
        self.draws = 0

    def _launch(self, kname, args, with_uc=True):
        prog = self._program(self._cpv is not None)
        ptrs = [ctypes.cast(ctypes.pointer(args), _P)]
        if with_uc:
            ptrs.append(ctypes.cast(self._ucbuf, _P))
        params = (_P * len(ptrs))(*ptrs)
        grid = (self.Nprobs + self.block - 1) // self.block
        _lib.check(_lib.lib().pca_launch(prog.kernel(kname), grid, 1, 1, self.block, 1, 1, 0,
                                         lm.stream_handle(), params), kname)


class RandomSearch(_SearchBase):
    """Uniform random search (lsqnonlin.py:1195-1242)."""

    def run(self, iters: int):
        """Score `iters` more candidates per voxel (one launch); the best so
        far is in best_pars_t / best_f_t."""
        if iters <= 0:
            return None
        if not hasattr(self, "data_t"):
            raise RuntimeError("setup() must be called before run()")
        a = RSearchArgs(self.data_t.data_ptr(), lm.ptr(self._cpv), self.lower_bound_t.data_ptr(),
                        self.upper_bound_t.data_ptr(), self.best_pars_t.data_ptr(),
                        self.best_f_t.data_ptr(), self.seed, self.Nprobs, self.draws,
                        int(iters), 0)
        self._launch("pca_random_search", a)
        self.draws += int(iters)
        return None


class SecondOrderRandomSearch(_SearchBase):
    """Multi-start LM (lsqnonlin.py:1245-1291): each iteration starts the LM
    from a random point in the bounds and keeps the result with the lowest
    final SSR. Knobs of the LM (compaction, chunk, fast_exp, ...) are passed
    to SecondOrderLevenbergMarquardt.

    Up to `batch` starts run as one LM over batch x N virtual voxels (the
    voxels' data, consts and bounds repeated, start k in columns [k N,
    (k + 1) N)), so one compacting launch sequence serves them all instead of
    one per start. Per-voxel LM runs do not interact, and the starts are kept
    in order, so results are identical to running the starts one by one
    (batch=1). That pays off for small volumes, where one start cannot fill
    the GPU; batch=None (the default) batches up to about 2^20 virtual voxels,
    at most 16 starts."""

    def __init__(self, *args, seed: int = 0, block: int = lm.DEFAULT_BLOCK,
                 strict: bool = False, batch: int | None = None, **kwargs):
        lm_kw = {k: kwargs.pop(k) for k in list(kwargs)
                 if k in ("compaction", "chunk", "max_chunks", "fast_exp", "fast_div", "reassoc",
                          "unroll", "waves_per_eu", "sched_group", "factored", "vpl", "wide",
                          "exp_rec", "uc_lds", "tile", "mfma", "quad",
                          "quad_waves")}
        super().__init__(*args, seed=seed, block=block, strict=strict, **kwargs)
        if batch is not None and int(batch) < 1:
            raise ValueError("batch must be at least 1")
        self.batch = None if batch is None else int(batch)
        self._lm_args = dict(block=block, keep_workspace=False, strict=strict, **lm_kw)
        self._virtual = None     # (K, solver, cand_v, views) of the last batch size

    def setup(self, consts_t, data_t, lower_bound_t, upper_bound_t):
        super().setup(consts_t, data_t, lower_bound_t, upper_bound_t)
        self.cand_t = torch.empty((self.nparam, self.Nprobs), dtype=self.dtype,
                                  device=self.data_t.device)
        self._virtual = None

    def _batch_solver(self, K: int):
        """The LM over K x N virtual voxels (built once per batch size)."""
        if self._virtual is not None and self._virtual[0] == K:
            return self._virtual
        N = self.Nprobs
        lm_ = SecondOrderLevenbergMarquardt(self.expr, list(self.pars_str),
                                            list(self.consts_str), self.ndata, self.dtype,
                                            **self._lm_args)
        data_v = self.data_t.repeat(1, K)
        lb_v = self.lower_bound_t.repeat(1, K)
        ub_v = self.upper_bound_t.repeat(1, K)
        if self._cpv is None:                 # voxel-uniform: one protocol for all
            consts_v = torch.tensor(self._uc, dtype=self.dtype, device=self.data_t.device)
        else:
            consts_v = self._cpv.reshape(self.nconst * self.ndata, N).repeat(1, K)
        cand_v = torch.empty((self.nparam, K * N), dtype=self.dtype, device=self.data_t.device)
        lm_.setup(cand_v, consts_v, data_v, lb_v, ub_v)
        self._virtual = (K, lm_, cand_v, (data_v, lb_v, ub_v, consts_v))
        return self._virtual

    def run(self, iters: int, lm_iters: int = 20, tol=1e-8):
        if not hasattr(self, "data_t"):
            raise RuntimeError("setup() must be called before run()")
        left = int(iters)
        N = self.Nprobs
        most = self.batch if self.batch is not None else min(16, max(1, -(-(1 << 20) // N)))
        while left > 0:
            K = min(left, most)
            _, solver, cand_v, _ = self._batch_solver(K)
            for k in range(K):
                a = RParamsArgs(self.lower_bound_t.data_ptr(), self.upper_bound_t.data_ptr(),
                                self.cand_t.data_ptr(), self.seed, N, self.draws + k)
                self._launch("pca_random_params", a, with_uc=False)
                cand_v[:, k * N:(k + 1) * N].copy_(self.cand_t)
            solver.reset_state()
            solver.run(int(lm_iters), tol)
            for k in range(K):
                self.cand_t.copy_(cand_v[:, k * N:(k + 1) * N])
                kp = KeepArgs(self.cand_t.data_ptr(), self.data_t.data_ptr(), lm.ptr(self._cpv),
                              self.best_pars_t.data_ptr(), self.best_f_t.data_ptr(), N)
                self._launch("pca_keep_best", kp)
            self.draws += K
            left -= K
        return None
