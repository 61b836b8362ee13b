"""Batched mixed-model fitting: voxels of several models in one persistent launch.

Each model keeps the reference's per-model interface (lsqnonlin.py:1036-1073:
parameters updated in place, lam/step_type/f/first_f/last_f state), but all
models are compiled into one module whose pca_lm_chunk kernel serves one
packed work queue per model (csrc/lm_kernel.hip). A chunk launch walks the
64-voxel wave tasks of every queue, so models with different parameter and
measurement counts -- and voxels with very different iteration counts -- share
the GPU in the same launches; converged voxels leave the queues between
chunks. The launch sequence (state reset, optional initial-guess kernels, all
chunks) can be captured in one hipGraph. Per-voxel results are identical to
fitting each model on its own.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import chunked, lm
from .lsqnonlin import (F, Expr, _as_rows, _check_device_tensor, split_consts)


@dataclass
class Member:
    """One model's voxels and state inside a batch."""
    spec: lm.LMSpec
    pars_t: torch.Tensor = None
    data_t: torch.Tensor = None
    lower_bound_t: torch.Tensor = None
    upper_bound_t: torch.Tensor = None
    uc_t: torch.Tensor = None
    uc: list = None
    lam_t: torch.Tensor = None
    step_type_t: torch.Tensor = None
    f_t: torch.Tensor = None
    first_f: torch.Tensor = None
    last_f: torch.Tensor = None
    n_iter_t: torch.Tensor = None
    ws: chunked.Workset = None

    @property
    def Nprobs(self):
        return 0 if self.data_t is None else int(self.data_t.shape[1])

    def lm_args(self, iters, tol, mu, eta, acc, dec):
        from . import _lib
        P = lm.ptr
        return _lib.LMArgs(
            data=P(self.data_t), consts=None, lb=P(self.lower_bound_t),
            ub=P(self.upper_bound_t), params=P(self.pars_t), lam=P(self.lam_t),
            step_type=P(self.step_type_t), f=P(self.f_t), first_f=P(self.first_f),
            last_f=P(self.last_f), n_iter=P(self.n_iter_t), g_out=None, h_out=None,
            hl_out=None, step_out=None, N=self.Nprobs, iters=int(iters), reserved=0,
            tol=float(tol), mu=float(mu), eta=float(eta), acc=float(acc), dec=float(dec))


def _as_expr(m):
    if isinstance(m, Expr):
        return m
    if hasattr(m, "expr") and hasattr(m, "pars"):        # models.ModelSpec
        return Expr(m.expr, list(m.pars), list(m.consts))
    expr, pars, consts = m
    return Expr(expr, list(pars), list(consts))


class BatchedLM:
    """models: list of (model, ndata), model an Expr, a models.ModelSpec or an
    (expr, pars, consts) triple; at most 4 models. Consts must be
    voxel-uniform (one protocol per model). Knobs as SecondOrderLevenbergMarquardt's
    chunked path: block, chunk, max_chunks, vpl, unroll, fast_exp, fast_div, strict."""

    def __init__(self, models, dtype=torch.float32, block: int = lm.DEFAULT_BLOCK,
                 chunk: int = 8, max_chunks: int = 0, fast_exp: bool = False,
                 fast_div: bool = False,
                 strict: bool = False, unroll: int = 0, waves_per_eu: int = 0,
                 reassoc: bool = False, sched_group: int = 0, factored: bool = False,
                 vpl: int = 1, compaction: bool = True, exp_rec: int = 0,
                 uc_lds: bool = False, **_):
        if not 1 <= len(models) <= 4:
            raise ValueError("a batch holds 1 to 4 models")
        self.dtype_name = lm.dtype_name(dtype)
        self.dtype = lm.torch_dtype(self.dtype_name)
        self.block, self.chunk, self.max_chunks = int(block), int(chunk), int(max_chunks)
        specs = []
        for m, ndata in models:
            e = _as_expr(m)
            specs.append(lm.LMSpec(e.expr, tuple(e.pars_str), tuple(e.consts_str), int(ndata),
                                   self.dtype_name, self.block, False, bool(fast_exp),
                                   bool(strict), unroll=int(unroll),
                                   waves_per_eu=int(waves_per_eu), fast_div=bool(fast_div),
                                   reassoc=bool(reassoc), sched_group=int(sched_group),
                                   factored=bool(factored), vpl=int(vpl),
                                   exp_rec=int(exp_rec), uc_lds=bool(uc_lds)))
        self.members = [Member(s) for s in specs]
        self.program = lm.get_program(specs[0], tuple(specs[1:]))
        self.funcid = self.program.funcid
        self._run = None
        self._key = None

    # -- setup ------------------------------------------------------------------
    def setup(self, problems):
        """problems: one (pars_t, consts_t, data_t, lower_bound_t, upper_bound_t)
        per model, in the reference's layouts; pars_t is updated in place."""
        if len(problems) != len(self.members):
            raise ValueError(f"expected {len(self.members)} problems, got {len(problems)}")
        for mi, (mem, (pars_t, consts_t, data_t, lb_t, ub_t)) in enumerate(
                zip(self.members, problems)):
            s = mem.spec
            for t, nm in ((pars_t, "pars_t"), (data_t, "data_t"), (lb_t, "lower_bound_t"),
                          (ub_t, "upper_bound_t"), (consts_t, "consts_t")):
                _check_device_tensor(t, nm, self.dtype)
            d = data_t[0] if data_t.dim() == 3 and data_t.shape[0] == 1 else data_t
            if d.dim() != 2 or d.shape[0] != s.M:
                raise ValueError(f"data_t has shape {tuple(data_t.shape)}, expected ({s.M}, N)")
            N = int(d.shape[1])
            uc, pv = split_consts(consts_t, s.nconst, s.M, N)
            if pv is not None:
                raise ValueError("batched fits need voxel-uniform consts")
            mem.data_t = _as_rows(d, s.M, N, "data_t")
            mem.pars_t = _as_rows(pars_t, s.nparam, N, "pars_t")
            mem.lower_bound_t = _as_rows(lb_t, s.nparam, N, "lower_bound_t")
            mem.upper_bound_t = _as_rows(ub_t, s.nparam, N, "upper_bound_t")
            dev = mem.data_t.device
            kw = dict(dtype=self.dtype, device=dev)
            mem.uc = uc
            # the chunked kernel reads the model's derived consts rows too
            mem.uc_t = torch.tensor(self.program.uconsts(uc, mi), **kw)
            mem.lam_t = torch.full((1, N), 5.0, **kw)
            mem.step_type_t = torch.ones((1, N), dtype=torch.int8, device=dev)
            mem.f_t = torch.zeros((1, N), **kw)
            mem.first_f = torch.zeros((1, N), **kw)
            mem.last_f = torch.zeros((1, N), **kw)
            mem.n_iter_t = torch.zeros((1, N), dtype=torch.int32, device=dev)
            if mem.ws is None or mem.ws.cap < max(N, 1):
                mem.ws = chunked.Workset(N, s.nparam, self.dtype, dev, M=s.M)
        if self._run is not None:
            self._run.destroy_graph()
        self._run, self._key = None, None

    @property
    def Nprobs(self):
        return sum(m.Nprobs for m in self.members)

    def _chunked(self, tol, mu, eta, acc, dec):
        key = (float(tol), float(mu), float(eta), float(acc), float(dec))
        if self._run is not None and self._key == key:
            return self._run
        if self.members[0].data_t is None:
            raise RuntimeError("setup() must be called before run()")
        jobs = [(self.program, m.lm_args(0, *key), m.uc_t, m.ws) for m in self.members]
        self._run = chunked.ChunkedRun(self.program, jobs, self.members[0].data_t.device,
                                       self.block)
        self._key = key
        return self._run

    # -- fit --------------------------------------------------------------------
    def run(self, iters: int, tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0):
        """Run up to ``iters`` LM iterations on every voxel of every model
        (asynchronous, on the current stream); resumable like the single-model
        solver."""
        if iters <= 0:
            return None
        r = self._chunked(tol, mu, eta, acc, dec)
        r.set_iters(iters)
        r.launch(chunked.chunk_schedule(int(iters), self.chunk, self.max_chunks))
        return None

    def reset_state(self, stream=None):
        """lambda 5, step_type 1, f 0, n_iter 0 for every model (one kernel per
        model on ``stream`` or the current stream)."""
        from .lsqnonlin import launch_reset_state
        s = lm.stream_handle() if stream is None else stream
        for m in self.members:
            launch_reset_state(self.program, m.lm_args(0, 1e-8, .25, .75, .2, 2.0), m.Nprobs, s)

    def capture_graph(self, iters: int, tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0,
                      init=None):
        """Capture optionally init(stream) (kernels writing the initial guesses
        into the members' pars_t) and the whole chunked fit, starting from the
        setup() state as after reset_state(), into one hipGraph; replay it with
        replay_graph()."""
        r = self._chunked(tol, mu, eta, acc, dec)
        r.set_iters(iters)
        r.capture(chunked.chunk_schedule(int(iters), self.chunk, self.max_chunks), init,
                  fresh=True)
        return r

    def replay_graph(self):
        if self._run is None or self._run._graph is None:
            raise RuntimeError("capture_graph() first")
        self._run.replay()

    # -- results ----------------------------------------------------------------
    def final_f(self):
        """SSR at every model's current parameters, concatenated in model order
        as a (1, sum N) tensor."""
        outs = []
        for m in self.members:
            prog = lm.get_program(m.spec)
            out = torch.zeros((1, m.Nprobs), dtype=self.dtype, device=m.data_t.device)
            ucbuf, ucbytes = lm.uconsts_buffer(m.uc, self.dtype_name)
            F.launch(prog, m.pars_t, m.data_t, None, ucbuf, ucbytes, None, out, self.block)
            outs.append(out)
        return torch.cat(outs, dim=1)

    def executed_iterations(self):
        return sum(int(m.n_iter_t.sum()) for m in self.members)
