"""One-call fitting for the built-in models: default init + fused LM on the GPU.

This is the caller flow of test/ivim/test_ivim.ipynb (cells 5, 6, 16) as a
function: bounds from the model, the default initial guess on the GPU, then
SecondOrderLevenbergMarquardt.setup/run.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import init as _init
from .lsqnonlin import SecondOrderLevenbergMarquardt
from .models import MODELS, ModelSpec


@dataclass
class FitProblem:
    model: ModelSpec
    consts: np.ndarray          # (nconst, M) host copy (uniform)
    consts_t: torch.Tensor      # (nconst*M,) on device
    data_t: torch.Tensor        # (M, N)
    lb_t: torch.Tensor          # (n, N)
    ub_t: torch.Tensor          # (n, N)

    @property
    def N(self):
        return int(self.data_t.shape[1])

    @property
    def M(self):
        return int(self.data_t.shape[0])


def make_problem(model, consts, data, device="cuda", dtype=torch.float32, lb=None, ub=None):
    model = MODELS[model] if isinstance(model, str) else model
    consts = np.asarray(consts, np.float64)
    if consts.ndim == 1:
        consts = consts[None, :]
    if isinstance(data, torch.Tensor):
        data_t = data.to(device=device, dtype=dtype).contiguous()
    else:
        data_t = torch.as_tensor(np.ascontiguousarray(data), device=device).to(dtype).contiguous()
    N = data_t.shape[1]
    lb = np.asarray(model.lb if lb is None else lb, np.float64)
    ub = np.asarray(model.ub if ub is None else ub, np.float64)
    lb_t = torch.as_tensor(lb, device=device, dtype=dtype)
    ub_t = torch.as_tensor(ub, device=device, dtype=dtype)
    if lb_t.dim() == 1:
        lb_t = lb_t[:, None].expand(-1, N)
    if ub_t.dim() == 1:
        ub_t = ub_t[:, None].expand(-1, N)
    consts_t = torch.as_tensor(consts.reshape(-1), device=device, dtype=dtype)
    return FitProblem(model, consts, consts_t, data_t, lb_t.contiguous(), ub_t.contiguous())


def make_solver(prob: FitProblem, **kw) -> SecondOrderLevenbergMarquardt:
    m = prob.model
    return SecondOrderLevenbergMarquardt(m.expr, list(m.pars), list(m.consts), prob.M,
                                         prob.data_t.dtype, **kw)


def default_init(prob: FitProblem, out=None):
    return _init.default_init(prob.model, prob.consts, prob.data_t, prob.lb_t, prob.ub_t, out)


def fit(prob: FitProblem, iters=100, params0=None, solver=None, tol=1e-8, mu=0.25, eta=0.75,
        acc=0.2, dec=2.0, **kw):
    """Fit every voxel; returns the solver (pars_t holds the result)."""
    solver = solver or make_solver(prob, **kw)
    pars = default_init(prob) if params0 is None else params0
    solver.setup(pars, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    solver.run(iters, tol, mu, eta, acc, dec)
    return solver
