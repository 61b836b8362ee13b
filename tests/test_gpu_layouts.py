"""Consts layouts beyond the register and argument-block limits, and counter
buffers shared between a captured graph and later eager launches.

* Per-voxel consts with more than 64 values per voxel do not fit the chunked
  kernel's registers; the single-launch kernel reads them from global memory
  (pycompute/cuda/lsqnonlin.py:987-1011 takes consts as a (nconst*M, N)
  array). Strict build: bit-identical to the uniform-consts fit.
* Uniform consts beyond the 256-value kernel-argument block are replicated
  into the per-voxel layout; the strict fit equals the oracle's replay of the
  same generated model bit for bit.
* A chunked fit captured as a graph keeps its own counter buffer when a
  later eager run needs a longer one; replays still equal the eager fit.
"""
import numpy as np
import pytest
import torch

import oracle.genmodel as gm
from pycompute_amd import fit, synth
from pycompute_amd.models import IVIM

pytestmark = pytest.mark.gpu

STATE = ("pars_t", "step_type_t", "lam_t", "f_t", "first_f", "last_f", "n_iter_t")


def _ivim_many_b(N, M, seed):
    b = np.linspace(0.0, 1000.0, M)
    return synth.ivim(N, b, seed=seed, snr=50.0)


def test_per_voxel_consts_above_register_limit(gpu):
    """80 per-voxel consts (> 64): uniform and per-voxel strict fits agree."""
    N, M = 2048, 80
    w = _ivim_many_b(N, M, 41)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    a = fit.fit(prob, iters=60, params0=p0.clone(), strict=True)
    s = fit.make_solver(prob, strict=True)
    cons = prob.consts_t.reshape(-1, 1).expand(-1, N).contiguous()
    s.setup(p0.clone(), cons, prob.data_t, prob.lb_t, prob.ub_t, uniform_consts=False)
    assert s.program.spec.consts_per_voxel and not s._use_chunked()
    s.run(25)
    s.run(35)
    torch.cuda.synchronize()
    for k in ("pars_t", "step_type_t", "lam_t", "f_t", "n_iter_t"):
        assert torch.equal(getattr(a, k), getattr(s, k)), k


def test_uniform_consts_above_argument_block(gpu):
    """272 b-values (> 256 uniform consts): replicated per voxel, strict fit
    bit-identical to the oracle's replay."""
    N, M = 256, 272
    w = _ivim_many_b(N, M, 42)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = np.clip(w["truth"] * 1.2, prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy())
    p0_t = torch.as_tensor(p0, device="cuda", dtype=torch.float32).contiguous()
    s = fit.make_solver(prob, strict=True, unroll=8)
    assert s.program.spec.consts_per_voxel
    s.setup(p0_t, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    s.run(40)
    torch.cuda.synchronize()
    ref = gm.lm_run(IVIM.expr, IVIM.pars, IVIM.consts, prob.consts, prob.data_t.cpu().numpy(),
                    prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy(), p0.astype(np.float32), 40)
    np.testing.assert_array_equal(s.pars_t.cpu().numpy(), ref["params"])
    np.testing.assert_array_equal(s.step_type_t.cpu().numpy()[0], ref["step_type"])
    np.testing.assert_array_equal(s.lam_t.cpu().numpy()[0], ref["lam"])
    np.testing.assert_array_equal(s.n_iter_t.cpu().numpy()[0], ref["n_iter"])


def test_graph_survives_longer_eager_run(gpu):
    """capture_graph(16) (two chunks), then an eager run(100) (thirteen
    chunks, a longer counter buffer), then replays: each replay equals the
    eager 16-iteration fit."""
    w = synth.ivim(8192, synth.B_C2, seed=43, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    ref = fit.fit(prob, iters=16, params0=p0.clone(), compaction=False)
    torch.cuda.synchronize()
    pars = p0.clone()
    s = fit.make_solver(prob, compaction=True, chunk=8)
    s.setup(pars, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    r = s.capture_graph(16, init=lambda stream: pars.copy_(p0))
    captured = r._graph_counts
    s.run(100)
    torch.cuda.synchronize()
    assert r._counts is not captured and r._graph_counts is captured
    for rep in range(2):
        s.replay_graph()
        torch.cuda.synchronize()
        for k in ("pars_t", "step_type_t", "lam_t", "f_t", "first_f", "n_iter_t"):
            assert torch.equal(getattr(ref, k), getattr(s, k)), (rep, k)
