"""Random-search initialisers on the GPU (pycompute/cuda/lsqnonlin.py:
1076-1291): the device draws replay on the host (Philox oracle), the best
candidate is the SSR minimum, searches continue, and the multi-start LM
matches the oracle LM run from the same starts."""
import numpy as np
import pytest
import torch

import oracle
from oracle.philox import random_candidates
from pycompute_amd import fit, synth
from pycompute_amd.models import IVIM
from pycompute_amd.search import RandomSearch, SecondOrderRandomSearch

pytestmark = pytest.mark.gpu


def _problem(n, seed):
    w = synth.ivim_c1(n, seed=seed)
    return w, fit.make_problem("ivim", w["consts"], w["data"])


def test_random_search_replays_and_keeps_the_minimum(gpu):
    w, prob = _problem(512, 21)
    rs = RandomSearch(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M, torch.float32,
                      seed=1234)
    rs.setup(prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    rs.run(1)                                     # the first candidate is always taken
    torch.cuda.synchronize()
    lb, ub = prob.lb_t[:, 0].cpu().numpy(), prob.ub_t[:, 0].cpu().numpy()
    c = random_candidates(1234, prob.N, 4, range(64), lb, ub)       # (64, 4, N)
    np.testing.assert_array_equal(rs.best_pars_t.cpu().numpy(), c[0])
    rs.run(63)
    torch.cuda.synchronize()
    # SSR of every candidate on the host (FP32 restatement of the F kernel)
    y = np.ascontiguousarray(w["data"])
    f = np.stack([oracle.ssr("ivim", prob.consts, y, np.ascontiguousarray(c[k]))
                  for k in range(64)])                               # (64, N)
    best = rs.best_pars_t.cpu().numpy()
    bf = rs.best_f_t.cpu().numpy()[0]
    # the kept parameters are one of the replayed candidates, bit for bit
    match = (c == best[None]).all(axis=1)                              # (64, N)
    assert match.any(axis=0).all()
    kept = match.argmax(axis=0)
    fmin = f.min(axis=0)
    rel = (f[kept, np.arange(prob.N)] - fmin) / fmin
    print("kept candidate is the host minimum for", float(np.mean(kept == f.argmin(0))),
          "max rel gap", float(rel.max()), flush=True)
    assert rel.max() < 1e-5                # ties within device/host exp rounding only
    np.testing.assert_allclose(bf, fmin, rtol=1e-5)
    # continuing: 1 + 63 candidates == 64 in one call
    rs2 = RandomSearch(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M, torch.float32,
                       seed=1234)
    rs2.setup(prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    rs2.run(64)
    torch.cuda.synchronize()
    assert torch.equal(rs2.best_pars_t, rs.best_pars_t)
    assert torch.equal(rs2.best_f_t, rs.best_f_t)


def test_second_order_random_search_matches_oracle_multistart(gpu):
    """Strict build: every LM run from the replayed starts, and so the best
    result, is bit-identical to the oracle's replay of the generated model."""
    import oracle.genmodel as gm
    w, prob = _problem(1024, 22)
    srs = SecondOrderRandomSearch(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M,
                                  torch.float32, seed=99, strict=True)
    srs.setup(prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    srs.run(4, lm_iters=20)
    torch.cuda.synchronize()
    lb, ub = prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy()
    c = random_candidates(99, prob.N, 4, range(4), lb[:, 0], ub[:, 0])
    y = np.ascontiguousarray(w["data"])
    m = prob.model
    res, f = [], []
    for k in range(4):
        r = gm.lm_run(m.expr, m.pars, m.consts, prob.consts, y, lb, ub,
                      np.ascontiguousarray(c[k]), 20)
        res.append(r["params"])
        f.append(gm.ssr(m.expr, m.pars, m.consts, prob.consts, y, r["params"]))
    f = np.stack(f)
    best = f.argmin(axis=0)                       # first minimum, as the device keeps it
    want_p = np.stack(res)[best, :, np.arange(prob.N)].T
    got_p = srs.best_pars_t.cpu().numpy()
    got_f = srs.best_f_t.cpu().numpy()[0]
    print("multi-start identical on", float(np.mean((got_p == want_p).all(0))), flush=True)
    np.testing.assert_array_equal(got_f, f.min(axis=0))
    np.testing.assert_array_equal(got_p, want_p)
    assert torch.all(srs.best_pars_t >= prob.lb_t) and torch.all(srs.best_pars_t <= prob.ub_t)


def test_second_order_random_search_default_build(gpu):
    """The default build: best_f is the SSR of best_pars_t and never worse
    than the best of its own starting points."""
    w, prob = _problem(4096, 23)
    srs = SecondOrderRandomSearch(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M,
                                  torch.float32, seed=7)
    srs.setup(prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    srs.run(3, lm_iters=30)
    torch.cuda.synchronize()
    bf = srs.best_f_t.cpu().numpy()[0]
    g = fit.make_solver(prob)
    g.setup(srs.best_pars_t.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    np.testing.assert_allclose(g.final_f().cpu().numpy()[0], bf, rtol=1e-6)
    c = random_candidates(7, prob.N, 4, range(3), prob.lb_t[:, 0].cpu().numpy(),
                          prob.ub_t[:, 0].cpu().numpy())
    f0 = np.stack([oracle.ssr("ivim", prob.consts, np.ascontiguousarray(w["data"]),
                              np.ascontiguousarray(c[k])) for k in range(3)]).min(0)
    assert np.all(bf <= f0 * (1 + 1e-5))


@pytest.mark.parametrize("per_voxel", [False, True])
def test_batched_multistart_equals_one_start_at_a_time(gpu, per_voxel):
    """Starts batched into one LM over K x N virtual voxels give the same best
    parameters and SSR, bit for bit, as the starts run one by one; runs
    continue the stream either way (5 starts = a batch of 3, then 2)."""
    w, prob = _problem(3000, 23)
    consts = prob.consts_t
    if per_voxel:                               # b-values 0.01% apart on even voxels
        consts = consts[:, None].expand(-1, prob.N).contiguous()
        consts[:, ::2] *= 1.0001
    out = {}
    for batch in (1, 3):
        srs = SecondOrderRandomSearch(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M,
                                      torch.float32, seed=7, batch=batch)
        srs.setup(consts, prob.data_t, prob.lb_t, prob.ub_t)
        srs.run(2, lm_iters=15)
        srs.run(3, lm_iters=15)
        torch.cuda.synchronize()
        out[batch] = (srs.best_pars_t.clone(), srs.best_f_t.clone(), srs.draws)
    assert torch.equal(out[1][0], out[3][0]) and torch.equal(out[1][1], out[3][1])
    assert out[1][2] == out[3][2] == 5
