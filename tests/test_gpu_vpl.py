"""Two voxels per lane (vpl=2) in the chunked LM kernel.

A lane iterates two queue entries in lockstep (csrc/lm_kernel.hip,
iterate_n), each with the one-voxel iteration's arithmetic; a voxel that has
stopped keeps its state through selects. So in the strict build every output
must equal the one-voxel-per-lane launch bit for bit and the oracle's replay
of the generated model (oracle/genmodel.py): on the corner-case voxels (NaN
and infinite samples, zero signal, tight and pinned bounds), on sizes that
leave wavefront tasks of 128 slots partly empty, on resumed fits (voxels that
enter with step_type 0), with per-voxel consts and in a mixed-model batch.
The fast builds are held to the oracle parity bars of tests/parity.py.
"""
import numpy as np
import pytest
import torch

import oracle.genmodel as gm
import parity
from pycompute_amd import fit, synth
from pycompute_amd.batch import BatchedLM
from pycompute_amd.models import IVIM, MODELS

from test_gpu_semantics import _edge_case, _same

pytestmark = pytest.mark.gpu

KEYS = ("params", "lam", "step_type", "f", "first_f", "n_iter")


def _state(s):
    return {"params": s.pars_t.cpu().numpy(), "lam": s.lam_t.cpu().numpy()[0],
            "step_type": s.step_type_t.cpu().numpy()[0], "f": s.f_t.cpu().numpy()[0],
            "first_f": s.first_f.cpu().numpy()[0], "n_iter": s.n_iter_t.cpu().numpy()[0]}


def _assert_same(a, b, tag):
    for k in KEYS:
        same = _same(a[k], b[k])
        assert same.all(), (tag, k, np.nonzero(~same)[-1][:8])


def test_vpl2_strict_corner_cases_bit_exact(gpu):
    consts, y, lb, ub, p0 = _edge_case(seed=61)
    prob = fit.make_problem("ivim", consts, y, lb=lb, ub=ub)
    p0_t = torch.as_tensor(p0, device="cuda").contiguous()
    ref = gm.lm_run(IVIM.expr, IVIM.pars, IVIM.consts, consts, y, lb, ub, p0, 50)
    got = {}
    for vpl in (1, 2):
        print(f"[vpl corner] vpl={vpl} N={y.shape[1]}", flush=True)
        s = fit.fit(prob, iters=50, params0=p0_t.clone(), strict=True, chunk=7, block=128,
                    vpl=vpl)
        torch.cuda.synchronize()
        got[vpl] = _state(s)
        _assert_same(got[vpl], ref, f"oracle vpl={vpl}")
    _assert_same(got[2], got[1], "vpl2 vs vpl1")


@pytest.mark.parametrize("N", [1, 65, 129, 1000, 4096])
def test_vpl2_strict_sizes_and_resume(gpu, N):
    """Partial tasks (N not a multiple of 128), and a fit resumed after 15
    iterations: voxels that converged in the first run enter the second with
    step_type 0 and must come out unchanged."""
    w = synth.ivim_c1(N, seed=62)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    runs = {}
    for vpl in (1, 2):
        print(f"[vpl sizes] N={N} vpl={vpl}", flush=True)
        s = fit.fit(prob, iters=15, params0=p0.clone(), strict=True, chunk=4, vpl=vpl)
        first = _state(s)
        s.run(25)
        torch.cuda.synchronize()
        runs[vpl] = (first, _state(s))
    _assert_same(runs[2][0], runs[1][0], f"first run N={N}")
    _assert_same(runs[2][1], runs[1][1], f"resumed N={N}")
    lb, ub = prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy()
    ref = gm.lm_run(IVIM.expr, IVIM.pars, IVIM.consts, w["consts"],
                    w["data"].astype(np.float32), lb, ub, p0.cpu().numpy(), 15)
    _assert_same(runs[2][0], ref, f"oracle N={N}")


def test_vpl2_strict_per_voxel_consts(gpu):
    """Per-voxel consts travel in the queue entry of each of a lane's voxels."""
    N = 3000
    w = synth.ivim_c1(N, seed=63)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    g = torch.Generator(device="cpu").manual_seed(63)
    scale = (0.98 + 0.04 * torch.rand(prob.consts_t.numel(), N, generator=g)).to("cuda")
    cons = (prob.consts_t.reshape(-1, 1) * scale).contiguous()
    out = {}
    for vpl in (1, 2):
        print(f"[vpl per-voxel consts] vpl={vpl}", flush=True)
        s = fit.make_solver(prob, strict=True, chunk=5, vpl=vpl)
        s.setup(p0.clone(), cons, prob.data_t, prob.lb_t, prob.ub_t, uniform_consts=False)
        assert s.program.spec.consts_per_voxel and s._use_chunked()
        s.run(40)
        torch.cuda.synchronize()
        out[vpl] = _state(s)
    _assert_same(out[2], out[1], "per-voxel consts")


BUILDS = {"default": {}, "fastest": dict(fast_exp=True, fast_div=True, reassoc=True,
                                         factored=True)}


@pytest.mark.parametrize("build", sorted(BUILDS))
def test_vpl2_fast_builds_parity(gpu, build):
    w = synth.ivim(65536, synth.B_C2, seed=64, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    res = {}
    for vpl in (1, 2):
        print(f"[vpl {build}] vpl={vpl}", flush=True)
        s = fit.fit(prob, iters=100, params0=p0.clone(), chunk=16, max_chunks=4, block=128,
                    vpl=vpl, **BUILDS[build])
        torch.cuda.synchronize()
        res[vpl] = _state(s)
    ident = float(np.mean(np.all(res[2]["params"] == res[1]["params"], 0)))
    print(f"[vpl {build}] identical to vpl=1: {ident:.5f}", flush=True)
    y = prob.data_t.cpu().numpy()
    lb, ub = prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy()
    r32, r64 = parity.oracle_pair("ivim", prob.consts, y, lb, ub, p0.cpu().numpy(), 100)
    parity.check(f"ivim-64k-{build}-vpl2", "ivim", prob.consts, y, res[2]["params"], r32, r64,
                 gpu_first_f=res[2]["first_f"])


def test_vpl2_strict_mixed_batch(gpu):
    """A mixed-model batch with two voxels per lane: each member equals its
    one-voxel-per-lane batch bit for bit."""
    order = ("t2", "ivim", "se")
    w = synth.mixed(3000, 5000, 1500, seed=65)
    probs = [fit.make_problem(k, w[k]["consts"], w[k]["data"]) for k in order]
    p0 = [fit.default_init(p) for p in probs]
    members = {}
    for vpl in (1, 2):
        print(f"[vpl batch] vpl={vpl}", flush=True)
        b = BatchedLM([(MODELS[k], p.M) for k, p in zip(order, probs)], torch.float32,
                      strict=True, chunk=8, block=128, vpl=vpl)
        pars = [q.clone() for q in p0]
        b.setup([(pars[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
        b.run(100)
        torch.cuda.synchronize()
        members[vpl] = [_state(m) for m in b.members]
    for i, k in enumerate(order):
        _assert_same(members[2][i], members[1][i], f"batch member {k}")
