"""Mixed-model batched fits: several models' voxels in one persistent chunked
launch sequence give the same per-voxel results as fitting each model alone,
the strict batch equals the oracle's replay of every member bit for bit, and
the full-size C3 volume (400k T2 + 400k IVIM + 200k stretched exponential)
keeps those properties, with the oracle parity bars on a sample of each
model."""
import numpy as np
import pytest
import torch

import oracle.genmodel as gm
import parity
from pycompute_amd import fit, synth
from pycompute_amd.batch import BatchedLM
from pycompute_amd.models import MODELS

pytestmark = pytest.mark.gpu

ORDER = ("t2", "ivim", "se")


def _problems(sizes, seed=7):
    w = synth.mixed(*sizes, seed=seed)
    probs = [fit.make_problem(k, w[k]["consts"], w[k]["data"]) for k in ORDER]
    p0 = [fit.default_init(p) for p in probs]
    return probs, p0


def _solo(prob, p0, **kw):
    return fit.fit(prob, iters=100, params0=p0.clone(), **kw)


BUILDS = {"strict": dict(strict=True), "default": dict(strict=False),
          "factored": dict(fast_exp=True, fast_div=True, reassoc=True, factored=True,
                           waves_per_eu=2)}


@pytest.mark.parametrize("build", sorted(BUILDS))
def test_batch_equals_separate_fits(gpu, build):
    probs, p0 = _problems((3000, 5000, 1500))
    kw = dict(chunk=8, **BUILDS[build])
    b = BatchedLM([(MODELS[k], p.M) for k, p in zip(ORDER, probs)], torch.float32, **kw)
    pars = [q.clone() for q in p0]
    b.setup([(pars[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
    b.run(100)
    torch.cuda.synchronize()
    for i, (p, m) in enumerate(zip(probs, b.members)):
        ref = _solo(p, p0[i], compaction=True, **kw)
        print(ORDER[i], "identical", float((m.pars_t == ref.pars_t).all(0).float().mean()),
              flush=True)
        assert torch.equal(m.pars_t, ref.pars_t), ORDER[i]
        assert torch.equal(m.f_t, ref.f_t), ORDER[i]
        assert torch.equal(m.lam_t, ref.lam_t), ORDER[i]
        assert torch.equal(m.step_type_t, ref.step_type_t), ORDER[i]
        assert torch.equal(m.n_iter_t, ref.n_iter_t), ORDER[i]
        assert torch.equal(m.first_f, ref.first_f), ORDER[i]
        assert torch.equal(m.last_f, ref.last_f), ORDER[i]
        assert m.pars_t.data_ptr() == pars[i].data_ptr()     # updated in place


def test_batch_graph_replay_and_resume(gpu):
    probs, p0 = _problems((2000, 2000, 1000), seed=11)
    models = [(MODELS[k], p.M) for k, p in zip(ORDER, probs)]
    ref = BatchedLM(models, torch.float32, chunk=8)
    rp = [q.clone() for q in p0]
    ref.setup([(rp[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
    ref.run(100)
    # resumed: 30 + 70 iterations
    r2 = BatchedLM(models, torch.float32, chunk=8)
    pp = [q.clone() for q in p0]
    r2.setup([(pp[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
    r2.run(30)
    r2.run(70)
    # graph: reset + restore the initial guesses + 100 iterations, replayed 3 times
    g = BatchedLM(models, torch.float32, chunk=8)
    gp = [q.clone() for q in p0]
    g.setup([(gp[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])

    def init(stream):
        for p, out in zip(probs, gp):
            fit._init.default_init(p.model, p.consts, p.data_t, p.lb_t, p.ub_t, out=out,
                                   stream=stream)

    g.capture_graph(100, init=init)
    torch.cuda.synchronize()
    for rep in range(3):
        for q in gp:
            q.fill_(float("nan"))
        for m in g.members:
            m.lam_t.fill_(123.0)
        torch.cuda.synchronize()
        g.replay_graph()
        torch.cuda.synchronize()
        for i in range(3):
            assert torch.equal(g.members[i].pars_t, ref.members[i].pars_t), (rep, ORDER[i])
            assert torch.equal(g.members[i].n_iter_t, ref.members[i].n_iter_t), (rep, ORDER[i])
            assert torch.equal(g.members[i].f_t, ref.members[i].f_t), (rep, ORDER[i])
    for i in range(3):
        # resuming continues the same trajectory; the per-launch counters differ
        assert torch.equal(r2.members[i].pars_t, ref.members[i].pars_t), ORDER[i]
        assert torch.equal(r2.members[i].f_t, ref.members[i].f_t), ORDER[i]
    f = ref.final_f()
    assert f.shape == (1, 5000) and bool(torch.isfinite(f).all())


def test_strict_batch_members_equal_oracle(gpu):
    """Each member of a strict batch equals the oracle's replay of that model
    on the batch's own inputs, bit for bit (NaN matching NaN)."""
    probs, p0 = _problems((1500, 2500, 1000), seed=13)
    b = BatchedLM([(MODELS[k], p.M) for k, p in zip(ORDER, probs)], torch.float32, chunk=8,
                  strict=True)
    pars = [q.clone() for q in p0]
    b.setup([(pars[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
    b.run(100)
    torch.cuda.synchronize()
    for i, (p, m) in enumerate(zip(probs, b.members)):
        mod = p.model
        ref = gm.lm_run(mod.expr, mod.pars, mod.consts, p.consts, p.data_t.cpu().numpy(),
                        p.lb_t.cpu().numpy(), p.ub_t.cpu().numpy(), p0[i].cpu().numpy(), 100)
        got = {"params": m.pars_t, "lam": m.lam_t[0], "step_type": m.step_type_t[0],
               "f": m.f_t[0], "first_f": m.first_f[0], "last_f": m.last_f[0],
               "n_iter": m.n_iter_t[0]}
        for k, t in got.items():
            a = t.cpu().numpy()
            same = (a == ref[k]) | (np.isnan(a) & np.isnan(ref[k])) if a.dtype.kind == "f" \
                else (a == ref[k])
            assert same.all(), (ORDER[i], k, int((~same).sum()))


def test_full_size_mixed_batch(gpu):
    """The C3 volume in one batched launch sequence with the fastest build:
    identical to the three models fitted one by one with the same knobs,
    parameters inside the bounds, final SSR finite, and each model's first
    4,096 voxels within the oracle parity bars."""
    sizes = (400_000, 400_000, 200_000)
    probs, p0 = _problems(sizes, seed=7)
    kw = dict(chunk=16, fast_exp=True, fast_div=True, reassoc=True, factored=True)
    b = BatchedLM([(MODELS[k], p.M) for k, p in zip(ORDER, probs)], torch.float32, **kw)
    pars = [q.clone() for q in p0]
    b.setup([(pars[i], p.consts_t, p.data_t, p.lb_t, p.ub_t) for i, p in enumerate(probs)])
    b.run(100)
    torch.cuda.synchronize()
    f = b.final_f()
    assert bool(torch.isfinite(f).all())
    ns = 4096
    for i, (p, m) in enumerate(zip(probs, b.members)):
        ref = _solo(p, p0[i], compaction=True, **kw)
        torch.cuda.synchronize()
        assert torch.equal(m.pars_t, ref.pars_t), ORDER[i]
        assert torch.equal(m.n_iter_t, ref.n_iter_t), ORDER[i]
        assert bool((m.pars_t >= p.lb_t).all()) and bool((m.pars_t <= p.ub_t).all())
        its = m.n_iter_t.double().mean().item()
        print(f"{ORDER[i]}: {its:.2f} iterations per voxel, converged "
              f"{float((m.step_type_t == 0).double().mean()):.4f}", flush=True)
        y = np.ascontiguousarray(p.data_t[:, :ns].cpu().numpy())
        lb = np.ascontiguousarray(p.lb_t[:, :ns].cpu().numpy())
        ub = np.ascontiguousarray(p.ub_t[:, :ns].cpu().numpy())
        q0 = np.ascontiguousarray(p0[i][:, :ns].cpu().numpy())
        r32, r64 = parity.oracle_pair(ORDER[i], p.consts, y, lb, ub, q0, 100)
        parity.check(f"C3-{ORDER[i]}", ORDER[i], p.consts, y, m.pars_t[:, :ns].cpu().numpy(),
                     r32, r64, gpu_first_f=m.first_f[0, :ns].cpu().numpy())
