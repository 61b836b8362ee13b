"""GPU parity of the fused HIP LM fitter against the CPU oracle (same inputs,
same initial guess).

The non-strict builds are held to the per-voxel bars of tests/parity.py,
which are set against what FP32 itself does on the same inputs (the FP32
and FP64 oracles): first_f, final SSR against the FP32 oracle, parameters
against the FP64 optimum as close as the FP32 oracle's own, and a residual
tail no heavier than the FP32 oracle's. Bit-exact agreement on every voxel
is tested in the strict build (test_gpu_strict.py).
"""
import numpy as np
import pytest
import torch

import oracle
import parity
from pycompute_amd import fit, synth
from pycompute_amd.models import IVIM, T2, SE

pytestmark = pytest.mark.gpu

REL = 1e-4


def _rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def _summary(name, x):
    q = np.percentile(x, [50, 99, 100]) if x.size else [0, 0, 0]
    print(f"{name}: median {q[0]:.2e} p99 {q[1]:.2e} max {q[2]:.2e}", flush=True)


def _run_pair(model, w, iters, dtype=torch.float32, n=None, **kw):
    prob = fit.make_problem(model, w["consts"], w["data"][:, :n] if n else w["data"], dtype=dtype)
    p0 = fit.default_init(prob)
    solver = fit.fit(prob, iters=iters, params0=p0.clone(), **kw)
    torch.cuda.synchronize()
    npd = np.float32 if dtype == torch.float32 else np.float64
    ref = oracle.lm_run(prob.model.name, prob.consts, prob.data_t.cpu().numpy(),
                        prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy(), p0.cpu().numpy(), iters,
                        dtype=npd)
    gpu = {"params": solver.pars_t.cpu().numpy(), "step_type": solver.step_type_t.cpu().numpy()[0],
           "f": solver.f_t.cpu().numpy()[0], "first_f": solver.first_f.cpu().numpy()[0],
           "last_f": solver.last_f.cpu().numpy()[0], "n_iter": solver.n_iter_t.cpu().numpy()[0],
           "final_f": solver.final_f().cpu().numpy()[0]}
    ref["final_f"] = oracle.ssr(prob.model.name, prob.consts, prob.data_t.cpu().numpy(),
                                ref["params"], dtype=npd)
    return prob, p0, gpu, ref


def _check(prob, p0, gpu, tag, iters=100):
    """The per-voxel bars of tests/parity.py against the FP32 and FP64
    oracles run from the same start."""
    y = prob.data_t.double().cpu().numpy().astype(np.float32)
    lb, ub = prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy()
    r32, r64 = parity.oracle_pair(prob.model.name, prob.consts, y, lb.astype(np.float32),
                                  ub.astype(np.float32), p0.cpu().numpy().astype(np.float32),
                                  iters)
    return parity.check(tag, prob.model.name, prob.consts, y, gpu["params"], r32, r64,
                        gpu_first_f=gpu["first_f"])


def test_ivim_c1_parity(gpu):
    w = synth.ivim_c1(4096, seed=1)
    prob, p0, g, r = _run_pair("ivim", w, 100)
    _check(prob, p0, g, "ivim-c1")


def test_ivim_c2_subset_parity(gpu):
    w = synth.ivim(65536, synth.B_C2, seed=2, snr=(20.0, 100.0))
    prob, p0, g, r = _run_pair("ivim", w, 100)
    _check(prob, p0, g, "ivim-c2-64k")


def test_t2_and_se_parity(gpu):
    for model, w in (("t2", synth.t2(8192, seed=3)), ("se", synth.se(8192, seed=4))):
        prob, p0, g, r = _run_pair(model, w, 100)
        _check(prob, p0, g, model)


FAST_BUILDS = {"fast": dict(fast_exp=True, fast_div=True),
               "fast-reassoc": dict(fast_exp=True, fast_div=True, reassoc=True),
               "fast-factored": dict(fast_exp=True, fast_div=True, reassoc=True, factored=True),
               "fast-factored-fenced": dict(fast_exp=True, fast_div=True, reassoc=True,
                                            factored=True, sched_group=4),
               # exponential recurrence: every protocol here (16 b-values
               # 0-800, 32 echoes 10-320 ms, 16 b-values 50-800) is an
               # arithmetic progression
               "fast-factored-rec8": dict(fast_exp=True, fast_div=True, reassoc=True,
                                          factored=True, exp_rec=8),
               "fast-factored-fenced-rec4": dict(fast_exp=True, fast_div=True, reassoc=True,
                                                 factored=True, sched_group=4, exp_rec=4)}


@pytest.mark.parametrize("build", sorted(FAST_BUILDS))
def test_fast_builds_parity(gpu, build):
    """The autotuner's faster builds (v_exp_f32, approximate divide, FMA-folded
    Hessian update and reciprocal solve) meet the same bars."""
    kw = FAST_BUILDS[build]
    w = synth.ivim(65536, synth.B_C2, seed=2, snr=(20.0, 100.0))
    prob, p0, g, r = _run_pair("ivim", w, 100, **kw)
    _check(prob, p0, g, f"ivim-64k-{build}")
    for model, wm in (("t2", synth.t2(8192, seed=3)), ("se", synth.se(8192, seed=4))):
        prob, p0, g, r = _run_pair(model, wm, 100, **kw)
        _check(prob, p0, g, f"{model}-{build}")


def _factored_vs_folded(prob, p0, sched_group, tag):
    """Fraction of voxels whose one-iteration gradient, Hessian and damped
    Hessian agree between the factored and the FMA-folded fast builds."""
    out = {}
    for key, kw in (("ref", dict(reassoc=True)),
                    ("fac", dict(reassoc=True, factored=True, sched_group=sched_group))):
        s = fit.make_solver(prob, keep_workspace=True, fast_exp=True, fast_div=True, **kw)
        s.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
        s.run(1)
        torch.cuda.synchronize()
        out[key] = {k: getattr(s, k).double().cpu().numpy() for k in ("g_t", "h_t", "hl_t")}
        out[key]["f"] = s.first_f.double().cpu().numpy().reshape(-1)
    a, b = out["fac"], out["ref"]
    npar = a["g_t"].shape[0]
    diag = [j * (j + 1) // 2 + j for j in range(npar)]
    hd = np.abs(b["hl_t"][diag])
    ok = np.ones(a["g_t"].shape[1], bool)
    for j in range(npar):
        ok &= np.abs(a["g_t"][j] - b["g_t"][j]) <= 1e-4 * np.sqrt(hd[j] * b["f"]) + 1e-30
        ok &= np.abs(a["hl_t"][diag[j]] - b["hl_t"][diag[j]]) <= 1e-4 * np.abs(b["hl_t"][diag[j]])
        for k in range(j + 1):
            l = j * (j + 1) // 2 + k
            ok &= np.abs(a["h_t"][l] - b["h_t"][l]) <= 1e-4 * np.sqrt(hd[j] * hd[k]) + 1e-30
    print(f"[{tag} g={sched_group}] within bounds {ok.mean():.5f}", flush=True)
    return ok.mean()


@pytest.mark.parametrize("model,n", [("ivim", 16384), ("t2", 8192), ("se", 8192)])
@pytest.mark.parametrize("sched_group", [0, 4])
def test_factored_gradient_and_hessian(gpu, model, n, sched_group):
    """One iteration of the factored build leaves the same gradient, Hessian
    and damped Hessian as the FMA-folded build, up to FP32 rounding. The
    bounds are scale-aware, with D the damped diagonal (which holds lam
    sum J_j^2 and so the size of the Gauss-Newton terms even where the
    second-order terms cancel them in h_jj): |dh_jk| <= 1e-4 sqrt(D_j D_k),
    |dg_j| <= 1e-4 sqrt(D_j f), |dD_j| <= 1e-4 D_j. A wrong index or sign in
    the generated code breaks the bounds on nearly every voxel; FP32
    cancellation leaves a few per thousand outside them (both forms replayed
    in float32 on the CPU from the same start: 0.01%; the device's exp
    rounds the two differently arranged arguments differently on top)."""
    w = {"ivim": lambda: synth.ivim(n, synth.B_C2, seed=2, snr=(20.0, 100.0)),
         "t2": lambda: synth.t2(n, seed=3), "se": lambda: synth.se(n, seed=4)}[model]()
    prob = fit.make_problem(model, w["consts"], w["data"])
    assert _factored_vs_folded(prob, fit.default_init(prob), sched_group, model) >= 0.995


USER_MODELS = [
    ("a*exp(-x/b) + c", ("a", "b", "c"), (1.0, 0.5, 0.1)),
    ("a*sin(w*x + ph)**2 + log(1 + b*x)", ("a", "w", "ph", "b"), (1.0, 1.3, 0.4, 0.7)),
    ("sqrt(a*x + b)/(1 + c*x**3) - a**2", ("a", "b", "c"), (0.8, 0.6, 0.3)),
    ("p0*tanh(p1*x)^2 - p2/(x + p1)", ("p0", "p1", "p2"), (1.2, 0.9, 0.5)),
]


@pytest.mark.parametrize("expr,pars,truth", USER_MODELS, ids=[m[0] for m in USER_MODELS])
def test_factored_user_expressions(gpu, expr, pars, truth):
    """The same agreement for user expressions with sin, log, sqrt, tanh,
    powers and quotients, whose factors the generator has to find inside
    nested products (the model written through the reference's Expr form)."""
    from pycompute_amd.models import ModelSpec
    from pycompute_amd.sym import expr as E, hipgen
    n, M = 8192, 24
    spec = ModelSpec("user", expr, pars, ("x",), tuple(-10.0 for _ in pars),
                     tuple(10.0 for _ in pars))
    x = np.linspace(0.1, 2.5, M)
    rng = np.random.default_rng(11)
    P = np.array(truth)[:, None] * rng.uniform(0.8, 1.2, (len(pars), n))
    dag = hipgen.res_jac_hes(expr, pars, ("x",))
    env = {k: P[i][None, :] for i, k in enumerate(pars)}
    env["x"] = x[:, None]
    y = E.evaluate(dag.model, env) + rng.normal(0, 0.01, (M, n))
    prob = fit.make_problem(spec, x[None, :], y)
    p0 = torch.as_tensor(P * rng.uniform(0.9, 1.1, P.shape), device="cuda",
                         dtype=torch.float32).contiguous()
    assert _factored_vs_folded(prob, p0, 0, expr) >= 0.995


def test_ivim_float64_parity(gpu):
    """The FP64 build against the FP64 oracle: final SSR within 1e-10 and
    parameters within 1e-6 on nearly every voxel (FP64 trajectories diverge
    far more rarely than FP32 ones)."""
    w = synth.ivim_c1(4096, seed=7)
    prob, p0, g, r = _run_pair("ivim", w, 100, dtype=torch.float64)
    sr = _rel(g["final_f"], r["final_f"], 1e-300)
    pr = _rel(g["params"], r["params"], 1e-12).max(axis=0)
    fr = _rel(g["first_f"], r["first_f"], 1e-300)
    _summary("[ivim-f64] final SSR rel", sr)
    _summary("[ivim-f64] param rel", pr)
    assert np.mean(fr < 1e-12) >= 0.999
    assert np.mean(sr < 1e-10) >= 0.99 and np.mean(pr < 1e-6) >= 0.98


def test_default_init_matches_oracle(gpu):
    w = synth.ivim_c1(4096, seed=1)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob).cpu().numpy()
    lo, hi = fit._init.ivim_split(w["consts"][0])
    ref = oracle.ivim_default_init(w["consts"][0], w["data"], prob.lb_t.cpu().numpy(),
                                   prob.ub_t.cpu().numpy(), lo, hi)
    rel = _rel(p0, ref, 1e-6)
    _summary("init rel", rel.max(axis=0))
    assert np.mean(rel.max(axis=0) < 1e-4) >= 0.999
    # the same problem built from a tensor already on the GPU
    prob_t = fit.make_problem("ivim", w["consts"], torch.as_tensor(w["data"], device="cuda"))
    assert torch.equal(prob_t.data_t, prob.data_t)
    assert np.array_equal(fit.default_init(prob_t).cpu().numpy(), p0)


def test_resume_equals_single_run(gpu):
    """run(40) then run(60) is the reference's 100 launches (state carried)."""
    w = synth.ivim_c1(4096, seed=1)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    a = fit.fit(prob, iters=100, params0=p0.clone())
    b = fit.make_solver(prob)
    b.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    b.run(40)
    b.run(60)
    torch.cuda.synchronize()
    assert torch.equal(a.pars_t, b.pars_t)
    assert torch.equal(a.step_type_t, b.step_type_t)
    assert torch.equal(a.lam_t, b.lam_t)
    assert torch.equal(a.f_t, b.f_t)
    assert torch.equal(a.n_iter_t, b.n_iter_t)


@pytest.mark.parametrize("knobs", [{}, dict(fast_exp=True, fast_div=True, reassoc=True,
                                            factored=True)], ids=["default", "factored"])
def test_per_voxel_consts_layout(gpu, knobs):
    """The reference's replicated (1, M, N) consts layout gives the same fit."""
    w = synth.ivim_c1(2048, seed=1)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    a = fit.fit(prob, iters=50, params0=p0.clone(), **knobs)
    cons = prob.consts_t.reshape(1, -1, 1).expand(1, prob.M, prob.N).contiguous()
    b = fit.make_solver(prob, **knobs)
    b.setup(p0.clone(), cons, prob.data_t, prob.lb_t, prob.ub_t, uniform_consts=False)
    assert b.program.spec.consts_per_voxel
    b.run(50)
    torch.cuda.synchronize()
    assert torch.equal(a.pars_t, b.pars_t)
    assert torch.equal(a.step_type_t, b.step_type_t)


def test_ssr_kernel_matches_oracle(gpu):
    from pycompute_amd.lsqnonlin import F
    w = synth.ivim_c1(4096, seed=1)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    pars = torch.as_tensor(w["truth"], device="cuda", dtype=torch.float32).contiguous()
    f = torch.full((1, prob.N), 0.5, device="cuda", dtype=torch.float32)
    F(IVIM.expr, list(IVIM.pars), list(IVIM.consts), prob.M, torch.float32).run(
        pars, prob.consts_t, prob.data_t, None, f)
    ref = oracle.ssr("ivim", prob.consts, w["data"], pars.cpu().numpy(), dtype=np.float64) + 0.5
    rel = _rel(f.cpu().numpy()[0], ref, 1e-30)
    _summary("ssr rel", rel)
    assert rel.max() < 1e-5


@pytest.mark.parametrize("name", ["ivim", "t2", "se"])
def test_model_eval_matches_oracle(gpu, name):
    from pycompute_amd.lsqnonlin import Expr, model_eval
    model = {"ivim": IVIM, "t2": T2, "se": SE}[name]
    w = {"ivim": synth.ivim_c1(512, seed=1), "t2": synth.t2(512, seed=3),
         "se": synth.se(512, seed=4)}[name]
    M = w["data"].shape[0]
    pars = torch.as_tensor(w["truth"], device="cuda", dtype=torch.float64).contiguous()
    cons = torch.as_tensor(w["consts"].reshape(-1), device="cuda", dtype=torch.float64)
    m, J, d2 = model_eval(Expr(model.expr, model.pars, model.consts), pars, cons, M,
                          dtype=torch.float64)
    rm, rJ, rd2 = oracle.model_eval(name, w["consts"], w["truth"], M, dtype=np.float64)
    for a, b in ((m, rm), (J, rJ), (d2, rd2)):
        a = a.cpu().numpy()
        err = np.abs(a - b) / (np.abs(b).max() + 1e-300)
        assert err.max() < 1e-11


@pytest.mark.parametrize("chunk", [1, 3, 8, 100])
def test_chunked_compaction_equals_single_launch(gpu, chunk):
    """The compacting chunked path gives bit-identical per-voxel results."""
    w = synth.ivim(32768, synth.B_C2, seed=2, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    a = fit.fit(prob, iters=100, params0=p0.clone(), compaction=False)
    b = fit.fit(prob, iters=100, params0=p0.clone(), compaction=True, chunk=chunk)
    torch.cuda.synchronize()
    for k in ("pars_t", "step_type_t", "lam_t", "f_t", "first_f", "last_f", "n_iter_t", "g_t",
              "h_t", "hl_t", "step_t"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k


def test_chunked_per_voxel_consts_and_resume(gpu):
    w = synth.ivim_c1(4096, seed=1)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    a = fit.fit(prob, iters=100, params0=p0.clone(), compaction=False)
    cons = prob.consts_t.reshape(-1, 1).expand(-1, prob.N).contiguous()
    b = fit.make_solver(prob, compaction=True, chunk=7)
    b.setup(p0.clone(), cons, prob.data_t, prob.lb_t, prob.ub_t, uniform_consts=False)
    b.run(30)
    b.run(70)
    torch.cuda.synchronize()
    for k in ("pars_t", "step_type_t", "lam_t", "f_t", "n_iter_t"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k


def test_graph_replay_equals_eager(gpu):
    """A captured fit (reset + init + chunks) replayed several times, with
    eager work in between, matches the eager fit on every replay."""
    w = synth.ivim(16384, synth.B_C2, seed=5, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    ref = fit.fit(prob, iters=100, compaction=False)
    torch.cuda.synchronize()
    pars = torch.empty_like(ref.pars_t)
    s = fit.make_solver(prob, compaction=True, chunk=8)
    s.setup(pars, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    s.capture_graph(100, init=lambda stream: fit._init.default_init(
        prob.model, prob.consts, prob.data_t, prob.lb_t, prob.ub_t, out=pars, stream=stream))
    for rep in range(3):
        pars.fill_(float("nan"))
        s.lam_t.fill_(123.0)
        torch.cuda.synchronize()
        s.replay_graph()
        torch.cuda.synchronize()
        for k in ("pars_t", "step_type_t", "lam_t", "f_t", "first_f", "n_iter_t"):
            assert torch.equal(getattr(ref, k), getattr(s, k)), (rep, k)
