"""The wide-model kernels on the GPU. csrc/lm_glm.hip: the 22-parameter
kurtosis model against the oracle, tiling and resumption invariance, and a
linear model whose fit must reach the least-squares solution.
csrc/lm_glm_mfma.hip (the GLM Hessian on the matrix cores): the kurtosis
model against the oracle and against the VALU kernel, the first iteration's
gradient and step, resumption, ragged sizes and voxel independence, and the
linear model's least-squares solution.
csrc/lm_wide.hip (any model with n > 8, the Hessian kept in LDS): strict
builds bit for bit against the oracle's replay of the same generated model
(9- and 12-parameter non-linear models), resumption and ragged sizes, and the
default build within the parity bars."""
import numpy as np
import pytest
import torch

import oracle
import oracle.genmodel as gm
import parity
from pycompute_amd import fit, mlp, synth
from pycompute_amd import init as _init
from pycompute_amd.lsqnonlin import SecondOrderLevenbergMarquardt
from pycompute_amd.sym import glm

pytestmark = pytest.mark.gpu

TRIG_PARS = ["a0"] + [f"{c}{k}" for k in range(1, 6) for c in ("a", "b")]
TRIG_EXPR = "a0 + " + " + ".join(f"a{k}*cos({k}*t) + b{k}*sin({k}*t)" for k in range(1, 6))


def _dki(n, seed=6):
    w = synth.dki(n, seed=seed)
    prob = fit.make_problem("dki", w["consts"], w["data"])
    return w, prob, fit.default_init(prob)


def test_dki_wide_kernel_parity(gpu):
    n = 2048
    w, prob, p0 = _dki(n)
    s = fit.fit(prob, iters=100, params0=p0.clone(), wide=True)
    assert s._glm_tab is not None
    f = s.final_f().double().cpu().numpy()[0]
    y = np.ascontiguousarray(w["data"])
    lb, ub, pp = (t.cpu().numpy() for t in (prob.lb_t, prob.ub_t, p0))
    r32 = oracle.lm_run("dki", w["consts"], y, lb, ub, pp, 100)
    f32 = oracle.ssr("dki", w["consts"], y, r32["params"])
    d = np.float64
    r64 = oracle.lm_run("dki", w["consts"], y.astype(d), lb.astype(d), ub.astype(d), pp.astype(d),
                        100, dtype=d)
    f64 = oracle.ssr("dki", w["consts"], y.astype(d), r64["params"], dtype=d)
    a32 = np.mean(np.abs(f - f32) / f32 <= 1e-4)
    a64 = np.mean(np.abs(f - f64) / f64 <= 1e-4)
    o = np.mean(np.abs(f32 - f64) / f64 <= 1e-4)
    print(f"DKI wide: final SSR within 1e-4 of FP32 oracle {a32:.4f}, FP64 oracle {a64:.4f} "
          f"(FP32 oracle vs FP64 {o:.4f})", flush=True)
    assert a32 >= 0.98 and a64 >= o - 0.01
    # the same fit through the generic register kernel
    g = fit.fit(prob, iters=100, params0=p0.clone(), wide=False, compaction=False)
    fg = g.final_f().double().cpu().numpy()[0]
    assert np.mean(np.abs(f - fg) / fg <= 1e-4) >= 0.98
    assert torch.all(s.pars_t >= prob.lb_t) and torch.all(s.pars_t <= prob.ub_t)


MFMA_BUILDS = {"default": {}, "fast": dict(fast_exp=True, fast_div=True, reassoc=True)}


@pytest.mark.parametrize("build", sorted(MFMA_BUILDS))
def test_dki_mfma_kernel_parity(gpu, build):
    """The matrix-core GLM kernel against the FP32 and FP64 oracles (same
    bars as the VALU kernel) and against the VALU kernel of the same build."""
    n = 4096
    kw = MFMA_BUILDS[build]
    w, prob, p0 = _dki(n, seed=16)
    print(f"[dki mfma {build}] N={n}", flush=True)
    s = fit.fit(prob, iters=100, params0=p0.clone(), wide=True, mfma=True, **kw)
    assert s._glm is not None and s._glm.mfma
    f = s.final_f().double().cpu().numpy()[0]
    y = np.ascontiguousarray(w["data"])
    lb, ub, pp = (t.cpu().numpy() for t in (prob.lb_t, prob.ub_t, p0))
    r32 = oracle.lm_run("dki", w["consts"], y, lb, ub, pp, 100)
    f32 = oracle.ssr("dki", w["consts"], y, r32["params"])
    d = np.float64
    r64 = oracle.lm_run("dki", w["consts"], y.astype(d), lb.astype(d), ub.astype(d), pp.astype(d),
                        100, dtype=d)
    f64 = oracle.ssr("dki", w["consts"], y.astype(d), r64["params"], dtype=d)
    ok = np.isfinite(f64) & (f64 < 1e9)       # the default guess fails a few voxels everywhere
    a32 = np.mean((np.abs(f - f32) / f32 <= 1e-4)[ok])
    a64 = np.mean((np.abs(f - f64) / f64 <= 1e-4)[ok])
    o = np.mean((np.abs(f32 - f64) / f64 <= 1e-4)[ok])
    v = fit.fit(prob, iters=100, params0=p0.clone(), wide=True, **kw)
    fv = v.final_f().double().cpu().numpy()[0]
    av = np.mean((np.abs(f - fv) / fv <= 1e-4)[ok])
    worse = np.mean(((f - f64) / f64 > 1e-2)[ok])
    print(f"[dki mfma {build}] final SSR within 1e-4 of FP32 oracle {a32:.4f}, FP64 oracle "
          f"{a64:.4f} (FP32 oracle vs FP64 {o:.4f}), VALU kernel {av:.4f}; >1% above FP64 "
          f"{worse:.5f}", flush=True)
    assert a32 >= 0.98 and a64 >= o - 0.01 and av >= 0.98
    assert worse <= 2e-3
    assert torch.all(s.pars_t >= prob.lb_t) and torch.all(s.pars_t <= prob.ub_t)


def test_dki_mfma_first_iteration_gradient_and_step(gpu):
    """One iteration from the default guess: g and the step of the
    matrix-core kernel and of the VALU kernel against a float64 restatement
    (design matrix checked against the oracle's model, h = sum_i w_i A_i A_i^T,
    the damping floor, the oracle's GMW81 solve). The matrix-core step may be
    no further from it than the VALU kernel's."""
    n = 1024
    w, prob, p0 = _dki(n, seed=17)
    s = fit.fit(prob, iters=1, params0=p0.clone(), wide=True, mfma=True, keep_workspace=True)
    v = fit.fit(prob, iters=1, params0=p0.clone(), wide=True, keep_workspace=True)
    torch.cuda.synchronize()
    M, npar = prob.M, 22
    tab = glm.coefficient_table(s._glm.form, s._glm.dag, w["consts"])
    A, a0 = tab[:, :npar], tab[:, 2 * npar]
    P = p0.double().cpu().numpy()
    m = np.exp(a0[:, None] + A @ P)
    m_or, J_or, _ = oracle.model_eval("dki", w["consts"], P, M)
    np.testing.assert_allclose(m, m_or, rtol=1e-9)
    np.testing.assert_allclose(m[None] * A.T[:, :, None], J_or, rtol=1e-8,
                               atol=1e-9 * np.abs(J_or).max())
    y = np.ascontiguousarray(w["data"]).astype(np.float32).astype(np.float64)
    r = m - y
    g = A.T @ (m * r)
    H = np.einsum("in,ij,ik->njk", m * m + m * r, A, A)
    hd = np.maximum(5.0 * (m * m)[:, :, None] * (A * A)[:, None, :], 1e-4).sum(0)
    hl = H.copy()
    hl[:, np.arange(npar), np.arange(npar)] += hd
    step = np.stack([oracle.gmw81_solve(hl[k], g[:, k]) for k in range(n)], 1)
    err = {}
    for name, sol in (("mfma", s), ("valu", v)):
        gg = sol.g_t.double().cpu().numpy()
        st = sol.step_t.double().cpu().numpy()
        rg = np.abs(gg - g).max(0) / np.maximum(np.abs(g).max(0), 1e-30)
        rs = np.linalg.norm(st - step, axis=0) / np.maximum(np.linalg.norm(step, axis=0), 1e-30)
        err[name] = rs
        print(f"[dki first iteration {name}] g rel err median {np.median(rg):.2e} max "
              f"{rg.max():.2e}; step rel err median {np.median(rs):.2e} p99 "
              f"{np.percentile(rs, 99):.2e}", flush=True)
        assert np.median(rg) < 1e-4
    assert np.median(err["mfma"]) <= 4 * np.median(err["valu"]) + 1e-5
    assert np.percentile(err["mfma"], 99) <= 4 * np.percentile(err["valu"], 99) + 1e-4


def test_dki_mfma_resume_sizes_and_independence(gpu):
    """Resumed fits (40 + 60 iterations) equal one 100-iteration fit bit for
    bit, and fits of 1, 65 and 300 voxels (partial wavefronts and
    workgroups) equal the same voxels of a 4096-voxel fit."""
    w, prob, p0 = _dki(4096, seed=18)
    a = fit.fit(prob, iters=100, params0=p0.clone(), wide=True, mfma=True)
    c = fit.make_solver(prob, wide=True, mfma=True)
    c.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    c.run(40)
    c.run(60)
    torch.cuda.synchronize()
    for k in ("pars_t", "f_t", "lam_t", "step_type_t", "n_iter_t"):
        assert torch.equal(getattr(a, k), getattr(c, k)), ("resume", k)
    for m in (1, 65, 300):
        print(f"[dki mfma sizes] N={m}", flush=True)
        sub = fit.make_problem("dki", w["consts"], w["data"][:, :m])
        b = fit.fit(sub, iters=100, params0=p0[:, :m].clone(), wide=True, mfma=True)
        torch.cuda.synchronize()
        for k in ("pars_t", "f_t", "lam_t", "step_type_t", "n_iter_t"):
            assert torch.equal(getattr(b, k), getattr(a, k)[:, :m]), (m, k)


def test_mfma_linear_model_reaches_least_squares(gpu):
    """The 11-parameter linear model through the matrix-core kernel (identity
    link: w = 1) ends at the float64 least-squares solution."""
    M, N = 48, 4096
    t = 2 * np.pi * np.arange(M) / M
    X = np.stack([np.ones(M)] + [f(k * t) for k in range(1, 6) for f in (np.cos, np.sin)], 1)
    rng = np.random.default_rng(5)
    truth = rng.uniform(-1, 1, (11, N))
    y = X @ truth + 0.01 * rng.standard_normal((M, N))
    s = SecondOrderLevenbergMarquardt(TRIG_EXPR, TRIG_PARS, ["t"], M, torch.float32, wide=True,
                                      mfma=True)
    dev = "cuda"
    pars = torch.zeros((11, N), dtype=torch.float32, device=dev)
    lb = torch.full((11, N), -10.0, device=dev)
    ub = torch.full((11, N), 10.0, device=dev)
    s.setup(pars, torch.tensor(t, dtype=torch.float32, device=dev),
            torch.tensor(y, dtype=torch.float32, device=dev), lb, ub)
    s.run(100)
    torch.cuda.synchronize()
    y32 = y.astype(np.float32).astype(np.float64)
    ref = np.linalg.lstsq(X, y32, rcond=None)[0]
    got = pars.double().cpu().numpy()
    assert np.abs(got - ref).max() < 1e-4, np.abs(got - ref).max()


def test_wide_tiling_and_resume_are_bit_identical(gpu):
    w, prob, p0 = _dki(4096, seed=8)
    a = fit.fit(prob, iters=100, params0=p0.clone(), wide=True, tile=136)
    b = fit.fit(prob, iters=100, params0=p0.clone(), wide=True, tile=40)
    c = fit.make_solver(prob, wide=True, tile=136)
    c.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    c.run(40)
    c.run(60)
    torch.cuda.synchronize()
    for k in ("pars_t", "f_t", "lam_t", "step_type_t"):
        assert torch.equal(getattr(a, k), getattr(b, k)), ("tile", k)
        assert torch.equal(getattr(a, k), getattr(c, k)), ("resume", k)


def test_wide_linear_model_reaches_least_squares(gpu):
    """An 11-parameter trigonometric model is linear, so the LM must end at
    the least-squares solution (numpy, float64)."""
    M, N = 48, 4096
    t = 2 * np.pi * np.arange(M) / M
    X = np.stack([np.ones(M)] + [f(k * t) for k in range(1, 6) for f in (np.cos, np.sin)], 1)
    rng = np.random.default_rng(4)
    truth = rng.uniform(-1, 1, (11, N))
    y = X @ truth + 0.01 * rng.standard_normal((M, N))
    s = SecondOrderLevenbergMarquardt(TRIG_EXPR, TRIG_PARS, ["t"], M, torch.float32, wide=True)
    dev = "cuda"
    pars = torch.zeros((11, N), dtype=torch.float32, device=dev)
    lb = torch.full((11, N), -10.0, device=dev)
    ub = torch.full((11, N), 10.0, device=dev)
    s.setup(pars, torch.tensor(t, dtype=torch.float32, device=dev),
            torch.tensor(y, dtype=torch.float32, device=dev), lb, ub)
    s.run(100)
    torch.cuda.synchronize()
    y32 = y.astype(np.float32).astype(np.float64)
    ref = np.linalg.lstsq(X, y32, rcond=None)[0]
    got = pars.double().cpu().numpy()
    # The FP32 stop rule |g|^2 < 1e-8 (1 + f) leaves |g| ~ 1e-4 against
    # Hessian eigenvalues 24..48, so parameters agree to ~1e-5; a wrong entry
    # of the tiled Hessian or of the solve would miss by orders of magnitude.
    assert np.abs(got - ref).max() < 1e-4, np.abs(got - ref).max()
    grad = X.T @ (X @ got - y32)
    assert np.abs(grad).max() < 2e-3, np.abs(grad).max()
    ssr_ref = ((X @ ref - y32) ** 2).sum(0)
    ssr = s.final_f().double().cpu().numpy()[0]
    assert np.allclose(ssr, ssr_ref, rtol=2e-3), np.abs(ssr / ssr_ref - 1).max()
    # the reference's stop also needs the previous step accepted with rho > eta,
    # which FP32 noise in rho denies some voxels near the optimum
    assert float((s.step_type_t == 0).double().mean()) > 0.9


def test_dki_mlp_leg_matches_oracle(gpu):
    """C5 leg: the kurtosis network's guess (61 -> 256 -> 256 -> 22, trained on
    forward-model signals of a separate seed), kept per voxel only where its
    SSR beats the default guess, then the wide LM capped at 10 iterations,
    against the oracle run from the same guesses, to the parity bars of
    tests/parity.py."""
    n = 2048
    w, prob, pd = _dki(n, seed=62)
    net = mlp.dki_mlp(w["consts"])
    T = synth.dki_truth_gpu(1_000_000, seed=20240614)
    data, truth = net.synthesize(1_000_000, seed=777, snr=(30.0, 30.0), truth=T)
    net.fit_standardization(truth, data)
    X, Tt = net.prepare(data, truth)
    net.train(X, Tt, epochs=2, batch=8192)
    p0 = pd.clone()
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p0)
    _init.keep_better(prob.model, prob.consts, prob.data_t, p0, pd)
    s = fit.fit(prob, iters=10, params0=p0.clone(), wide=True)
    torch.cuda.synchronize()
    y = np.ascontiguousarray(w["data"])
    r32, r64 = parity.oracle_pair("dki", w["consts"], y, prob.lb_t.cpu().numpy(),
                                  prob.ub_t.cpu().numpy(), p0.cpu().numpy(), 10)
    parity.check("C5 MLP guess + 10 wide-LM iterations", "dki", w["consts"], y,
                 s.pars_t.cpu().numpy(), r32, r64, gpu_first_f=s.first_f.cpu().numpy()[0])


# non-GLM models with more than 8 parameters: a 4-exponential decay with
# offset (9) and a 4-component Gaussian mixture (12)
MULTIEXP = ("a1*exp(-t/b1) + a2*exp(-t/b2) + a3*exp(-t/b3) + a4*exp(-t/b4) + c",
            ("a1", "b1", "a2", "b2", "a3", "b3", "a4", "b4", "c"))
GAUSS4 = (" + ".join(f"a{k}*exp(-((t - m{k})/w{k})**2)" for k in range(1, 5)),
          tuple(x for k in range(1, 5) for x in (f"a{k}", f"m{k}", f"w{k}")))


def _general_problem(which, N, seed):
    rng = np.random.default_rng(seed)
    if which == "multiexp":
        expr, pars = MULTIEXP
        t = np.linspace(0.05, 5.0, 32)
        b_lo, b_hi = [0.1, 0.5, 1.5, 4.0], [0.3, 1.0, 3.0, 8.0]
        truth = [rng.uniform(0.2, 1.0, N) if j % 2 == 0 and j < 8 else None for j in range(9)]
        for k in range(4):
            truth[2 * k + 1] = rng.uniform(b_lo[k], b_hi[k], N)
        truth[8] = rng.uniform(0.0, 0.1, N)
        truth = np.stack(truth)
        lb = np.array([0, 1e-3] * 4 + [-1.0])
        ub = np.array([10, 100] * 4 + [1.0])
    else:
        expr, pars = GAUSS4
        t = np.linspace(0.0, 10.0, 48)
        mu = [1.5, 3.5, 6.0, 8.5]
        rows = []
        for k in range(4):
            rows += [rng.uniform(0.5, 1.5, N), mu[k] + rng.uniform(-0.3, 0.3, N),
                     rng.uniform(0.5, 1.2, N)]
        truth = np.stack(rows)
        lb = np.array([0, 0, 0.05] * 4)
        ub = np.array([10, 10, 5.0] * 4)
    from pycompute_amd.sym import expr as E, hipgen
    dag = hipgen.res_jac_hes(expr, pars, ("t",))
    env = {k: truth[i][None, :] for i, k in enumerate(pars)}
    env["t"] = t[:, None]
    y = (E.evaluate(dag.model, env) + rng.normal(0, 0.01, (t.size, N))).astype(np.float32)
    p0 = np.clip(truth * rng.uniform(0.85, 1.2, truth.shape), lb[:, None], ub[:, None])
    lbN = np.repeat(lb[:, None], N, 1).astype(np.float32)
    ubN = np.repeat(ub[:, None], N, 1).astype(np.float32)
    return expr, pars, t, y, lbN, ubN, p0.astype(np.float32)


def _general_fit(expr, pars, t, y, lb, ub, p0, iters, runs=None, **kw):
    s = SecondOrderLevenbergMarquardt(expr, list(pars), ["t"], t.size, torch.float32, **kw)
    dev = "cuda"
    s.setup(torch.tensor(p0, device=dev), torch.tensor(t, dtype=torch.float32, device=dev),
            torch.tensor(y, device=dev), torch.tensor(lb, device=dev), torch.tensor(ub, device=dev))
    for k in (runs or [iters]):
        s.run(k)
    torch.cuda.synchronize()
    return s


@pytest.mark.parametrize("which", ["multiexp", "gauss4"])
def test_general_wide_strict_bit_exact(gpu, which):
    """The strict build of the general wide kernel equals the oracle's replay
    of the same generated model bit for bit, also when the fit is resumed."""
    expr, pars, t, y, lb, ub, p0 = _general_problem(which, 1024, 81)
    s = _general_fit(expr, pars, t, y, lb, ub, p0, 40, strict=True)
    assert s._wide_prog is not None and s._glm_tab is None
    ref = gm.lm_run(expr, pars, ("t",), t[None, :], y, lb, ub, p0, 40)
    got = {"params": s.pars_t.cpu().numpy(), "lam": s.lam_t.cpu().numpy()[0],
           "step_type": s.step_type_t.cpu().numpy()[0], "f": s.f_t.cpu().numpy()[0],
           "first_f": s.first_f.cpu().numpy()[0], "n_iter": s.n_iter_t.cpu().numpy()[0]}
    for k, a in got.items():
        same = (a == ref[k]) | (np.isnan(a) & np.isnan(ref[k])) if a.dtype.kind == "f" \
            else (a == ref[k])
        print(f"[{which}] {k}: identical on {float(np.mean(same)):.6f}", flush=True)
        assert same.all(), (which, k, int((~same).sum()))
    r = _general_fit(expr, pars, t, y, lb, ub, p0, 40, runs=[15, 25], strict=True)
    for k in ("pars_t", "lam_t", "step_type_t", "f_t", "n_iter_t"):
        assert torch.equal(getattr(s, k), getattr(r, k)), ("resume", k)
    for n in (1, 65):
        e = _general_fit(expr, pars, t, np.ascontiguousarray(y[:, :n]),
                         np.ascontiguousarray(lb[:, :n]), np.ascontiguousarray(ub[:, :n]),
                         np.ascontiguousarray(p0[:, :n]), 40, strict=True)
        assert torch.equal(e.pars_t, s.pars_t[:, :n]) and torch.equal(e.f_t, s.f_t[:, :n])


@pytest.mark.parametrize("which", ["multiexp", "gauss4"])
@pytest.mark.parametrize("build", ["default", "fast"])
def test_general_wide_parity(gpu, build, which):
    """Default and fast builds of the general wide kernel within the parity
    bars of tests/parity.py against the oracle's FP32 and FP64 replays of the
    generated model (which use the strict exp, so FP32 rounding differs from
    the device's by more than the strict build's). The 4-exponential model
    is ill-conditioned: there the FP32 oracle itself ends within 1e-4 of the
    FP64 one on only about 90% of voxels."""
    kw = {} if build == "default" else dict(fast_exp=True, fast_div=True, reassoc=True)
    expr, pars, t, y, lb, ub, p0 = _general_problem(which, 4096, 82)
    s = _general_fit(expr, pars, t, y, lb, ub, p0, 60, **kw)
    r32, r64, ssr = parity.gen_oracle_pair(expr, pars, ("t",), t[None, :], y, lb, ub, p0, 60)
    parity.check(f"wide-{which}-{build}", None, t[None, :], y, s.pars_t.cpu().numpy(), r32, r64,
                 gpu_first_f=s.first_f.cpu().numpy()[0], ssr_fn=ssr)
