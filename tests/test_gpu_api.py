"""GPU checks of the remaining reference API surface: the F_GradF plugin hook
(user device code with the reference signatures), bpolyfit, the standalone
GMW81 solver, and the committed golden fixtures."""
import json
import os

import numpy as np
import pytest
import torch

import oracle
from pycompute_amd import fit, lsq, synth
from pycompute_amd.lsqnonlin import F_GradF, SecondOrderLevenbergMarquardt
from pycompute_amd.models import IVIM
from pycompute_amd.program import HipFunction
from pycompute_amd.sym import hipgen

pytestmark = pytest.mark.gpu


class DeviceCode(HipFunction):
    def __init__(self, funcid, code, deps=()):
        self.funcid, self.code, self.deps = funcid, code, list(deps)

    def get_device_funcid(self):
        return self.funcid

    def get_device_code(self):
        return self.code

    def get_deps(self):
        return self.deps


# A user's IVIM plugin with the reference's fghhl / f signatures, built on the
# generated model struct so its arithmetic matches the expression path.
GRADF = """
__device__ void user_ivim_fghhl(const float* params, const float* consts, const float* data,
    const float* lam, float* f, float* g, float* h, float* hl, int tid, int Nprobs) {
  float p[4];
  for (int j = 0; j < 4; ++j) p[j] = params[j * Nprobs + tid];
  float ff = 0.f, gg[4] = {0.f, 0.f, 0.f, 0.f}, hh[10], hd[4] = {0.f, 0.f, 0.f, 0.f};
  for (int l = 0; l < 10; ++l) hh[l] = 0.f;
  const float lm = lam[tid];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    float c[1] = {consts[i * Nprobs + tid]};
    UserIvim::accum(p, c, data[i * Nprobs + tid], lm, ff, gg, hh, hd);
  }
  f[tid] += ff;
  int l = 0;
  for (int j = 0; j < 4; ++j) {
    g[j * Nprobs + tid] += gg[j];
    for (int k = 0; k <= j; ++k, ++l) {
      h[l * Nprobs + tid] += hh[l];
      hl[l * Nprobs + tid] += (j == k) ? hd[j] : hh[l];
    }
  }
}
"""
FOBJ = """
__device__ void user_ivim_f(const float* params, const float* consts, const float* data,
    float* f, int tid, int Nprobs) {
  float p[4];
  for (int j = 0; j < 4; ++j) p[j] = params[j * Nprobs + tid];
  float ff = 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    float c[1] = {consts[i * Nprobs + tid]};
    const float r = UserIvim::resid(p, c, data[i * Nprobs + tid]);
    ff += r * r;
  }
  f[tid] += ff;
}
"""


def test_f_gradf_plugin_matches_expression_path(gpu):
    dag = hipgen.res_jac_hes(IVIM.expr, IVIM.pars, IVIM.consts)
    model = DeviceCode("UserIvim", hipgen.emit_model(dag, "UserIvim", "float32"))
    fg = F_GradF(DeviceCode("user_ivim_f", FOBJ, [model]),
                 DeviceCode("user_ivim_fghhl", GRADF, [model]), 4, 1, "user_ivim")
    w = synth.ivim(16384, synth.B_C2, seed=2, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    ref = fit.fit(prob, iters=100, params0=p0.clone(), compaction=False)
    s = SecondOrderLevenbergMarquardt(None, fg, 16, torch.float32)
    assert s.funcid == "second_order_levenberg_marquardt_4_1_16_f_user_ivim"
    cons = prob.consts_t.reshape(-1, 1).expand(-1, prob.N).contiguous()   # (M, N) layout
    pars = p0.clone()
    s.setup(pars, cons, prob.data_t, prob.lb_t, prob.ub_t)
    s.run(40)
    s.run(60)
    torch.cuda.synchronize()
    same = (s.pars_t == ref.pars_t).all(dim=0)
    print("plugin identical on", float(same.float().mean()), flush=True)
    assert torch.equal(s.step_type_t, ref.step_type_t)
    assert torch.equal(s.pars_t, ref.pars_t)
    assert torch.equal(s.lam_t, ref.lam_t)
    assert torch.equal(s.n_iter_t, ref.n_iter_t)
    assert torch.equal(s.f_t, ref.f_t)


def test_bpolyfit_matches_least_squares(gpu):
    rng = np.random.default_rng(0)
    m, N = 12, 5000
    x = rng.uniform(0, 2, (m, N))
    c = rng.standard_normal((3, N))
    y = c[0] + c[1] * x + c[2] * x ** 2 + 1e-3 * rng.standard_normal((m, N))
    # float32 inputs are rounded before the (double precision) solve, so the
    # float32 result is compared with the least-squares fit of the rounded data;
    # what remains is the final rounding of the coefficients to float32.
    for dt, npt, tol in ((torch.float64, np.float64, 1e-9), (torch.float32, np.float32, 1e-5)):
        xr, yr = x.astype(npt).astype(np.float64), y.astype(npt).astype(np.float64)
        ref = oracle.bpolyfit(xr, yr, 2)
        out = lsq.bpolyfit(torch.tensor(x, device="cuda").reshape(1, m, N),
                           torch.tensor(y, device="cuda", dtype=dt), 2, dt)
        torch.cuda.synchronize()
        err = np.abs(out.double().cpu().numpy() - ref) / (np.abs(ref) + 1e-3)
        assert err.max() < tol, (dt, err.max())
    # shared x (the reference's replicated b-values)
    xs = np.linspace(0, 1, m)
    ys = 2.0 - 3.0 * xs[:, None] + 0 * c[0]
    out = lsq.bpolyfit(torch.tensor(xs, device="cuda"), torch.tensor(ys, device="cuda"), 1,
                       torch.float64)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), np.stack([np.full(N, 2.0), np.full(N, -3.0)]),
                               atol=1e-10)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("n", [4, 22])
@pytest.mark.parametrize("dt,tag", [(torch.float32, "f32"), (torch.float64, "f64")])
def test_gmw81_solver_bit_exact_on_stress_fixture(gpu, n, dt, tag):
    """The reference's GMW81 stress recipe (test/linalg/test_gmw81.py:17-37):
    the strict build reproduces the oracle's solutions bit for bit."""
    from pycompute_amd.solver import GMW81Solver
    g = np.load(os.path.join(GOLDEN, "gmw81_stress.npz"))
    mat = torch.tensor(g[f"mat{n}"], dtype=dt, device="cuda")
    rhs = torch.tensor(g[f"rhs{n}"], dtype=dt, device="cuda")
    sol = torch.empty_like(rhs)
    GMW81Solver(n, dt, strict=True).run(mat, rhs, sol)
    torch.cuda.synchronize()
    exp = g[f"sol{n}_{tag}"]
    got = sol.cpu().numpy()
    print(n, tag, "bit-exact systems", float(np.mean(np.all(got == exp, axis=0))), flush=True)
    np.testing.assert_array_equal(got, exp)


def test_gmw81_solver_reference_kat(gpu):
    from pycompute_amd.solver import GMW81Solver
    with open(os.path.join(GOLDEN, "kat_test_mats.json")) as fh:
        d = json.load(fh)
    A = np.array(d["A"])
    packed = np.array([A[i, j] for i in range(4) for j in range(i + 1)])
    for dt, rtol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        mat = torch.tensor(packed[:, None].repeat(3, 1), dtype=dt, device="cuda")
        rhs = torch.tensor(np.array(d["b"])[:, None].repeat(3, 1), dtype=dt, device="cuda")
        sol = torch.empty_like(rhs)
        GMW81Solver(4, dt).run(mat, rhs, sol)
        torch.cuda.synchronize()
        np.testing.assert_allclose(sol.cpu().numpy(), np.array(d["x"])[:, None].repeat(3, 1),
                                   rtol=rtol)


def test_ivim_c1_golden_fixture(gpu):
    """The default GPU build on the committed C1 fixture, started from the
    fixture's initial guess: final SSR within 1e-4 of the oracle's FP32 result
    (as close as the FP32 oracle is to its FP64 result)."""
    d = np.load(os.path.join(GOLDEN, "ivim_c1_256.npz"))
    prob = fit.make_problem("ivim", d["consts"], d["data"])
    p0 = torch.tensor(d["p0"], dtype=torch.float32, device="cuda")
    s = fit.fit(prob, iters=100, params0=p0)
    f = s.final_f().double().cpu().numpy()[0]
    rel32 = np.abs(f - d["ssr_f32"]) / d["ssr_f32"]
    rel64 = np.abs(f - d["ssr_f64"]) / d["ssr_f64"]
    o = np.abs(d["ssr_f32"] - d["ssr_f64"]) / d["ssr_f64"]
    print("C1 fixture: vs f32", np.mean(rel32 <= 1e-4), "vs f64", np.mean(rel64 <= 1e-4),
          "oracle f32 vs f64", np.mean(o <= 1e-4), flush=True)
    assert np.mean(rel32 <= 1e-4) >= 0.97
    assert np.mean(rel64 <= 1e-4) >= np.mean(o <= 1e-4) - 0.02
