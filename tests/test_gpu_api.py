"""GPU checks of the remaining reference API surface: the F_GradF plugin hook
(user device code with the reference signatures) and bpolyfit."""
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
