"""Edge semantics of the LM iteration on the GPU against the oracle.

Strict build, bit for bit against the oracle's replay of the same generated
model (oracle/genmodel.py), on voxels built to reach the reference's corner
cases (test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:
95-150 pivot, 474-492 NaN/Inf scrub and convergence, 494-514 trial point and
clamp):
  * a NaN or an infinite sample: NaN residuals, a NaN Hessian diagonal (the
    reference's pivot then depends on the NaN; the kernel's compile-time
    rotation must give the same all-NaN step, scrubbed to zero);
  * an all-zero signal (zero-valued residual derivatives at the start);
  * tight per-voxel bounds, so trial points leave the box and are clamped;
  * a parameter with lb == ub (pinned), and starts on a bound;
  * a signal of 1e30 (the SSR overflows to inf).
Every output must be identical, NaN matching NaN.

A model whose device code is written by hand through the F_GradF hook
(pycompute/cuda/lsqnonlin.py:32-38) rather than taken from the code
generator: mono-exponential T2 in HIP, fitted against the oracle's own
closed-form T2 model (oracle/lm_oracle.cpp) on the same inputs.
"""
import numpy as np
import pytest
import torch

import oracle
import oracle.genmodel as gm
import parity
from pycompute_amd import fit, synth
from pycompute_amd.lsqnonlin import F_GradF, SecondOrderLevenbergMarquardt
from pycompute_amd.models import IVIM, T2
from pycompute_amd.program import HipFunction

pytestmark = pytest.mark.gpu


def _edge_case(N_per=64, seed=51):
    """IVIM C1 protocol; 8 groups of N_per voxels, one corner case each."""
    w = synth.ivim_c1(8 * N_per, seed=seed)
    y = w["data"].astype(np.float32).copy()
    truth = w["truth"].copy()
    lb = np.repeat(np.array(IVIM.lb)[:, None], y.shape[1], 1)
    ub = np.repeat(np.array(IVIM.ub)[:, None], y.shape[1], 1)
    rng = np.random.default_rng(seed)
    p0 = np.clip(truth * rng.uniform(0.7, 1.4, truth.shape), lb, ub)
    g = [slice(k * N_per, (k + 1) * N_per) for k in range(8)]
    y[3, g[1]] = np.nan                                 # a NaN sample
    y[5, g[2]] = np.inf                                 # an infinite sample
    y[:, g[3]] = 0.0                                    # all-zero signal
    lb[:, g[4]] = np.maximum(lb[:, g[4]], p0[:, g[4]] / 1.15)   # tight bounds
    ub[:, g[4]] = np.minimum(ub[:, g[4]], p0[:, g[4]] * 1.15)
    lb[1, g[5]] = ub[1, g[5]] = p0[1, g[5]]             # f pinned: lb == ub
    p0[2, g[6]] = ub[2, g[6]]                           # starts on bounds
    p0[3, g[6]] = lb[3, g[6]]
    y[:, g[7]] = 1e30                                   # SSR overflows
    return w["consts"], y, lb.astype(np.float32), ub.astype(np.float32), p0.astype(np.float32)


def _same(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b))


def test_edge_semantics_strict_bit_exact(gpu):
    consts, y, lb, ub, p0 = _edge_case()
    prob = fit.make_problem("ivim", consts, y, lb=lb, ub=ub)
    p0_t = torch.as_tensor(p0, device="cuda").contiguous()
    for compaction in (False, True):
        s = fit.fit(prob, iters=50, params0=p0_t.clone(), strict=True, compaction=compaction,
                    chunk=7)
        torch.cuda.synchronize()
        ref = gm.lm_run(IVIM.expr, IVIM.pars, IVIM.consts, consts, y, lb, ub, p0, 50)
        got = {"params": s.pars_t.cpu().numpy(), "lam": s.lam_t.cpu().numpy()[0],
               "step_type": s.step_type_t.cpu().numpy()[0], "f": s.f_t.cpu().numpy()[0],
               "first_f": s.first_f.cpu().numpy()[0], "n_iter": s.n_iter_t.cpu().numpy()[0]}
        for k, a in got.items():
            same = _same(a, ref[k])
            assert same.all(), (compaction, k, np.nonzero(~same)[-1][:8])
    # the corner cases were reached
    st = got["step_type"]
    f = got["f"]
    assert np.all(np.isnan(f[64:128])) and np.all(np.isinf(f[128:192]) | np.isnan(f[128:192]))
    assert np.all(got["params"][1, 320:384] == lb[1, 320:384])
    assert np.all((got["params"][:, 256:320] >= lb[:, 256:320])
                  & (got["params"][:, 256:320] <= ub[:, 256:320]))
    assert np.any(got["params"][:, 256:320] == lb[:, 256:320]) or \
        np.any(got["params"][:, 256:320] == ub[:, 256:320])
    assert np.all(np.isinf(f[448:512]))
    assert set(np.unique(st).tolist()) <= {0, 1, 2, 4, 9, 10, 12}


def test_edge_semantics_default_build_states(gpu):
    """The default and fastest builds reach the same per-voxel states as the
    oracle on the corner-case groups whose outcome does not depend on
    rounding: the NaN and infinite samples (NaN or inf f, zero steps, the
    start kept), the pinned parameter, bounds respected everywhere."""
    consts, y, lb, ub, p0 = _edge_case(seed=52)
    prob = fit.make_problem("ivim", consts, y, lb=lb, ub=ub)
    p0_t = torch.as_tensor(p0, device="cuda").contiguous()
    ref = oracle.lm_run("ivim", consts, y, lb, ub, p0, 50)
    for kw in ({}, dict(fast_exp=True, fast_div=True, reassoc=True, factored=True)):
        s = fit.fit(prob, iters=50, params0=p0_t.clone(), **kw)
        torch.cuda.synchronize()
        p = s.pars_t.cpu().numpy()
        f = s.f_t.cpu().numpy()[0]
        for g in (slice(64, 128), slice(128, 192)):
            assert np.array_equal(p[:, g], ref["params"][:, g])
            assert np.array_equal(np.isnan(f[g]), np.isnan(ref["f"][g]))
            assert np.array_equal(s.step_type_t.cpu().numpy()[0][g], ref["step_type"][g])
        assert np.all(p[1, 320:384] == lb[1, 320:384])
        assert np.all((p >= lb) & (p <= ub))


# ---------------------------------------------------------------------------
# F_GradF plugin with hand-written device code
# ---------------------------------------------------------------------------

class _Code(HipFunction):
    def __init__(self, funcid, code):
        self._id, self._code = funcid, code

    def get_device_funcid(self):
        return self._id

    def get_device_code(self):
        return self._code


# S0 exp(-TE/T2), written directly: J = (e, S0 TE e / T2^2); second
# derivatives d2/dS0dT2 = TE e / T2^2, d2/dT2^2 = S0 TE e (TE - 2 T2) / T2^4.
T2_GRADF = r"""
__device__ void t2_fghhl(const float* params, const float* consts, const float* data,
                         const float* lam, float* f, float* g, float* h, float* hl,
                         int tid, int N) {
  const float S0 = params[tid], T2 = params[N + tid], l = lam[tid];
  const float iT = 1.0f / T2;
  float ff = 0.f, g0 = 0.f, g1 = 0.f, h0 = 0.f, h1 = 0.f, h2 = 0.f, d0 = 0.f, d1 = 0.f;
  for (int i = 0; i < 32; ++i) {
    const float te = consts[i * N + tid];
    const float e = expf(-te * iT);
    const float r = S0 * e - data[i * N + tid];
    const float j0 = e, j1 = S0 * te * e * iT * iT;
    const float q01 = te * e * iT * iT;
    const float q11 = S0 * te * e * (te - 2.0f * T2) * iT * iT * iT * iT;
    ff += r * r;
    g0 += j0 * r;
    g1 += j1 * r;
    h0 += j0 * j0;
    h1 += j1 * j0 + q01 * r;
    h2 += j1 * j1 + q11 * r;
    d0 += fmaxf(l * j0 * j0, 1e-4f);
    d1 += fmaxf(l * j1 * j1, 1e-4f);
  }
  f[tid] += ff;
  g[tid] += g0;
  g[N + tid] += g1;
  h[tid] += h0;
  h[N + tid] += h1;
  h[2 * N + tid] += h2;
  hl[tid] += h0 + d0;
  hl[N + tid] += h1;
  hl[2 * N + tid] += h2 + d1;
}
"""
T2_F = r"""
__device__ void t2_f(const float* params, const float* consts, const float* data, float* f,
                     int tid, int N) {
  const float S0 = params[tid], T2 = params[N + tid];
  float ff = 0.f;
  for (int i = 0; i < 32; ++i) {
    const float r = S0 * expf(-consts[i * N + tid] / T2) - data[i * N + tid];
    ff += r * r;
  }
  f[tid] += ff;
}
"""


def test_hand_written_plugin_matches_oracle(gpu):
    w = synth.t2(8192, seed=53)
    prob = fit.make_problem("t2", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    fg = F_GradF(_Code("t2_f", T2_F), _Code("t2_fghhl", T2_GRADF), 2, 1, "t2_hand")
    s = SecondOrderLevenbergMarquardt(None, fg, 32, torch.float32)
    cons = prob.consts_t.reshape(-1, 1).expand(-1, prob.N).contiguous()
    s.setup(p0.clone(), cons, prob.data_t, prob.lb_t, prob.ub_t)
    s.run(100)
    torch.cuda.synchronize()
    lb, ub = prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy()
    y = prob.data_t.cpu().numpy()
    r32, r64 = parity.oracle_pair("t2", w["consts"], y, lb, ub, p0.cpu().numpy(), 100)
    parity.check("hand-written T2 plugin", "t2", w["consts"], y, s.pars_t.cpu().numpy(), r32,
                 r64, gpu_first_f=s.first_f.cpu().numpy()[0])
