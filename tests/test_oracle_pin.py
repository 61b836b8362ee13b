"""Independent pin of the oracle's LM iteration and default guess (CPU only).

The oracle (oracle/lm_oracle.cpp) restates the reference kernel operation by
operation. Here one iteration is computed a second, independent way, in
float64 numpy, from the mathematics of the reference's second-order LM
(test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:2-73,
307-365, 367-414, 416-517; pycompute/cuda/lsqnonlin.py:206-346, 432-498,
584-607, 691-729, 820-947):

  r_i  = m(p, c_i) - y_i,  f = sum r_i^2,  g = J^T r
  H    = sum_i (J_i J_i^T + r_i d2m_i)                  full second order
  HL   = H + diag(sum_i max(lam J_ij^2, 1e-4))          per-measurement floor
  s    = HL^{-1} g   (numpy.linalg.solve; GMW81 adds nothing when the scaled
                      HL is comfortably positive definite, checked per voxel)
  stop if |p - clamp(p - g)|^2 < tol (1 + f) and (st & 1 or clamped)
  p_tp = p - s (unclamped),  rho = 0.5 (f - f(p_tp)) / (s.g - 0.5 s^T H s)
  accept / reject / lambda / step-type rules, then clamp.

m, J and d2m come from sympy (lambdified), not from the oracle or the code
generator. The comparison is voxel by voxel over consecutive iterations,
each started from the oracle's own state after the previous one.

Mutation checks: the same comparison is run against restatements with one
plausible mistake each (Gauss-Newton Hessian, no damping floor, damped H in
the predicted reduction, clamped trial point, the two ratio thresholds or
the two damping factors swapped, gradient sign) and must reject every one of
them, so a mistake of that kind in the oracle would fail this test.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pycompute_amd import models, synth
from pycompute_amd.init import ivim_split

TOL, MU, ETA, ACC, DEC = 1e-8, 0.25, 0.75, 0.2, 2.0

# the mutated restatements and far-off trial points overflow exp on purpose
pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def _sympy_funcs(name):
    import sympy as sp
    m = models.MODELS[name]
    syms = {s: sp.Symbol(s) for s in m.pars + m.consts}
    expr = sp.sympify(m.expr, locals=syms)
    P = [syms[p] for p in m.pars]
    C = [syms[c] for c in m.consts]
    f = sp.lambdify(P + C, expr, "numpy")
    J = [sp.lambdify(P + C, sp.diff(expr, p), "numpy") for p in P]
    H = [[sp.lambdify(P + C, sp.diff(expr, P[j], P[k]), "numpy") for k in range(len(P))]
         for j in range(len(P))]
    return m, f, J, H


def _eval(funcs, p, consts):
    """m (M, N), J (M, N, n), d2m (M, N, n, n) in float64."""
    m, f, Jf, Hf = funcs
    n, N = p.shape
    M = consts.shape[1]
    mv = np.zeros((M, N))
    J = np.zeros((M, N, n))
    H = np.zeros((M, N, n, n))
    for i in range(M):
        args = [p[j] for j in range(n)] + [np.full(N, consts[k, i]) for k in range(consts.shape[0])]
        mv[i] = np.broadcast_to(f(*args), (N,))
        for j in range(n):
            J[i, :, j] = np.broadcast_to(Jf[j](*args), (N,))
            for k in range(n):
                H[i, :, j, k] = np.broadcast_to(Hf[j][k](*args), (N,))
    return mv, J, H


def _gmw81_leaves_unmodified(HL):
    """True where the scaled HL is positive definite with a margin large
    enough that GMW81 (solver.py:20-115) takes every pivot as it stands:
    d_j >= max(delta, theta_j^2 / beta^2) with room to spare, so its LDL^T
    equals the plain factorisation and the solve equals numpy's."""
    N, n, _ = HL.shape
    ok = np.zeros(N, bool)
    perm = [n - 1] + list(range(n - 1))          # the reference's fixed rotation
    for v in range(N):
        A = HL[v][np.ix_(perm, perm)]
        d = np.sqrt(np.abs(np.diag(A)))
        if not np.all(np.isfinite(A)) or np.any(d == 0):
            continue
        B = A / d[:, None] / d[None, :]
        gamma = np.max(np.abs(np.diag(B)))
        xi = np.max(np.abs(B[np.tril_indices(n, -1)])) if n > 1 else 0.0
        if n > 1:
            xi /= np.sqrt(n * n - 1)
        beta2 = max(4e-16, gamma, xi)
        delta = 4e-16 * max(1.0, gamma + xi)
        # plain LDL^T of B
        L = np.eye(n)
        D = np.zeros(n)
        good = True
        for j in range(n):
            D[j] = B[j, j] - np.sum(L[j, :j] ** 2 * D[:j])
            c = B[j + 1:, j] - L[j + 1:, :j] @ (L[j, :j] * D[:j])
            theta = np.max(np.abs(c)) if j < n - 1 else 0.0
            if not (D[j] > 2 * delta and D[j] > 2 * theta * theta / beta2):
                good = False
                break
            L[j + 1:, j] = c / D[j]
        ok[v] = good
    return ok


def restated_iteration(funcs, consts, y, lb, ub, p, lam, st, *, mutation=None):
    """One reference launch in float64 for every voxel. Returns (p, lam, st,
    f, pd_ok) with pd_ok marking the voxels where the plain solve is exact."""
    p = p.astype(np.float64).copy()
    lam = lam.astype(np.float64).copy()
    st = st.astype(np.int64).copy()
    n, N = p.shape
    mv, J, d2 = _eval(funcs, p, consts)
    r = mv - y.astype(np.float64)                                   # (M, N)
    f = np.sum(r * r, axis=0)
    g = np.einsum("mvj,mv->vj", J, r)
    if mutation == "grad_sign":
        g = -g
    JJ = np.einsum("mvj,mvk->vjk", J, J)
    H = JJ if mutation == "gauss_newton" else JJ + np.einsum("mvjk,mv->vjk", d2, r)
    if mutation == "no_floor":
        damp = lam[:, None] * np.sum(J * J, axis=0)
    else:
        damp = np.sum(np.maximum(lam[None, :, None] * J * J, 1e-4), axis=0)
    HL = H.copy()
    HL[:, np.arange(n), np.arange(n)] += damp
    pd_ok = _gmw81_leaves_unmodified(HL)
    s = np.zeros((N, n))
    for v in range(N):
        if pd_ok[v]:
            s[v] = np.linalg.solve(HL[v], g[v])
    s[~np.isfinite(s)] = 0.0
    s = s.T                                                          # (n, N)
    active = st != 0
    # convergence
    t = np.clip(p - g.T, lb, ub)
    clamped = np.any((p - g.T > ub) | (p - g.T < lb), axis=0)
    q = np.sum((p - t) ** 2, axis=0)
    conv = active & (q < TOL * (1 + f)) & (((st & 1) != 0) | clamped)
    # trial point
    ptp = p - s
    if mutation == "clamped_trial":
        ptp = np.clip(ptp, lb, ub)
    mt, _, _ = _eval(funcs, ptp, consts)
    ftp = np.sum((mt - y.astype(np.float64)) ** 2, axis=0)
    actual = 0.5 * (f - ftp)
    Hp = HL if mutation == "damped_pred" else H
    pred = np.einsum("jv,jv->v", s, g.T) - 0.5 * np.einsum("jv,vjk,kv->v", s, Hp, s)
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = actual / pred
    mu, eta = (ETA, MU) if mutation == "swap_mu_eta" else (MU, ETA)
    acc, dec = (DEC, ACC) if mutation == "swap_acc_dec" else (ACC, DEC)
    accept = (rho > mu) & (actual > 0)
    newp = np.where(accept[None, :], ptp, p)
    newlam = np.where(accept, np.where(rho > eta, lam * acc, lam), lam * dec)
    newst = np.where(accept, np.where(rho > eta, 1, 2), 4)
    neg = pred < 0
    newlam = np.where(neg, newlam * dec, newlam)
    newst = np.where(neg, newst | 8, newst)
    newp = np.clip(newp, lb, ub)
    step = active & ~conv
    out_p = np.where(step[None, :], newp, p)
    out_lam = np.where(step, newlam, lam)
    out_st = np.where(step, newst, np.where(conv, 0, st))
    # f is the SSR at the start of the iteration for voxels that ran
    return out_p, out_lam, out_st, np.where(active, f, np.nan), pd_ok


def _oracle_states(name, consts, y, lb, ub, p0, iters, lam0=None):
    """The oracle's state after each of `iters` single launches (resumed
    launch by launch; resume equals one longer run, tested elsewhere).
    lam0: starting damping per voxel (default: the reference's 5)."""
    N = p0.shape[1]
    states = [dict(params=p0.astype(np.float64),
                   lam=np.full(N, 5.0) if lam0 is None else np.asarray(lam0, np.float64),
                   step_type=np.ones(N, np.int8), f=np.zeros(N))]
    for _ in range(iters):
        s = states[-1]
        r = oracle.lm_run(name, consts, y, lb, ub, s["params"], 1, dtype=np.float64,
                          lam=s["lam"], step_type=s["step_type"], f=s["f"])
        states.append(r)
    return states


def _compare(funcs, name, consts, y, lb, ub, states, mutation=None):
    """Per-voxel agreement of the restatement with the oracle over the
    recorded iterations; returns (fraction of compared voxel-iterations that
    agree, number compared)."""
    agree = total = 0
    for k in range(len(states) - 1):
        s0, s1 = states[k], states[k + 1]
        p, lam, st, f, ok = restated_iteration(funcs, consts, y, lb, ub, s0["params"], s0["lam"],
                                               s0["step_type"], mutation=mutation)
        ran = s0["step_type"] != 0
        sel = ok & ran
        scale = np.maximum(np.abs(s1["params"]), np.abs(ub - lb) * 1e-6)
        same = (np.all(np.abs(p - s1["params"]) <= 1e-6 * scale + 1e-300, axis=0)
                & np.isclose(lam, s1["lam"], rtol=1e-12, atol=0)
                & (st == s1["step_type"])
                & np.isclose(f, s1["f"], rtol=1e-10, atol=1e-300))
        agree += int(np.sum(same & sel))
        total += int(np.sum(sel))
    return agree / max(total, 1), total


def _ivim_case(N=512, seed=21):
    """IVIM voxels from perturbed starts: accepted and rejected steps, active
    bounds, the damping floor active at b = 0 (J_D1 = J_D2 = 0 there) and
    inactive elsewhere at large lambda."""
    w = synth.ivim_c1(N, seed=seed)
    m = models.IVIM
    lb = np.repeat(np.array(m.lb)[:, None], N, 1)
    ub = np.repeat(np.array(m.ub)[:, None], N, 1)
    rng = np.random.default_rng(seed)
    p0 = w["truth"] * rng.uniform(0.5, 1.8, w["truth"].shape)
    p0[1] = np.clip(p0[1], 0.02, 0.9)
    p0[:, : N // 8] = np.clip(p0[:, : N // 8] * 30, lb[:, : N // 8], ub[:, : N // 8])  # at bounds
    p0 = np.clip(p0, lb, ub)
    # every fourth voxel gets tight per-voxel bounds around its start (the
    # notebook's cell-10 style, test_ivim.ipynb:207-232), so trial points
    # often leave the box
    q = np.arange(N) % 4 == 3
    lb[:, q] = np.maximum(lb[:, q], p0[:, q] / 1.3)
    ub[:, q] = np.minimum(ub[:, q], p0[:, q] * 1.3)
    return w["consts"], w["data"].astype(np.float64), lb, ub, p0


@pytest.fixture(scope="module")
def ivim_setup():
    funcs = _sympy_funcs("ivim")
    consts, y, lb, ub, p0 = _ivim_case()
    # half the voxels start at the reference's lambda 5, half log-uniform in
    # [1e-2, 1e2], so steps run from damped to nearly Gauss-Newton
    rng = np.random.default_rng(5)
    N = p0.shape[1]
    lam0 = np.where(np.arange(N) % 2 == 0, 5.0, np.exp(rng.uniform(np.log(1e-2), np.log(1e2), N)))
    states = _oracle_states("ivim", consts, y, lb, ub, p0, 4, lam0)
    return funcs, consts, y, lb, ub, states


def test_one_iteration_matches_independent_restatement(ivim_setup):
    funcs, consts, y, lb, ub, states = ivim_setup
    frac, total = _compare(funcs, "ivim", consts, y, lb, ub, states)
    assert total >= 800, total
    assert frac == 1.0, frac
    # the case exercises what the mutations below change
    st = np.concatenate([s["step_type"] for s in states[1:]])
    assert np.any((st & 4) != 0) and np.any(st == 1) and np.any(st == 2)


def test_case_exercises_floor_and_bounds(ivim_setup):
    funcs, consts, y, lb, ub, states = ivim_setup
    p = states[0]["params"]
    _, J, _ = _eval(funcs, p, consts)
    lam = states[0]["lam"][None, :, None]
    below = lam * J * J < 1e-4
    above = ~below
    assert below.any() and above.any()
    assert np.any(np.isclose(p, lb) | np.isclose(p, ub))


@pytest.mark.parametrize("mutation", ["gauss_newton", "no_floor", "damped_pred",
                                      "clamped_trial", "swap_mu_eta", "swap_acc_dec",
                                      "grad_sign"])
def test_mutations_are_rejected(ivim_setup, mutation):
    funcs, consts, y, lb, ub, states = ivim_setup
    frac, total = _compare(funcs, "ivim", consts, y, lb, ub, states, mutation=mutation)
    assert total > 0
    assert frac < 0.98, (mutation, frac)


@pytest.mark.parametrize("name", ["t2", "se"])
def test_other_models_match_independent_restatement(name):
    N = 192
    w = synth.t2(N, seed=31) if name == "t2" else synth.se(N, seed=32)
    m = models.MODELS[name]
    lb = np.repeat(np.array(m.lb)[:, None], N, 1)
    ub = np.repeat(np.array(m.ub)[:, None], N, 1)
    rng = np.random.default_rng(7)
    p0 = np.clip(w["truth"] * rng.uniform(0.6, 1.5, w["truth"].shape), lb, ub)
    y = w["data"].astype(np.float64)
    states = _oracle_states(name, w["consts"], y, lb, ub, p0, 3)
    frac, total = _compare(_sympy_funcs(name), name, w["consts"], y, lb, ub, states)
    assert total >= 300 and frac == 1.0, (total, frac)


# ---------------------------------------------------------------------------
# default initial guess (test/ivim/test_ivim.ipynb cell 6, :146-165)
# ---------------------------------------------------------------------------

def test_default_init_closed_form():
    """Data that is exactly A exp(-b a) on the low-b measurements and
    B exp(-b c) on the high-b ones (the segments the guess fits; the middle
    measurements are arbitrary) gives S0 = A, f = 1 - B/A, D_1 = a, D_2 = c."""
    b = synth.B_C2
    lo, hi = ivim_split(b)
    N = 64
    rng = np.random.default_rng(3)
    A = rng.uniform(0.8, 1.2, N)
    Bv = A * rng.uniform(0.6, 0.95, N)
    a = rng.uniform(5e-3, 0.05, N)
    c = rng.uniform(5e-4, 3e-3, N)
    y = rng.uniform(0.1, 1.0, (b.size, N))            # arbitrary in the middle
    y[lo] = A * np.exp(-b[lo, None] * a)
    y[hi] = Bv * np.exp(-b[hi, None] * c)
    m = models.IVIM
    lb = np.repeat(np.array(m.lb)[:, None], N, 1)
    ub = np.repeat(np.array(m.ub)[:, None], N, 1)
    p = oracle.ivim_default_init(b, y, lb, ub, lo, hi)
    np.testing.assert_allclose(p[0], A, rtol=1e-10)
    np.testing.assert_allclose(p[1], 1 - Bv / A, rtol=1e-9)
    np.testing.assert_allclose(p[2], a, rtol=1e-9)
    np.testing.assert_allclose(p[3], c, rtol=1e-9)


def test_default_init_nan_and_clip():
    """A zero sample makes log(y) = -inf: the fitted values become NaN or
    inf, nan_to_num maps them to 0, and the clip lifts them to the lower
    bound; values past the upper bound are clipped down."""
    b = synth.B_C2
    lo, hi = ivim_split(b)
    N = 3
    y = np.ones((b.size, N)) * np.exp(-b[:, None] * 1e-3)
    y[lo[0], 0] = 0.0                          # voxel 0: low segment broken
    y[:, 2] = 3e4 * np.exp(-b * 1e-3)          # voxel 2: S0 above the bound
    m = models.IVIM
    lb = np.repeat(np.array(m.lb)[:, None], N, 1)
    ub = np.repeat(np.array(m.ub)[:, None], N, 1)
    with np.errstate(all="ignore"):
        p = oracle.ivim_default_init(b, y, lb, ub, lo, hi)
    # voxel 0: intercept -inf -> S0 = exp(-inf) = 0, slope nan -> 0 -> lower bound
    assert p[0, 0] == lb[0, 0] and p[2, 0] == lb[2, 0]
    # voxel 1: mono-exponential: S0 = 1, f = 0 -> lower bound, D_1 = D_2 = 1e-3
    np.testing.assert_allclose(p[:, 1], [1.0, lb[1, 1], 1e-3, 1e-3], rtol=1e-9)
    # voxel 2: S0 clipped to its upper bound
    assert p[0, 2] == ub[0, 2]
