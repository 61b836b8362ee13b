"""Code generator checks (no GPU): derivatives against sympy and finite
differences, naming against the reference's emitted kernel, op counts, and
hiprtc compile-only builds for gfx950."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from pycompute_amd.models import IVIM, MODELS
from pycompute_amd.sym import expr as E
from pycompute_amd.sym import hipgen

USER_EXPRS = [
    ("a*exp(-x/b) + c", ["a", "b", "c"], ["x"]),
    ("a*sin(w*t + ph)**2 + log(1 + b*t)", ["a", "w", "ph", "b"], ["t"]),
    ("sqrt(a*x + b)/(1 + c*x**3) - a**2", ["a", "b", "c"], ["x"]),
    ("S0*exp(-(b*D)**alpha)", ["S0", "D", "alpha"], ["b"]),
    ("p0*tanh(p1*u)^2 - p2/(u + p1)", ["p0", "p1", "p2"], ["u"]),
]


def _eval_dag(dag, P, C):
    env = dict(zip(dag.pars, P))
    env.update(zip(dag.consts, C))
    memo = {}
    m = E.evaluate(dag.model, env, memo)
    J = [E.evaluate(j, env, memo) for j in dag.jac]
    H = [E.evaluate(h, env, memo) for h in dag.hes]
    return m, J, H


def _points(dag, rng, n=64):
    P = [rng.uniform(0.3, 1.5, n) for _ in dag.pars]
    C = [rng.uniform(0.2, 2.0, n) for _ in dag.consts]
    return P, C


@pytest.mark.parametrize("expr,pars,consts", USER_EXPRS + [(IVIM.expr, IVIM.pars, IVIM.consts)])
def test_dag_matches_sympy(expr, pars, consts):
    import sympy as sp
    dag = hipgen.res_jac_hes(expr, pars, consts)
    syms = {s: sp.Symbol(s) for s in list(pars) + list(consts)}
    se = sp.sympify(expr.replace("^", "**"), locals=syms)
    P_, C_ = [syms[p] for p in pars], [syms[c] for c in consts]
    rng = np.random.default_rng(1)
    P, C = _points(dag, rng)
    m, J, H = _eval_dag(dag, P, C)
    args = P + C
    np.testing.assert_allclose(m, sp.lambdify(P_ + C_, se)(*args), rtol=1e-12)
    l = 0
    for j, pj in enumerate(P_):
        dj = sp.diff(se, pj)
        np.testing.assert_allclose(np.broadcast_to(J[j], (64,)),
                                   np.broadcast_to(sp.lambdify(P_ + C_, dj)(*args), (64,)),
                                   rtol=1e-10, atol=1e-12)
        for k in range(j + 1):
            dk = sp.diff(dj, P_[k])
            np.testing.assert_allclose(np.broadcast_to(H[l], (64,)),
                                       np.broadcast_to(sp.lambdify(P_ + C_, dk)(*args), (64,)),
                                       rtol=1e-9, atol=1e-11)
            l += 1


@pytest.mark.parametrize("name", ["ivim", "t2", "se", "dki"])
def test_builtin_models_finite_differences(name):
    m = MODELS[name]
    dag = hipgen.res_jac_hes(m.expr, m.pars, m.consts)
    rng = np.random.default_rng(2)
    from pycompute_amd import synth
    w = {"ivim": lambda: synth.ivim_c1(8, 1), "t2": lambda: synth.t2(8, 3),
         "se": lambda: synth.se(8, 4), "dki": lambda: synth.dki(8, 6)}[name]()
    P = [w["truth"][j] for j in range(len(m.pars))]
    for i in rng.choice(w["consts"].shape[1], 4, replace=False):
        C = [np.full(8, w["consts"][k][i]) for k in range(len(m.consts))]
        _, J, H = _eval_dag(dag, P, C)
        # complex-step differentiation: d f/dx = Im f(x + i h) / h, free of
        # cancellation, so the comparison can be tight
        h = 1e-30
        l = 0
        for j in range(len(P)):
            Pc = [p.astype(complex) for p in P]
            Pc[j] = Pc[j] + 1j * h
            mc, Jc, _ = _eval_dag(dag, Pc, C)
            np.testing.assert_allclose(np.broadcast_to(J[j], (8,)),
                                       np.broadcast_to(np.imag(mc) / h, (8,)), rtol=1e-11,
                                       atol=1e-300)
            # H is packed lower-triangular: entry (j, k), k <= j, is d J_j / d p_k
            for k in range(j + 1):
                Pk = [p.astype(complex) for p in P]
                Pk[k] = Pk[k] + 1j * h
                _, Jk, _ = _eval_dag(dag, Pk, C)
                np.testing.assert_allclose(np.broadcast_to(H[l], (8,)),
                                           np.broadcast_to(np.imag(Jk[j]) / h, (8,)),
                                           rtol=1e-10, atol=1e-300)
                l += 1


def test_reference_kernel_name():
    """test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu is the
    reference's emitted name for the IVIM model with 21 measurements."""
    from pycompute_amd.lm import LMSpec
    assert hipgen.expr_hash(IVIM.expr, 12) == "f7def86f0a03"
    spec = LMSpec(IVIM.expr, IVIM.pars, IVIM.consts, 21)
    assert spec.funcid() == "second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03"


def test_ivim_structure_and_counts():
    dag = hipgen.res_jac_hes(IVIM.expr, IVIM.pars, IVIM.consts)
    # packed Hessian zeros of IVIM: (S0,S0), (f,f), (D_1,D_2) -- as in the
    # reference's emitted hes[0], hes[2], hes[8] = 0.0f*res
    assert dag.hes_nonzero() == [False, True, False, True, True, True, True, True, False, True]
    assert dag.counts_fjh["exp"] == 2 and dag.counts_f["exp"] == 2
    fl = hipgen.flops_per_iteration(dag, 16)
    assert fl["F_solve"] == 110 and fl["F_step"] == 64 and fl["nnz"] == 7
    assert fl["exp_it"] == 64
    # at most a few ops more than the reference's CSE (20 + 2 exp, 7 + 2 exp)
    assert 22 <= fl["E_fjh"] <= 26 and fl["E_f"] == 9


def test_parse_errors():
    with pytest.raises(ValueError):
        hipgen.res_jac_hes("a*x + q", ["a"], ["x"])
    with pytest.raises(ValueError):
        hipgen.res_jac_hes("a if x else b", ["a", "b"], ["x"])
    with pytest.raises(ValueError):
        hipgen.res_jac_hes("a*x", ["a", "x"], ["x"])


def test_float_literals_are_single_precision():
    dag = hipgen.res_jac_hes("a*exp(-x/3.5) + 0.1*a**2", ["a"], ["x"])
    src = hipgen.emit_model(dag, "M", "float32")
    import re
    for lit in re.findall(r"(?<![\w.])(\d+\.\d*(?:e[-+]?\d+)?f?)", src):
        assert lit.endswith("f"), lit


def test_strict_math_accuracy():
    """The reproducible transcendentals are within a few ulp of numpy's."""
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    hdr = os.path.join(root, "pycompute_amd", "csrc", "strict_math.h")
    d = os.path.join(root, "oracle", "_gen")
    os.makedirs(d, exist_ok=True)
    cpp = os.path.join(d, "strict_probe.cpp")
    so = os.path.join(d, "strict_probe.so")
    with open(cpp, "w") as fh:
        fh.write(f'#include "{hdr}"\n'
                 'extern "C" void e32(const float* x, float* y, int n){for(int i=0;i<n;++i)y[i]=pca_exp_strict(x[i]);}\n'
                 'extern "C" void l32(const float* x, float* y, int n){for(int i=0;i<n;++i)y[i]=pca_log_strict(x[i]);}\n'
                 'extern "C" void e64(const double* x, double* y, int n){for(int i=0;i<n;++i)y[i]=pca_exp_strict(x[i]);}\n'
                 'extern "C" void l64(const double* x, double* y, int n){for(int i=0;i<n;++i)y[i]=pca_log_strict(x[i]);}\n')
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, cpp],
                   check=True)
    L = ctypes.CDLL(so)
    rng = np.random.default_rng(0)
    for fn, dt, lo, hi, ref in (("e32", np.float32, -87, 88, np.exp), ("e64", np.float64, -700, 700, np.exp),
                                ("l32", np.float32, -30, 30, None), ("l64", np.float64, -300, 300, None)):
        if ref is None:
            x = np.exp(rng.uniform(lo, hi, 20000)).astype(dt)
            ref = np.log
        else:
            x = rng.uniform(lo, hi, 20000).astype(dt)
        y = np.zeros_like(x)
        getattr(L, fn)(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), x.size)
        exact = ref(x.astype(np.float64) if dt == np.float32 else x.astype(np.longdouble))
        ulp = np.abs(y - exact.astype(dt)) / np.spacing(np.abs(exact.astype(dt)))
        assert np.max(ulp) <= 4, (fn, np.max(ulp))


@pytest.mark.parametrize("name,M,dtype,strict", [
    ("ivim", 16, "float32", False), ("ivim", 10, "float64", False), ("t2", 32, "float32", False),
    ("se", 16, "float32", True), ("ivim", 16, "float32", True)])
def test_hiprtc_compiles_for_gfx950(name, M, dtype, strict):
    from pycompute_amd import lm
    m = MODELS[name]
    prog = lm.get_program(lm.LMSpec(m.expr, m.pars, m.consts, M, dtype, strict=strict))
    code = prog.module.code
    assert code[:4] == b"\x7fELF" and b"pca_lm_fit" in code and b"gfx950" in code


FACTOR_EXPRS = USER_EXPRS + [(m.expr, m.pars, m.consts) for k, m in sorted(MODELS.items())
                             if len(m.pars) <= 8] + [
    ("p0*c**2 + p1*c*exp(-p2*c) - p0*p1", ["p0", "p1", "p2"], ["c"]),
    ("-(a*b)*exp(-(k*x)*(a*a))/(1 + x)", ["a", "b"], ["k", "x"]),
]


@pytest.mark.parametrize("expr,pars,consts", FACTOR_EXPRS)
def test_factored_split_is_exact(expr, pars, consts):
    """Every Jacobian entry, second derivative and the model split as A(p) u:
    A has only parameters, u has no parameter-only factor, and A u equals the
    unsplit expression at random points."""
    dag = hipgen.res_jac_hes(expr, pars, consts)
    fac = hipgen._Factorer(dag)
    rng = np.random.default_rng(5)
    P, C = _points(dag, rng)
    env = dict(zip(dag.pars, P))
    env.update(zip(dag.consts, C))
    memo = {}
    nz = dag.hes_nonzero()
    for X in [dag.model] + dag.jac + [h for h, z in zip(dag.hes, nz) if z]:
        A, u = fac.factor(X)
        assert A.op == "num" or fac.fs(A) <= set(dag.pars)
        coef, Pf, K, U = fac.split(u)
        assert not Pf and coef == 1.0
        x = np.broadcast_to(E.evaluate(X, env, memo), (64,))
        au = np.broadcast_to(E.evaluate(A, env, memo) * E.evaluate(u, env, memo), (64,))
        np.testing.assert_allclose(au, x, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("lam", [1e-6, 0.37, 1e3])
@pytest.mark.parametrize("expr,pars,consts", FACTOR_EXPRS)
def test_factored_sums_equal_fghhl(expr, pars, consts, lam):
    """The factored sums (g = A sum u r, h = A_j A_k sum u_j u_k + A'' sum u'' r,
    hld = lam A^2 sum u^2 + sum max(1e-4 - lam A^2 u^2, 0) + h_jj) reproduce
    the fghhl loop of the reference (lsqnonlin.py:237-256) in float64, with
    the damping floor active on every measurement (lam 1e-6), on some and on
    none (lam 1e3)."""
    dag = hipgen.res_jac_hes(expr, pars, consts)
    fac = hipgen._Factorer(dag)
    rng = np.random.default_rng(6)
    M = 12
    p = [rng.uniform(0.3, 1.5) for _ in dag.pars]
    cs = [rng.uniform(0.2, 2.0, M) for _ in dag.consts]
    y = rng.uniform(0.0, 2.0, M)
    n, nz = dag.nparam, dag.hes_nonzero()
    env = dict(zip(dag.pars, p))
    env.update(zip(dag.consts, cs))
    ev = lambda node: np.broadcast_to(np.asarray(E.evaluate(node, env, {}), float), (M,))
    r = ev(dag.model) - y
    J = [ev(j) for j in dag.jac]
    d2 = [ev(h) for h in dag.hes]
    g_ref = np.array([np.sum(J[j] * r) for j in range(n)])
    h_ref, hld_ref, l = [], np.zeros(n), 0
    for j in range(n):
        for k in range(j + 1):
            h_ref.append(np.sum(d2[l] * r + J[j] * J[k]))
            if j == k:
                hld_ref[j] = np.sum(d2[l] * r + J[j] * J[k] + np.maximum(lam * J[j] ** 2, 1e-4))
            l += 1
    jf = [fac.factor(x) for x in dag.jac]
    hf = [fac.factor(x) if z else None for x, z in zip(dag.hes, nz)]
    A = [float(np.ravel(ev(a))[0]) for a, _ in jf]
    u = [ev(b) for _, b in jf]
    g = np.array([A[j] * np.sum(u[j] * r) for j in range(n)])
    h, hld, l = [], np.zeros(n), 0
    for j in range(n):
        for k in range(j + 1):
            v = A[j] * A[k] * np.sum(u[j] * u[k])
            if hf[l] is not None:
                v += float(np.ravel(ev(hf[l][0]))[0]) * np.sum(ev(hf[l][1]) * r)
            h.append(v)
            if j == k:
                # the emitted form: L sum u^2 + sum max(1e-4 - L u^2, 0)
                L = lam * A[j] ** 2
                uu = u[j] ** 2
                hld[j] = L * np.sum(uu) + np.sum(np.maximum(1e-4 - L * uu, 0.0)) + v
            l += 1
    np.testing.assert_allclose(g, g_ref, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(h, h_ref, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(hld, hld_ref, rtol=1e-10, atol=1e-12)


def _fghhl_ref(dag, p, cs, y, lam):
    """f, g, h and hld of the reference's fghhl loop (lsqnonlin.py:237-256)
    in float64, straight from the DAG."""
    env = dict(zip(dag.pars, p))
    env.update(zip(dag.consts, cs))
    M = len(y)
    ev = lambda node: np.broadcast_to(np.asarray(E.evaluate(node, env, {}), float), (M,))
    r = ev(dag.model) - y
    J = [ev(j) for j in dag.jac]
    d2 = [ev(h) for h in dag.hes]
    n = dag.nparam
    g = np.array([np.sum(J[j] * r) for j in range(n)])
    h, hld, l = [], np.zeros(n), 0
    for j in range(n):
        for k in range(j + 1):
            h.append(np.sum(d2[l] * r + J[j] * J[k]))
            if j == k:
                hld[j] = np.sum(d2[l] * r + J[j] * J[k] + np.maximum(lam * J[j] ** 2, 1e-4))
            l += 1
    return float(np.sum(r * r)), g, np.array(h), hld


_FACTORED_SHIM = """
#include <cmath>
#define __device__
#define __forceinline__ inline
typedef {real} real;
static inline real pca_fma(real a, real b, real c) {{ return std::fma(a, b, c); }}
static inline real pca_fmax(real a, real b) {{ return std::fmax(a, b); }}
static inline real pca_vgpr(real x) {{ return x; }}
static inline real pca_floor_excess(real L, real uu, real fl) {{
  return std::fmax(std::fma(-L, uu, fl), (real)0);
}}
#define PCA_EXP(x) std::exp(x)
#define PCA_LOG(x) std::log(x)
#define PCA_POW(x, y) std::pow(x, y)
#define PCA_EXPK(k, x) std::exp((k) * (x))
#define PCA_EXPKD(ks, x) {expkd}
// the prelude's exponential recurrence (csrc/prelude.hip), restated for the host
#define PCA_EXP_REC {rec}
template <int R>
static inline real pca_exp_rec(int i, real& e, real& q, real k, real ks, real x) {{
  if (i == 0) q = PCA_EXPKD(ks, x);
  if (i % R == 0) e = PCA_EXPKD(k, x); else e = e * q;
  return e;
}}
{plain}
{derived}
// one iteration's sums as the kernel forms them: prep, accf per measurement
// (consts k of measurement i at C[k*M+i]), finish, then hld += h_jj
// (HLD_DAMP_ONLY) and the SSR through rprep/residf at the same point
template <class Mdl>
static void drive(const real* p, const real* C, const real* y, int M, real lam, real* out) {{
  constexpr int NP = Mdl::NP, NH = Mdl::NH;
  typename Mdl::Pre P;
  Mdl::prep(p, lam, P);
  real s[Mdl::NS] = {{}}, q[NH] = {{}}, dmp[NP] = {{}}, f = 0;
  real ex[Mdl::NE], qx[Mdl::NE];
  for (int i = 0; i < M; ++i) {{
    real c[Mdl::NCT + 1];
    for (int k = 0; k < Mdl::NCT; ++k) c[k] = C[k * M + i];
    Mdl::accf(P, c, y[i], f, s, q, dmp, i, ex, qx);
  }}
  real g[NP], h[NH], hld[NP];
  Mdl::finish(p, P.L, s, q, dmp, g, h, hld);
  for (int j = 0; j < NP; ++j) hld[j] += h[j * (j + 1) / 2 + j];
  typename Mdl::RPre R;
  Mdl::rprep(p, R);
  real fr = 0;
  for (int i = 0; i < M; ++i) {{
    real c[Mdl::NCT + 1];
    for (int k = 0; k < Mdl::NCT; ++k) c[k] = C[k * M + i];
    const real r = Mdl::residf(R, c, y[i], i, ex, qx);
    fr += r * r;
  }}
  int o = 0;
  out[o++] = f;
  out[o++] = fr;
  for (int j = 0; j < NP; ++j) out[o++] = g[j];
  for (int l = 0; l < NH; ++l) out[o++] = h[l];
  for (int j = 0; j < NP; ++j) out[o++] = hld[j];
}}
extern "C" void run_plain(const real* p, const real* C, const real* y, int M, real lam,
                          real* out) {{ drive<Plain>(p, C, y, M, lam, out); }}
extern "C" void run_derived(const real* p, const real* C, const real* y, int M, real lam,
                            real* out) {{ drive<Derived>(p, C, y, M, lam, out); }}
"""


def _build_factored_host(dag, dtype, exp2, tmp_path, exp_rec=0):
    """The emitted factored members built for the host: run_plain (const-only
    subexpressions per measurement) and run_derived (read from derived consts
    rows; with ``exp_rec`` > 0 through the exponential recurrence)."""
    ent = []
    plain = hipgen.emit_model(dag, "Plain", dtype, reassoc=True, factored=True)
    derived = hipgen.emit_model(dag, "Derived", dtype, reassoc=True, factored=True,
                                derived=ent, exp_rec=exp_rec > 0)
    real = "float" if dtype == "float32" else "double"
    expkd = "std::exp2((ks) * (x))" if exp2 else "std::exp((ks) * (x))"
    src = _FACTORED_SHIM.format(real=real, expkd=expkd, plain=plain, derived=derived,
                                rec=max(1, exp_rec))
    cpp, so = tmp_path / "factored.cpp", tmp_path / "factored.so"
    cpp.write_text(src)
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-o", str(so), str(cpp)],
                   check=True)
    return ctypes.CDLL(str(so)), ent


@pytest.mark.parametrize("expr,pars,consts", FACTOR_EXPRS)
def test_factored_emitted_code_matches_fghhl(expr, pars, consts, tmp_path):
    """The emitted factored members (prep/accf/finish, rprep/residf), built
    for the host in float64 and driven as the kernel drives them, give the
    reference's f, g, h, hld and trial SSR, both computing const-only
    subexpressions per measurement and reading them as derived consts rows
    (hipgen.derived_values) after the model's own consts."""
    dag = hipgen.res_jac_hes(expr, pars, consts)
    L, ent = _build_factored_host(dag, "float64", False, tmp_path)
    rng = np.random.default_rng(11)
    M, n, nc = 12, dag.nparam, dag.nconst
    cs = np.array([rng.uniform(0.2, 2.0, M) for _ in dag.consts]).reshape(nc, M)
    table = np.concatenate([cs, hipgen.derived_values(ent, dag, cs, False)]).ravel()
    if nc == 0:
        table = np.zeros(1)
    dp = ctypes.POINTER(ctypes.c_double)
    for lam in (1e-6, 0.37, 1e3):
        p = np.array([rng.uniform(0.3, 1.5) for _ in dag.pars])
        y = rng.uniform(0.0, 2.0, M)
        f, g, h, hld = _fghhl_ref(dag, p, list(cs), y, lam)
        want = np.concatenate([[f, f], g, h, hld])
        for fn, C in ((L.run_plain, cs.ravel() if nc else np.zeros(1)), (L.run_derived, table)):
            out = np.zeros(2 + 2 * n + dag.nhes)
            fn.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_double, dp]
            fn(p.ctypes.data_as(dp), np.ascontiguousarray(C).ctypes.data_as(dp),
               y.ctypes.data_as(dp), M, lam, out.ctypes.data_as(dp))
            np.testing.assert_allclose(out, want, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("rec", [1, 3, 16])
@pytest.mark.parametrize("expr,pars,consts", FACTOR_EXPRS)
def test_exp_recurrence_matches_fghhl(expr, pars, consts, rec, tmp_path):
    """The emitted factored members with the exponential recurrence (exp(k_i
    x) directly every ``rec`` measurements, the previous value times exp(d x)
    in between), built for the host in float64 on consts rows that are
    arithmetic progressions: the reference's f, g, h, hld and trial SSR. A
    model whose exponential factor is not an arithmetic progression in the
    measurement (a product of two consts rows) is refused by
    derived_values()."""
    dag = hipgen.res_jac_hes(expr, pars, consts)
    M, n, nc = 12, dag.nparam, dag.nconst
    cs = np.array([np.linspace(0.2 + 0.1 * k, 2.0 - 0.2 * k, M) for k in range(nc)])
    cs = cs.reshape(nc, M)
    L, ent = _build_factored_host(dag, "float64", False, tmp_path, exp_rec=rec)
    kinds = [k for k, _ in ent]
    if "expk" in kinds:
        assert kinds.count("expstep") == kinds.count("expk")
    try:
        table = np.concatenate([cs, hipgen.derived_values(ent, dag, cs, False)]).ravel()
    except ValueError as err:
        assert "arithmetic progression" in str(err)
        rows = hipgen.derived_values([e for e in ent if e[0] == "expk"], dag, cs, False)
        dev = [np.max(np.abs(r - np.linspace(r[0], r[-1], M))) for r in rows]
        assert max(dev) > hipgen.EXP_REC_TOL * np.max(np.abs(rows))
        return
    if nc == 0:
        table = np.zeros(1)
    rng = np.random.default_rng(17)
    dp = ctypes.POINTER(ctypes.c_double)
    for lam in (1e-6, 0.37, 1e3):
        p = np.array([rng.uniform(0.3, 1.5) for _ in dag.pars])
        y = rng.uniform(0.0, 2.0, M)
        f, g, h, hld = _fghhl_ref(dag, p, list(cs), y, lam)
        want = np.concatenate([[f, f], g, h, hld])
        out = np.zeros(2 + 2 * n + dag.nhes)
        L.run_derived.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_double, dp]
        L.run_derived(p.ctypes.data_as(dp), np.ascontiguousarray(table).ctypes.data_as(dp),
                      y.ctypes.data_as(dp), M, lam, out.ctypes.data_as(dp))
        np.testing.assert_allclose(out, want, rtol=1e-10, atol=1e-11)


def test_arithmetic_step():
    """The common difference of an exponential-recurrence row: exact rows and
    float32-rounded linspace rows pass, a row with one value moved by more
    than the rounding of float32 consts does not; a single measurement has
    step 0."""
    v = np.linspace(0.0, 800.0, 16)
    assert hipgen.arithmetic_step(v) == pytest.approx(800.0 / 15, rel=1e-15)
    v32 = np.linspace(0.0, 800.0, 16).astype(np.float32).astype(np.float64)
    assert hipgen.arithmetic_step(v32) == pytest.approx(800.0 / 15, rel=1e-7)
    assert hipgen.arithmetic_step(np.array([3.0])) == 0.0
    bad = v.copy()
    bad[7] += 800.0 * 2.0 ** -20
    with pytest.raises(ValueError, match="arithmetic progression"):
        hipgen.arithmetic_step(bad)
    with pytest.raises(ValueError, match="arithmetic progression"):
        hipgen.arithmetic_step(np.array([0.0, 10.0, 20.0, 50.0, 100.0]))


@pytest.mark.parametrize("name", ["ivim", "t2"])
def test_exp_recurrence_float32(name, tmp_path):
    """Float32 exp2 form with the exponential recurrence (anchors every 8
    measurements) on the benchmark protocols (16 b-values 0-800 s/mm^2, 32
    echoes 10-320 ms): the float64 reference at the float32 bar, and within a
    few ulp-level errors of the direct float32 evaluation."""
    m = MODELS[name]
    dag = hipgen.res_jac_hes(m.expr, m.pars, m.consts)
    (tmp_path / "r").mkdir()
    (tmp_path / "d").mkdir()
    Lr, ent_r = _build_factored_host(dag, "float32", True, tmp_path / "r", exp_rec=8)
    Ld, ent_d = _build_factored_host(dag, "float32", True, tmp_path / "d")
    M = 16 if name == "ivim" else 32
    cs = (np.linspace(0.0, 800.0, M) * 1e-3 if name == "ivim"
          else np.linspace(10.0, 320.0, M))[None, :]
    p = np.array({"ivim": [1.0, 0.2, 3.0, 0.8], "t2": [1.0, 60.0]}[name])
    rng = np.random.default_rng(4)
    f0, _, _, _ = _fghhl_ref(dag, p, list(cs), np.zeros(M), 0.1)
    y = np.sqrt(f0 / M) * rng.uniform(0.5, 1.5, M)
    f, g, h, hld = _fghhl_ref(dag, p, list(cs), y, 0.1)
    want = np.concatenate([[f, f], g, h, hld])
    fp = ctypes.POINTER(ctypes.c_float)
    outs = []
    for L, ent in ((Lr, ent_r), (Ld, ent_d)):
        table = np.concatenate([cs, hipgen.derived_values(ent, dag, cs, True)]).astype(np.float32)
        out = np.zeros(len(want), dtype=np.float32)
        L.run_derived.argtypes = [fp, fp, fp, ctypes.c_int, ctypes.c_float, fp]
        p32, y32 = p.astype(np.float32), y.astype(np.float32)
        L.run_derived(p32.ctypes.data_as(fp),
                      np.ascontiguousarray(table.ravel()).ctypes.data_as(fp),
                      y32.ctypes.data_as(fp), M, 0.1, out.ctypes.data_as(fp))
        np.testing.assert_allclose(out, want, rtol=2e-4, atol=2e-4 * np.max(np.abs(want)))
        outs.append(out.astype(np.float64))
    np.testing.assert_allclose(outs[0], outs[1], rtol=2e-5, atol=2e-5 * np.max(np.abs(want)))


def test_derived_consts_of_builtin_models():
    """IVIM reads -b (the exp factor, exp2-scaled) and b*b from derived rows,
    T2 reads -TE; their host values, and the float32 exp2 form of the
    emitted code against the float64 reference."""
    from pycompute_amd import lm
    dag = hipgen.res_jac_hes(IVIM.expr, IVIM.pars, IVIM.consts)
    ent = []
    hipgen.emit_model(dag, "X", "float32", reassoc=True, factored=True, derived=ent)
    assert [(k, E.to_str(x)) for k, x in ent] == [("expk", "(-b)"), ("val", "(b*b)")]
    b = np.linspace(0.0, 800.0, 16)
    v = hipgen.derived_values(ent, dag, b[None, :], True)
    np.testing.assert_allclose(v[0], -b / np.log(2.0), rtol=1e-15)
    np.testing.assert_allclose(v[1], b * b, rtol=1e-15)
    np.testing.assert_allclose(hipgen.derived_values(ent, dag, b[None, :], False)[0], -b)
    t2 = MODELS["t2"]
    d2 = hipgen.res_jac_hes(t2.expr, t2.pars, t2.consts)
    e2 = []
    hipgen.emit_model(d2, "Y", "float32", reassoc=True, factored=True, derived=e2)
    assert [(k, E.to_str(x)) for k, x in e2] == [("expk", "(-TE)")]
    # the module lists its derived rows; per-voxel and unfactored builds derive none
    src_ent = []
    spec = lm.LMSpec(IVIM.expr, IVIM.pars, IVIM.consts, 16, fast_exp=True, fast_div=True,
                     reassoc=True, factored=True)
    src = lm.build_source(spec, dag, (), src_ent)
    assert len(src_ent) == 1 and len(src_ent[0]) == 2
    assert "derived const c[1]" in src and "derived const c[2]" in src
    for s in (lm.LMSpec(IVIM.expr, IVIM.pars, IVIM.consts, 16, fast_exp=True, reassoc=True,
                        factored=True, consts_per_voxel=True),
              lm.LMSpec(IVIM.expr, IVIM.pars, IVIM.consts, 16, fast_exp=True, reassoc=True)):
        e = []
        lm.build_source(s, dag, (), e)
        assert e == [[]]
    # a table that would not fit the consts block keeps the per-measurement form
    big = []
    lm.build_source(lm.LMSpec(IVIM.expr, IVIM.pars, IVIM.consts, 100, fast_exp=True,
                              reassoc=True, factored=True), dag, (), big)
    assert big == [[]]


@pytest.mark.parametrize("name", ["ivim", "t2", "se"])
def test_factored_derived_float32_exp2(name, tmp_path):
    """Float32 build of the emitted code with exp as exp2 of the host-scaled
    factor (the fast-exp form) against the float64 reference, at the bar
    float32 sums of 16 measurements meet."""
    m = MODELS[name]
    dag = hipgen.res_jac_hes(m.expr, m.pars, m.consts)
    L, ent = _build_factored_host(dag, "float32", True, tmp_path)
    M = 16
    cs = (np.linspace(50.0, 800.0, M) * (1e-3 if name != "t2" else 0.1))[None, :]
    table = np.concatenate([cs, hipgen.derived_values(ent, dag, cs, True)]).astype(np.float32)
    p = {"ivim": [1.0, 0.2, 3.0, 0.8], "t2": [1.0, 8.0], "se": [1.0, 1.2, 0.7]}[name]
    p = np.array(p)
    rng = np.random.default_rng(3)
    f0, _, _, _ = _fghhl_ref(dag, p, list(cs), np.zeros(M), 0.1)
    y = np.sqrt(f0 / M) * rng.uniform(0.5, 1.5, M)
    f, g, h, hld = _fghhl_ref(dag, p, list(cs), y, 0.1)
    want = np.concatenate([[f, f], g, h, hld])
    fp = ctypes.POINTER(ctypes.c_float)
    out = np.zeros(len(want), dtype=np.float32)
    L.run_derived.argtypes = [fp, fp, fp, ctypes.c_int, ctypes.c_float, fp]
    p32, y32 = p.astype(np.float32), y.astype(np.float32)
    L.run_derived(p32.ctypes.data_as(fp), np.ascontiguousarray(table.ravel()).ctypes.data_as(fp),
                  y32.ctypes.data_as(fp), M, 0.1, out.ctypes.data_as(fp))
    np.testing.assert_allclose(out, want, rtol=2e-4, atol=2e-4 * np.max(np.abs(want)))


@pytest.mark.parametrize("name,M", [("ivim", 16), ("t2", 32), ("se", 16)])
def test_hiprtc_compiles_factored(name, M):
    from pycompute_amd import lm
    m = MODELS[name]
    for sg in (0, 4):
        prog = lm.get_program(lm.LMSpec(m.expr, m.pars, m.consts, M, fast_exp=True,
                                        fast_div=True, reassoc=True, factored=True,
                                        sched_group=sg))
        assert b"pca_lm_chunk" in prog.module.code
    with pytest.raises(ValueError):
        lm.get_program(lm.LMSpec(m.expr, m.pars, m.consts, M, factored=True))
