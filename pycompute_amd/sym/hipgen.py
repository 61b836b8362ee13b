"""HIP source generation for a model expression.

Replaces the reference's printer and fghhl/f templates
(pycompute/sym/util.py:32-59, pycompute/cuda/lsqnonlin.py:46-117, 206-346):

* ``res_jac_hes`` builds, for ``model(pars; consts)``, the model value, its
  gradient and its packed lower-triangle Hessian (l = j(j+1)/2 + k, k <= j),
  in one hash-consed DAG (the CSE of sym/util.py:100-112).
* ``emit_model`` prints a HIP ``struct`` with straight-line device code:
    - ``resid(p, c, y)``          model - y              (the F kernel body)
    - ``accum(p, c, y, lam, f, g, h, hld)``  one measurement of the fghhl loop:
        f += r*r; for j, k<=j: jt = J_j J_k, hj = H_jk r + jt, h_l += hj,
        hld_j += hj + max(lam jt, 1e-4) on the diagonal; g_j += J_j r
      (lsqnonlin.py:237-256). Structurally zero Hessian entries are dropped at
      generation time (the reference multiplies a literal 0 by r).
    - ``fjh(p, c, m, J, d2)``     model, Jacobian, second derivatives
* Float builds print single-precision literals and functions throughout (the
  reference leaves decimal literals as double and maps rsqrt to "rsqrtff").
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import expr as E

_FN_F = {"exp": "PCA_EXP", "log": "PCA_LOG", "sqrt": "sqrtf", "sin": "sinf", "cos": "cosf",
         "tan": "tanf", "sinh": "sinhf", "cosh": "coshf", "tanh": "tanhf", "asin": "asinf",
         "acos": "acosf", "atan": "atanf", "abs": "fabsf"}
_FN_D = {"exp": "PCA_EXP", "log": "PCA_LOG", "sqrt": "sqrt", "sin": "sin", "cos": "cos", "tan": "tan",
         "sinh": "sinh", "cosh": "cosh", "tanh": "tanh", "asin": "asin", "acos": "acos",
         "atan": "atan", "abs": "fabs"}


def expr_hash(text: str, length: int = 12) -> str:
    """Kernel-name hash of an expression: md5 hex digest characters [1, 1+length)
    (pycompute/sym/util.py:61-62)."""
    return hashlib.md5(text.encode()).hexdigest()[1:length + 1]


@dataclass
class ModelDAG:
    expr: str
    pars: list
    consts: list
    graph: E.Graph
    model: E.Node
    jac: list
    hes: list                     # packed lower triangle, len n(n+1)/2
    counts_fjh: dict = field(default_factory=dict)
    counts_f: dict = field(default_factory=dict)

    @property
    def nparam(self):
        return len(self.pars)

    @property
    def nconst(self):
        return len(self.consts)

    @property
    def nhes(self):
        return len(self.hes)

    def hes_nonzero(self):
        return [not E.is_zero(h) for h in self.hes]


def res_jac_hes(expr: str, pars, consts) -> ModelDAG:
    """Model, gradient and packed Hessian of ``expr`` w.r.t. ``pars``."""
    pars, consts = list(pars), list(consts)
    if len(set(pars) | set(consts)) != len(pars) + len(consts):
        raise ValueError("parameter and constant names must be distinct")
    g = E.Graph()
    m = E.parse(g, expr)
    unknown = E.free_symbols(m) - set(pars) - set(consts)
    if unknown:
        raise ValueError(f"expression uses undeclared symbols: {sorted(unknown)}")
    memo: dict = {}
    jac = [E.diff(g, m, p, memo) for p in pars]
    hes = []
    for j in range(len(pars)):
        for k in range(j + 1):
            hes.append(E.diff(g, jac[j], pars[k], memo))
    dag = ModelDAG(expr, pars, consts, g, m, jac, hes)
    dag.counts_fjh = E.op_counts([m] + jac + hes)
    dag.counts_f = E.op_counts([m])
    return dag


def _lit(v: float, dtype: str) -> str:
    if dtype == "float32":
        f = np.float32(v)
        if not np.isfinite(f):
            raise ValueError(f"constant {v} is not representable in float32")
        s = repr(float(f)) if float(f) == int(f) and abs(f) < 1e7 else np.format_float_positional(
            f, unique=True, trim="0") if 1e-4 <= abs(f) < 1e7 else np.format_float_scientific(
            f, unique=True, trim="0")
        if "." not in s and "e" not in s:
            s += ".0"
        return s + "f"
    s = repr(float(v))
    if "." not in s and "e" not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


class _Printer:
    """Straight-line emission of a DAG: one SSA temporary per interior node."""

    def __init__(self, dag: ModelDAG, dtype: str, p_name="p", c_name="c"):
        self.dag = dag
        self.dtype = dtype
        self.fn = _FN_F if dtype == "float32" else _FN_D
        self.pow = "PCA_POW"
        self.sym = {name: f"{p_name}[{k}]" for k, name in enumerate(dag.pars)}
        self.sym.update({name: f"{c_name}[{k}]" for k, name in enumerate(dag.consts)})
        self.names: dict = {}
        self.lines: list = []
        self.count = 0

    def ref(self, n: E.Node) -> str:
        if n.op == "num":
            return _lit(n.value, self.dtype)
        if n.op == "sym":
            return self.sym[n.name]
        return self.names[n.id]

    def _tmp(self, text: str) -> str:
        name = f"t{self.count}"
        self.count += 1
        self.lines.append(f"const real {name} = {text};")
        return name

    def emit(self, outputs):
        for n in E.topo(outputs):
            if n.op in ("num", "sym") or n.id in self.names:
                continue
            self.names[n.id] = self._tmp(self._format(n, [self.ref(x) for x in n.args]))

    def _format(self, n: E.Node, a: list) -> str:
        if n.op == "add":
            txt = f"{a[0]} + {a[1]}"
        elif n.op == "neg":
            txt = f"-{a[0]}"
        elif n.op == "mul":
            txt = f"{a[0]}*{a[1]}"
        elif n.op == "div":
            txt = f"{a[0]}/{a[1]}"
        elif n.op == "pow":
            e = E.int_power(n)
            if e is not None:
                txt = self._int_pow(a[0], abs(e))
                if e < 0:
                    txt = f"{_lit(1.0, self.dtype)}/({txt})"
            elif n.args[1].op == "num" and n.args[1].value == 0.5:
                txt = f"{self.fn['sqrt']}({a[0]})"
            else:
                txt = f"{self.pow}({a[0]}, {a[1]})"
        elif n.op == "fn":
            txt = f"{self.fn[n.name]}({a[0]})"
        else:  # pragma: no cover
            raise ValueError(n.op)
        return txt

    def _int_pow(self, x: str, k: int) -> str:
        # exponentiation by squaring on SSA temporaries
        result, base = None, x
        while True:
            if k & 1:
                result = base if result is None else self._tmp(f"{result}*{base}")
            k >>= 1
            if not k:
                break
            base = self._tmp(f"{base}*{base}")
        return result


class _Factorer:
    """Splits a product into its numeric coefficient and its parameter-only,
    const-only and mixed factors: n = coef * prod(P) * prod(K) * prod(U)."""

    def __init__(self, dag: ModelDAG):
        self.g = dag.graph
        self.pars, self.consts = frozenset(dag.pars), frozenset(dag.consts)
        self._fs: dict = {}

    def fs(self, n: E.Node) -> frozenset:
        r = self._fs.get(n.id)
        if r is None:
            for x in E.topo([n]):
                if x.id in self._fs:
                    continue
                if x.op == "sym":
                    self._fs[x.id] = frozenset([x.name])
                else:
                    self._fs[x.id] = frozenset().union(*[self._fs[a.id] for a in x.args])
            r = self._fs[n.id]
        return r

    def param_only(self, n: E.Node) -> bool:
        return n.op != "num" and self.fs(n) <= self.pars

    def split(self, n: E.Node):
        """A quotient a/b is split as a * (1/b_1) * (1/b_2) ..., b_i the
        factors of b, so that a parameter-only divisor becomes a per-iteration
        reciprocal and a const-only one joins the consts factor."""
        coef, P, K, U = 1.0, [], [], []
        stack = [n]
        while stack:
            x = stack.pop()
            if x.op == "num":
                coef *= x.value
            elif x.op == "neg":
                coef = -coef
                stack.append(x.args[0])
            elif x.op == "mul":
                stack.extend(x.args)
            elif x.op == "div" and not (x.args[1].op == "num" and x.args[1].value == 0.0):
                stack.append(x.args[0])
                dc, dP, dK, dU = self.split(x.args[1])
                if dc == 0.0:
                    U.append(x)
                    continue
                coef /= dc
                one = self.g.num(1)
                for group, dst in ((dP, P), (dK, K), (dU, U)):
                    for d in group:
                        dst.append(self.g.div(one, d))
            else:
                f = self.fs(x)
                (P if f <= self.pars else K if f <= self.consts else U).append(x)
        return coef, P, K, U

    def prod(self, nodes) -> E.Node:
        out = self.g.num(1)
        for x in sorted(nodes, key=lambda z: z.id):
            out = self.g.mul(out, x)
        return out

    def factor(self, n: E.Node):
        """(A, u) with n = A * u, A parameter-only (or a number) and u free
        of parameter-only factors."""
        coef, P, K, U = self.split(n)
        return self.g.mul(self.g.num(coef), self.prod(P)), self.prod(K + U)


class DerivedConsts:
    """Const-only subexpressions of the per-measurement code, evaluated once
    per protocol on the host and read as extra consts c[NC + i] (rows of the
    voxel-uniform consts table after the model's own NC rows).

    entries[i] = (kind, node): kind "val" is the node's value; kind "expk"
    is the factor k of an exp(k*x), scaled by log2(e) when the fast-exp
    float build evaluates exp through exp2 (PCA_EXPKD); kind "expstep" is the
    common difference d of the "expk" row of the same node, which must be an
    arithmetic progression k_i = k_0 + i d over the measurements (the
    exponential recurrence, pca_exp_rec), in every column of its row."""

    def __init__(self, dag: ModelDAG, fac: "_Factorer"):
        self.nc = dag.nconst
        self.fac = fac
        self.index: dict = {}
        self.entries: list = []

    def hoists(self, n: E.Node) -> bool:
        """An interior node over consts only that costs an instruction per
        measurement (a negated const is a free source modifier)."""
        if n.op in ("num", "sym") or (n.op == "neg" and n.args[0].op == "sym"):
            return False
        return self.fac.fs(n) <= self.fac.consts

    def ref(self, kind: str, n: E.Node) -> str:
        i = self.index.get((kind, n.id))
        if i is None:
            i = self.index[(kind, n.id)] = len(self.entries)
            self.entries.append((kind, n))
        return f"c[{self.nc + i}]"


def derived_values(entries, dag: ModelDAG, consts_rows, exp2_scale: bool):
    """Values of the DerivedConsts entries for one protocol, in float64:
    consts_rows is (NC, M), one row per const; returns (len(entries), M)."""
    C = np.asarray(consts_rows, dtype=np.float64).reshape(dag.nconst, -1)
    M = C.shape[1]
    env = dict(zip(dag.consts, C))
    memo: dict = {}
    out = np.empty((len(entries), M))
    for i, (kind, n) in enumerate(entries):
        v = np.broadcast_to(np.asarray(E.evaluate(n, env, memo), dtype=np.float64), (M,))
        v = v * (1.0 / np.log(2.0) if kind in ("expk", "expstep") and exp2_scale else 1.0)
        if kind == "expstep":
            v = np.full(M, arithmetic_step(v, E.to_str(n)))
        out[i] = v
    return out


# Largest deviation of an exponential-recurrence row from its arithmetic
# progression, relative to the row's largest magnitude: the rounding of float32
# consts (half an ulp each) passes, anything else does not.
EXP_REC_TOL = 2.0 ** -22


def arithmetic_step(v, what: str = "row") -> float:
    """The common difference d of a row v_i = v_0 + i d (float64), or
    ValueError when the row is not an arithmetic progression to within
    EXP_REC_TOL of its largest magnitude."""
    v = np.asarray(v, dtype=np.float64)
    M = len(v)
    if M < 2:
        return 0.0
    d = (v[-1] - v[0]) / (M - 1)
    dev = np.max(np.abs(v - (v[0] + d * np.arange(M))))
    if not np.all(np.isfinite(v)) or dev > EXP_REC_TOL * np.max(np.abs(v)):
        raise ValueError(f"the exponential recurrence needs the consts factor {what} to be "
                         f"an arithmetic progression over the measurements "
                         f"(largest deviation {dev:.3g})")
    return float(d)


class _MeasPrinter(_Printer):
    """Per-measurement code of the factored accumulation: parameter-only
    subexpressions are read from the per-iteration table ``{pre}.a[i]``
    (collected in ``pre_nodes``), and exp(k*x) with a const-only factor k is
    printed as PCA_EXPK(k, x) so that the scaled k is loop-invariant. With a
    DerivedConsts registry, const-only subexpressions are read from the
    consts table instead and exp(k*x) becomes PCA_EXPKD(scaled k, x)."""

    def __init__(self, dag, dtype, fac: _Factorer, pre: str, derived: DerivedConsts | None = None,
                 exp_rec: bool = False):
        super().__init__(dag, dtype)
        self.fac, self.pre_name = fac, pre
        self.pre: dict = {}
        self.pre_nodes: list = []
        self._expk: dict = {}
        self.derived = derived
        # exponential recurrence: one running value per distinct exp(k*x)
        self.exp_rec = exp_rec and derived is not None
        self.rec_terms: dict = {}

    def _hoisted(self, n: E.Node) -> bool:
        return self.derived is not None and self.derived.hoists(n)

    def ref(self, n: E.Node) -> str:
        if self.fac.param_only(n):
            i = self.pre.get(n.id)
            if i is None:
                i = self.pre[n.id] = len(self.pre_nodes)
                self.pre_nodes.append(n)
            return f"{self.pre_name}.a[{i}]"
        if self._hoisted(n):
            return self.derived.ref("val", n)
        return super().ref(n)

    def _children(self, n: E.Node):
        if n.op in ("num", "sym") or self.fac.param_only(n) or self._hoisted(n):
            return ()
        if n.op == "div" and self.fac.param_only(n.args[1]):
            return (n.args[0],)
        if n.op == "fn" and n.name == "exp":
            coef, P, K, U = self.fac.split(n.args[0])
            if K:
                g = self.fac.g
                k = g.mul(g.num(coef), self.fac.prod(K))
                x = g.mul(self.fac.prod(P), self.fac.prod(U))
                self._expk[n.id] = (k, x)
                return (x,) if self.derived is not None else (k, x)
        return n.args

    def emit(self, outputs):
        seen = set()
        for o in outputs:
            stack = [(o, False)]
            while stack:
                node, done = stack.pop()
                if done:
                    if (node.op in ("num", "sym") or self.fac.param_only(node)
                            or self._hoisted(node)):
                        continue
                    if node.id in self._expk:
                        k, x = self._expk[node.id]
                        if self.exp_rec:
                            e = self.rec_terms.setdefault((k.id, x.id), len(self.rec_terms))
                            txt = (f"pca_exp_rec<PCA_EXP_REC>(i, ex[{e}], qx[{e}], "
                                   f"{self.derived.ref('expk', k)}, "
                                   f"{self.derived.ref('expstep', k)}, {self.ref(x)})")
                        elif self.derived is not None:
                            txt = f"PCA_EXPKD({self.derived.ref('expk', k)}, {self.ref(x)})"
                        else:
                            txt = f"PCA_EXPK({self.ref(k)}, {self.ref(x)})"
                    elif node.op == "div" and self.fac.param_only(node.args[1]):
                        # a parameter-only divisor: multiply by its reciprocal,
                        # computed once per iteration in the table
                        inv = self.fac.g.div(self.fac.g.num(1), node.args[1])
                        txt = f"{self.ref(node.args[0])}*{self.ref(inv)}"
                    else:
                        txt = self._format(node, [self.ref(a) for a in node.args])
                    self.names[node.id] = self._tmp(txt)
                    continue
                if node.id in seen:
                    continue
                seen.add(node.id)
                stack.append((node, True))
                for c in reversed(self._children(node)):
                    if c.id not in seen:
                        stack.append((c, False))


def _emit_factored(dag: ModelDAG, dtype: str, derive: bool = False, exp_rec: bool = False):
    """Members of the factored accumulation (fast builds), and the
    DerivedConsts entries they read (empty unless ``derive``).

    Every Jacobian entry and second derivative X is split as X = A(p) * u,
    A the product of its parameter-only factors (and sign), u the rest. The
    measurement loop then accumulates only u-products, and the per-iteration
    factors are applied once, after the loop:
        g_j  = A_j  * sum u_j r
        h_jk = A_j A_k * sum u_j u_k + A''_jk * sum u''_jk r
        hld_j = sum max(L_j u_j^2, 1e-4),  L_j = lam A_j^2
              = L_j sum u_j^2 + sum max(1e-4 - L_j u_j^2, 0)
    The damping keeps the reference's per-measurement floor; in the second
    form the per-measurement part is one clamped FMA (its value lies in
    [0, 1e-4], so the clamp to [0, 1] is the max with 0) and the sum of
    u_j^2 is the Gauss-Newton sum already accumulated for h_jj.
    Sums over the same u are shared (a second derivative often has the u of a
    Jacobian entry). The model m = A_m u_m gives r = fma(A_m, u_m, -y). Equal
    to accum() up to rounding.

    accf and residf take the measurement index i and the running values ex,
    qx of the exponential recurrence (``exp_rec``, with ``derive``): each
    exp(k*x) whose factor row k is an arithmetic progression is evaluated
    through pca_exp_rec (NE terms); without it they are unused."""
    n, nh = dag.nparam, dag.nhes
    fac = _Factorer(dag)
    g = dag.graph
    floor = _lit(1e-4, dtype)
    nz = dag.hes_nonzero()
    mA, mu = fac.factor(dag.model)
    jf = [fac.factor(J) for J in dag.jac]
    hf = [fac.factor(h) if z else None for h, z in zip(dag.hes, nz)]

    # one accumulator per distinct u multiplied by r
    sidx: dict = {}
    for A, u in jf + [x for x in hf if x is not None]:
        sidx.setdefault(u.id, (len(sidx), u))

    def resid(pr):
        if mA.op == "num" and mA.value == 1.0:
            return f"{pr.ref(mu)} - y"
        return f"pca_fma({pr.ref(mA)}, {pr.ref(mu)}, -y)"

    # accf: one measurement
    dc = DerivedConsts(dag, fac) if derive else None
    pm = _MeasPrinter(dag, dtype, fac, "P", dc, exp_rec)
    pm.emit([mu] + [u for _, u in jf] + [x[1] for x in hf if x is not None])
    acc = list(pm.lines)
    acc.append(f"const real r = {resid(pm)};")
    acc.append("f = pca_fma(r, r, f);")
    for i, u in sidx.values():
        acc.append(f"s[{i}] = pca_fma({pm.ref(u)}, r, s[{i}]);")
    l = 0
    for j in range(n):
        uj = pm.ref(jf[j][1])
        for k in range(j + 1):
            if j == k:
                acc.append(f"{{ const real uu = {uj}*{uj}; q[{l}] += uu; "
                           f"dmp[{j}] += pca_floor_excess(P.L[{j}], uu, P.fl); }}")
            else:
                acc.append(f"q[{l}] = pca_fma({uj}, {pm.ref(jf[k][1])}, q[{l}]);")
            l += 1

    # prep: the per-iteration table and lam A_j^2
    pp = _Printer(dag, dtype)
    pp.emit(pm.pre_nodes + [A for A, _ in jf])
    prep = list(pp.lines)
    prep += [f"P.a[{i}] = {pp.ref(x)};" for i, x in enumerate(pm.pre_nodes)]
    prep += [f"{{ const real a = {pp.ref(A)}; P.L[{j}] = lam*(a*a); }}"
             for j, (A, _) in enumerate(jf)]
    prep.append(f"P.fl = pca_vgpr((real){floor});")

    # finish: apply the factors
    pf = _Printer(dag, dtype)
    prods = []
    for j in range(n):
        for k in range(j + 1):
            prods.append(g.mul(jf[j][0], jf[k][0]))
    pf.emit([A for A, _ in jf] + prods + [x[0] for x in hf if x is not None])
    fin = list(pf.lines)
    for j in range(n):
        fin.append(f"g[{j}] = {pf.ref(jf[j][0])}*s[{sidx[jf[j][1].id][0]}];")
        jj = j * (j + 1) // 2 + j
        fin.append(f"hld[{j}] = pca_fma(L[{j}], q[{jj}], dmp[{j}]);")
    for l in range(nh):
        gn = f"{pf.ref(prods[l])}*q[{l}]"
        if hf[l] is None:
            fin.append(f"h[{l}] = {gn};")
        else:
            A2, u2 = hf[l]
            fin.append(f"h[{l}] = pca_fma({pf.ref(A2)}, s[{sidx[u2.id][0]}], {gn});")

    # trial residual: rprep once per trial point, residf per measurement
    pr = _MeasPrinter(dag, dtype, fac, "R", dc, exp_rec)
    pr.emit([mu])
    rf = list(pr.lines) + [f"return {resid(pr)};"]
    prr = _Printer(dag, dtype)
    prr.emit(pr.pre_nodes)
    rp = list(prr.lines) + [f"R.a[{i}] = {prr.ref(x)};" for i, x in enumerate(pr.pre_nodes)]

    ind = lambda lines: ["    " + x for x in lines]
    nd = len(dc.entries) if dc is not None else 0
    ne = max(len(pm.rec_terms), len(pr.rec_terms))
    out = ["  static constexpr bool FACTORED = true;",
           f"  static constexpr int NCT = {dag.nconst + nd};",
           f"  static constexpr int NS = {max(1, len(sidx))};",
           f"  static constexpr int NE = {max(1, ne)};",
           f"  static constexpr bool EXP_REC = {'true' if ne else 'false'};",
           f"  struct Pre {{ real a[{max(1, len(pm.pre_nodes))}]; real L[{n}]; real fl; }};",
           f"  struct RPre {{ real a[{max(1, len(pr.pre_nodes))}]; }};",
           "  __device__ static __forceinline__ void prep(const real* __restrict__ p, real lam, "
           "Pre& P) {"] + ind(prep) + ["  }"]
    out += ["  __device__ static __forceinline__ void accf(const Pre& P, "
            "const real* __restrict__ c, real y, real& f, real* s, real* q, real* dmp, "
            "int i, real* ex, real* qx) {"]
    out += ind(acc) + ["  }"]
    out += ["  __device__ static __forceinline__ void finish(const real* __restrict__ p, "
            "const real* L, const real* s, const real* q, const real* dmp, real* g, real* h, "
            "real* hld) {"]
    out += ind(fin) + ["  }"]
    out += ["  __device__ static __forceinline__ void rprep(const real* __restrict__ p, RPre& R) {"]
    out += ind(rp) + ["  }"]
    out += ["  __device__ static __forceinline__ real residf(const RPre& R, "
            "const real* __restrict__ c, real y, int i, real* ex, real* qx) {"]
    out += ind(rf) + ["  }"]
    return out, (dc.entries if dc is not None else [])


def emit_model(dag: ModelDAG, struct_name: str, dtype: str = "float32",
               reassoc: bool = False, factored: bool = False,
               derived: list | None = None, exp_rec: bool = False) -> str:
    """HIP source of the model struct (see module docstring).

    reassoc=False keeps the reference's association of the per-measurement
    update: hj = d2f*r + J_j J_k; h += hj; hld += hj + max(lam J_j^2, 1e-4).
    reassoc=True (fast builds) folds it into FMAs, h = fma(J_j, J_k,
    fma(d2f, r, h)), and accumulates only the damping terms into hld (the
    kernel adds the diagonal of h after the measurement loop,
    HLD_DAMP_ONLY): fewer instructions, results equal up to rounding.
    factored=True (with reassoc) adds the factored accumulation
    (_emit_factored), FACTORED = true. With a list as ``derived`` (factored
    builds with voxel-uniform consts), the accumulation reads its const-only
    subexpressions from consts rows NC .. NCT-1, and the list receives their
    DerivedConsts entries (derived_values() computes the rows); NCT = NC
    otherwise. exp_rec=True (with ``derived``) evaluates exp(k*x) through the
    exponential recurrence (pca_exp_rec, the kernel's PCA_EXP_REC), which
    needs every such factor row k to be an arithmetic progression."""
    if factored and not reassoc:
        raise ValueError("the factored accumulation is a reassociated (fast) form")
    n, nh = dag.nparam, dag.nhes
    floor = _lit(1e-4, dtype)
    out = [f"struct {struct_name} {{",
           f"  static constexpr int NP = {n};",
           f"  static constexpr int NC = {dag.nconst};",
           f"  static constexpr int NH = {nh};",
           f"  static constexpr bool HLD_DAMP_ONLY = {'true' if reassoc else 'false'};",
           # structurally non-zero second derivatives (packed order)
           "  static constexpr bool HNZ[NH] = {"
           + ", ".join("true" if z else "false" for z in dag.hes_nonzero()) + "};"]

    # resid: model - y (the F kernel body, lsqnonlin.py:60-66)
    pr = _Printer(dag, dtype)
    pr.emit([dag.model])
    body = pr.lines + [f"return {pr.ref(dag.model)} - y;"]
    out.append("  __device__ static __forceinline__ real resid(const real* __restrict__ p, "
               "const real* __restrict__ c, real y) {")
    out += ["    " + s for s in body]
    out.append("  }")

    # accum: one measurement of fghhl (lsqnonlin.py:214-258)
    pr = _Printer(dag, dtype)
    nz = dag.hes_nonzero()
    pr.emit([dag.model] + dag.jac + [h for h, z in zip(dag.hes, nz) if z])
    body = list(pr.lines)
    body.append(f"const real r = {pr.ref(dag.model)} - y;")
    body.append("f += r*r;")
    l = 0
    for j in range(n):
        Jj = pr.ref(dag.jac[j])
        for k in range(j + 1):
            Jk = pr.ref(dag.jac[k])
            if reassoc:
                acc = f"pca_fma({pr.ref(dag.hes[l])}, r, h[{l}])" if nz[l] else f"h[{l}]"
                if j == k:
                    body.append(f"{{ const real jt{l} = {Jj}*{Jj}; h[{l}] = {acc} + jt{l}; "
                                f"hld[{j}] += pca_fmax(lam*jt{l}, {floor}); }}")
                else:
                    body.append(f"h[{l}] = pca_fma({Jj}, {Jk}, {acc});")
            else:
                if nz[l]:
                    hj = f"{pr.ref(dag.hes[l])}*r + jt{l}"
                else:
                    hj = f"jt{l}"
                body.append(f"{{ const real jt{l} = {Jj}*{Jk}; const real hj = {hj}; "
                            f"h[{l}] += hj;"
                            + (f" hld[{j}] += hj + pca_fmax(lam*jt{l}, {floor});"
                               if j == k else "")
                            + " }")
            l += 1
        body.append(f"g[{j}] = pca_fma({Jj}, r, g[{j}]);" if reassoc else f"g[{j}] += {Jj}*r;")
    out.append("  __device__ static __forceinline__ void accum(const real* __restrict__ p, "
               "const real* __restrict__ c, real y, real lam, real& f, real* g, real* h, "
               "real* hld) {")
    out += ["    " + s for s in body]
    out.append("  }")

    # fjh: model, Jacobian, second derivatives (for the eval kernels)
    pr = _Printer(dag, dtype)
    pr.emit([dag.model] + dag.jac + dag.hes)
    body = list(pr.lines)
    body.append(f"m = {pr.ref(dag.model)};")
    body += [f"J[{j}] = {pr.ref(dag.jac[j])};" for j in range(n)]
    body += [f"d2[{l}] = {pr.ref(dag.hes[l])};" for l in range(nh)]
    out.append("  __device__ static __forceinline__ void fjh(const real* __restrict__ p, "
               "const real* __restrict__ c, real& m, real* J, real* d2) {")
    out += ["    " + s for s in body]
    out.append("  }")
    if factored:
        lines, entries = _emit_factored(dag, dtype, derive=derived is not None,
                                        exp_rec=exp_rec)
        out += lines
        if derived is not None:
            derived.extend(entries)
    else:
        out += ["  static constexpr bool FACTORED = false;",
                f"  static constexpr int NCT = {dag.nconst};"]
    out.append("};")
    return "\n".join(out) + "\n"


def flops_per_iteration(dag: ModelDAG, M: int) -> dict:
    """Algorithmic FLOPs of one active LM iteration (exp counted as 1 FLOP and
    also reported separately).

    F_it = M (E_fjh + 1 + 2 + 2n + 2 nh + 2 nnz + 2n) + M (E_f + 3)
           + F_solve(n) + F_step(n)
    with E_fjh/E_f the op counts of the CSE'd DAGs (transcendentals included
    as one each), nnz the structurally non-zero Hessian entries, and
      F_solve(n) = (n sqrt + n div + 2 nh) + sum_j [j + 2j(n-1-j) + 2 + 3(n-1-j)] + 4
                   + 3 n(n-1)/2 + n + n(n-1) + n
      F_step(n)  = (4n + 2) + n + (3 nh + 2n + 4).
    The converging iteration stops after the convergence test:
      F_conv = M (E_fjh + 1 + 2 + 2n + 2 nh + 2 nnz + 2n) + F_solve(n) + 4n + 2.
    """
    n, nh = dag.nparam, dag.nhes
    nnz = sum(dag.hes_nonzero())

    def tot(c):
        return c["flops"] + c["exp"] + c["log"] + c["pow"] + c["sqrt"] + c["other_fn"]

    e_fjh, e_f = tot(dag.counts_fjh), tot(dag.counts_f)
    gmw = sum(j + 2 * j * (n - 1 - j) + 2 + 3 * (n - 1 - j) for j in range(n)) + 4
    f_solve = (2 * n + 2 * nh) + gmw + 3 * n * (n - 1) // 2 + n + n * (n - 1) + n
    f_step = (4 * n + 2) + n + (3 * nh + 2 * n + 4)
    fghhl = M * (e_fjh + 1 + 2 + 2 * n + 2 * nh + 2 * nnz + 2 * n)
    trial = M * (e_f + 3)
    return {
        "F_it": fghhl + trial + f_solve + f_step,
        "F_conv": fghhl + f_solve + 4 * n + 2,
        "exp_it": M * (dag.counts_fjh["exp"] + dag.counts_f["exp"]),
        "exp_conv": M * dag.counts_fjh["exp"],
        "E_fjh": e_fjh, "E_f": e_f, "nnz": nnz, "F_solve": f_solve, "F_step": f_step,
    }
