"""Expression DAG, symbolic differentiation and op counting.

This replaces the reference's symengine use (pycompute/sym/util.py:64-112:
``sympify`` -> ``diff`` -> ``cse``) with a small standard-library engine, so the
code generator runs anywhere Python runs:

* ``parse`` reads the same infix strings the reference passes to ``sympify``
  (``**`` and ``^`` are both powers, as in sympify's default ``convert_xor``).
* Nodes are hash-consed in a ``Graph``: structurally equal subexpressions are
  the same object, so the DAG *is* the common-subexpression elimination.
* ``diff`` differentiates with local simplification (constant folding, x+0,
  x*1, x*0, --x), which is what symengine's automatic canonicalisation does for
  the forms that occur in these models.
"""
from __future__ import annotations

import ast
import math
from dataclasses import dataclass, field

_FUNCS = ("exp", "log", "sqrt", "sin", "cos", "tan", "sinh", "cosh", "tanh",
          "asin", "acos", "atan", "abs")


@dataclass(eq=False)
class Node:
    op: str                      # num, sym, add, mul, neg, div, pow, fn
    args: tuple = ()
    value: float = 0.0           # for num
    name: str = ""               # for sym / fn
    id: int = field(default=0)

    def __repr__(self):
        return to_str(self)


class Graph:
    """Hash-consing node factory with simplifying constructors."""

    def __init__(self):
        self._table: dict = {}
        self._count = 0

    def _make(self, op, args=(), value=0.0, name=""):
        key = (op, tuple(a.id for a in args), value if op == "num" else 0.0, name)
        n = self._table.get(key)
        if n is None:
            n = Node(op, tuple(args), value, name, self._count)
            self._count += 1
            self._table[key] = n
        return n

    # leaves
    def num(self, v):
        v = float(v)
        if v == 0.0:
            v = 0.0  # fold -0.0
        return self._make("num", value=v)

    def sym(self, name):
        return self._make("sym", name=name)

    # arithmetic with local simplification
    def add(self, a, b):
        if a.op == "num" and b.op == "num":
            return self.num(a.value + b.value)
        if is_zero(a):
            return b
        if is_zero(b):
            return a
        if b.op == "neg" and b.args[0] is a:
            return self.num(0)
        if a.op == "neg" and a.args[0] is b:
            return self.num(0)
        return self._make("add", (a, b))

    def sub(self, a, b):
        if b.op == "num":
            return self.add(a, self.num(-b.value))
        if b.op == "neg":
            return self.add(a, b.args[0])
        if a is b:
            return self.num(0)
        return self.add(a, self.neg(b))

    def neg(self, a):
        if a.op == "num":
            return self.num(-a.value)
        if a.op == "neg":
            return a.args[0]
        return self._make("neg", (a,))

    def mul(self, a, b):
        if a.op == "num" and b.op == "num":
            return self.num(a.value * b.value)
        if is_zero(a) or is_zero(b):
            return self.num(0)
        if is_one(a):
            return b
        if is_one(b):
            return a
        if a.op == "num" and a.value == -1.0:
            return self.neg(b)
        if b.op == "num" and b.value == -1.0:
            return self.neg(a)
        if a.op == "neg" and b.op == "neg":
            return self.mul(a.args[0], b.args[0])
        if a.op == "neg":
            return self.neg(self.mul(a.args[0], b))
        if b.op == "neg":
            return self.neg(self.mul(a, b.args[0]))
        return self._make("mul", (a, b))

    def div(self, a, b):
        if is_zero(a):
            return self.num(0)
        if is_one(b):
            return a
        if a.op == "num" and b.op == "num" and b.value != 0.0:
            return self.num(a.value / b.value)
        if a.op == "neg":
            return self.neg(self.div(a.args[0], b))
        if b.op == "neg":
            return self.neg(self.div(a, b.args[0]))
        return self._make("div", (a, b))

    def pow(self, a, b):
        if b.op == "num":
            if b.value == 0.0:
                return self.num(1)
            if b.value == 1.0:
                return a
            if a.op == "num":
                return self.num(a.value ** b.value)
        if is_one(a):
            return self.num(1)
        return self._make("pow", (a, b))

    def fn(self, name, a):
        if name not in _FUNCS:
            raise ValueError(f"unsupported function {name!r}")
        if a.op == "num":
            return self.num(getattr(math, "fabs" if name == "abs" else name)(a.value))
        return self._make("fn", (a,), name=name)


def is_zero(n):
    return n.op == "num" and n.value == 0.0


def is_one(n):
    return n.op == "num" and n.value == 1.0


# ---------------------------------------------------------------------------
# parsing
# ---------------------------------------------------------------------------

def parse(g: Graph, text: str, names: dict | None = None) -> Node:
    """Parse an infix expression string into ``g``.

    ``names`` optionally maps identifiers in the text to replacement symbol
    names (the reference renames parameters to ``parvar_*`` / consts to
    ``convar_*`` before differentiation, lsqnonlin.py:264-273).
    """
    tree = ast.parse(text.replace("^", "**"), mode="eval")
    names = names or {}

    def walk(t):
        if isinstance(t, ast.Expression):
            return walk(t.body)
        if isinstance(t, ast.Constant) and isinstance(t.value, (int, float)):
            return g.num(t.value)
        if isinstance(t, ast.Name):
            if t.id in ("pi", "E") and t.id not in names:
                return g.num(math.pi if t.id == "pi" else math.e)
            return g.sym(names.get(t.id, t.id))
        if isinstance(t, ast.UnaryOp):
            a = walk(t.operand)
            if isinstance(t.op, ast.USub):
                return g.neg(a)
            if isinstance(t.op, ast.UAdd):
                return a
        if isinstance(t, ast.BinOp):
            a, b = walk(t.left), walk(t.right)
            if isinstance(t.op, ast.Add):
                return g.add(a, b)
            if isinstance(t.op, ast.Sub):
                return g.sub(a, b)
            if isinstance(t.op, ast.Mult):
                return g.mul(a, b)
            if isinstance(t.op, ast.Div):
                return g.div(a, b)
            if isinstance(t.op, ast.Pow):
                return g.pow(a, b)
        if isinstance(t, ast.Call) and isinstance(t.func, ast.Name) and len(t.args) == 1:
            fname = {"Abs": "abs", "fabs": "abs"}.get(t.func.id, t.func.id)
            return g.fn(fname, walk(t.args[0]))
        raise ValueError(f"unsupported syntax in model expression: {ast.dump(t)}")

    return walk(tree)


def free_symbols(n: Node) -> set:
    out, seen, stack = set(), set(), [n]
    while stack:
        x = stack.pop()
        if x.id in seen:
            continue
        seen.add(x.id)
        if x.op == "sym":
            out.add(x.name)
        stack.extend(x.args)
    return out


# ---------------------------------------------------------------------------
# differentiation
# ---------------------------------------------------------------------------

def diff(g: Graph, n: Node, var: str, memo: dict | None = None) -> Node:
    """d n / d var, memoised per (node, var)."""
    if memo is None:
        memo = {}
    key = (n.id, var)
    r = memo.get(key)
    if r is not None:
        return r
    op = n.op
    if op == "num":
        r = g.num(0)
    elif op == "sym":
        r = g.num(1) if n.name == var else g.num(0)
    elif op == "add":
        r = g.add(diff(g, n.args[0], var, memo), diff(g, n.args[1], var, memo))
    elif op == "neg":
        r = g.neg(diff(g, n.args[0], var, memo))
    elif op == "mul":
        a, b = n.args
        r = g.add(g.mul(diff(g, a, var, memo), b), g.mul(a, diff(g, b, var, memo)))
    elif op == "div":
        a, b = n.args
        da, db = diff(g, a, var, memo), diff(g, b, var, memo)
        # d(a/b) = da/b - a*db/b^2
        r = g.sub(g.div(da, b), g.div(g.mul(a, db), g.pow(b, g.num(2))))
    elif op == "pow":
        a, b = n.args
        da, db = diff(g, a, var, memo), diff(g, b, var, memo)
        if b.op == "num":
            r = g.mul(g.mul(b, g.pow(a, g.num(b.value - 1.0))), da)
        else:
            # a^b (db log a + b da / a)
            t1 = g.mul(db, g.fn("log", a))
            t2 = g.div(g.mul(b, da), a)
            r = g.mul(n, g.add(t1, t2))
    elif op == "fn":
        a = n.args[0]
        da = diff(g, a, var, memo)
        if is_zero(da):
            r = g.num(0)
        else:
            nm = n.name
            if nm == "exp":
                d = n
            elif nm == "log":
                d = g.div(g.num(1), a)
            elif nm == "sqrt":
                d = g.div(g.num(0.5), n)
            elif nm == "sin":
                d = g.fn("cos", a)
            elif nm == "cos":
                d = g.neg(g.fn("sin", a))
            elif nm == "tan":
                d = g.add(g.num(1), g.pow(n, g.num(2)))
            elif nm == "sinh":
                d = g.fn("cosh", a)
            elif nm == "cosh":
                d = g.fn("sinh", a)
            elif nm == "tanh":
                d = g.sub(g.num(1), g.pow(n, g.num(2)))
            elif nm == "asin":
                d = g.div(g.num(1), g.fn("sqrt", g.sub(g.num(1), g.pow(a, g.num(2)))))
            elif nm == "acos":
                d = g.neg(g.div(g.num(1), g.fn("sqrt", g.sub(g.num(1), g.pow(a, g.num(2))))))
            elif nm == "atan":
                d = g.div(g.num(1), g.add(g.num(1), g.pow(a, g.num(2))))
            elif nm == "abs":
                d = g.div(a, n)
            else:  # pragma: no cover
                raise ValueError(nm)
            r = g.mul(d, da)
    else:  # pragma: no cover
        raise ValueError(op)
    memo[key] = r
    return r


# ---------------------------------------------------------------------------
# evaluation / printing / counting
# ---------------------------------------------------------------------------

def evaluate(n: Node, env: dict, memo: dict | None = None):
    """Evaluate with numpy-compatible values in ``env`` (name -> value)."""
    import numpy as np
    if memo is None:
        memo = {}
    if n.id in memo:
        return memo[n.id]
    op = n.op
    if op == "num":
        r = n.value
    elif op == "sym":
        r = env[n.name]
    else:
        a = [evaluate(x, env, memo) for x in n.args]
        if op == "add":
            r = a[0] + a[1]
        elif op == "neg":
            r = -a[0]
        elif op == "mul":
            r = a[0] * a[1]
        elif op == "div":
            r = a[0] / a[1]
        elif op == "pow":
            r = np.power(a[0], a[1])
        elif op == "fn":
            r = getattr(np, {"abs": "abs", "asin": "arcsin", "acos": "arccos",
                             "atan": "arctan"}.get(n.name, n.name))(a[0])
        else:  # pragma: no cover
            raise ValueError(op)
    memo[n.id] = r
    return r


def to_str(n: Node) -> str:
    op = n.op
    if op == "num":
        return repr(n.value)
    if op == "sym":
        return n.name
    a = [to_str(x) for x in n.args]
    if op == "add":
        return f"({a[0]} + {a[1]})"
    if op == "neg":
        return f"(-{a[0]})"
    if op == "mul":
        return f"({a[0]}*{a[1]})"
    if op == "div":
        return f"({a[0]}/{a[1]})"
    if op == "pow":
        return f"({a[0]}**{a[1]})"
    return f"{n.name}({a[0]})"


def topo(outputs) -> list:
    """Nodes reachable from ``outputs`` in dependency order (each once)."""
    order, seen = [], set()

    def visit(x):
        stack = [(x, False)]
        while stack:
            node, done = stack.pop()
            if done:
                order.append(node)
                continue
            if node.id in seen:
                continue
            seen.add(node.id)
            stack.append((node, True))
            for c in reversed(node.args):
                if c.id not in seen:
                    stack.append((c, False))

    for o in outputs:
        visit(o)
    return order


def int_power(n: Node):
    """Integer exponent of a pow node with a small integral constant exponent."""
    if n.op == "pow" and n.args[1].op == "num":
        e = n.args[1].value
        if e == int(e) and 2 <= abs(e) <= 8:
            return int(e)
    return None


def op_counts(outputs) -> dict:
    """Arithmetic cost of the CSE'd DAG: add/sub/mul/div = 1, neg = 0,
    integer pow = multiplications by squaring, transcendental calls counted
    separately under their names."""
    c = {"flops": 0, "exp": 0, "log": 0, "pow": 0, "sqrt": 0, "other_fn": 0}
    for x in topo(outputs):
        if x.op in ("add", "mul", "div"):
            c["flops"] += 1
        elif x.op == "pow":
            e = int_power(x)
            if e is not None:
                k = abs(e)
                c["flops"] += (k.bit_length() - 1) + (bin(k).count("1") - 1) + (1 if e < 0 else 0)
            elif x.args[1].op == "num" and x.args[1].value in (0.5, -0.5):
                c["sqrt"] += 1
                c["flops"] += 1 if x.args[1].value < 0 else 0
            else:
                c["pow"] += 1
        elif x.op == "fn":
            if x.name in c:
                c[x.name] += 1
            else:
                c["other_fn"] += 1
    return c
