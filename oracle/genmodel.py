"""Host build of a generated model, for the strict build's bit-exact replay.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). The project's code
generator (pycompute_amd.sym.hipgen) emits a model struct for the GPU; this
compiles the very same struct for the host with g++ -ffp-contract=off and the
reproducible transcendentals of pycompute_amd/csrc/strict_math.h, so that the
oracle's LM driver (lm_oracle.cpp, oracle_lm_run_gen_*) can replay the strict
GPU build bit for bit. What this replay checks is the LM driver, the GMW81
solve, the gain ratio and the clamp/convergence logic; the generated model
code itself is checked separately against the oracle's hand-written models.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

from . import _c, _consts_arg, _ptr, _suf, lib

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEN = os.path.join(_HERE, "_gen")
_ROOT = os.path.dirname(_HERE)
_STRICT_H = os.path.join(_ROOT, "pycompute_amd", "csrc", "strict_math.h")

_SHIM = """
#include <cmath>
#include "{strict_h}"
#define __device__
#define __forceinline__ inline
typedef {real} real;
static inline float pca_fmax(float a, float b) {{ return fmaxf(a, b); }}
static inline double pca_fmax(double a, double b) {{ return fmax(a, b); }}
#define PCA_EXP(x) pca_exp_strict(x)
#define PCA_LOG(x) pca_log_strict(x)
#define PCA_POW(x, y) pca_pow_strict(x, y)
{model}
extern "C" void gen_accum(const real* p, const real* c, real y, real lam, real* f, real* g,
                          real* h, real* hld) {{
  GenModel::accum(p, c, y, lam, *f, g, h, hld);
}}
extern "C" real gen_resid(const real* p, const real* c, real y) {{
  return GenModel::resid(p, c, y);
}}
"""

_libs: dict = {}


def build(expr: str, pars, consts, dtype=np.float32):
    """Compile the generated model for the host; returns the ctypes library."""
    from pycompute_amd.sym import hipgen
    dn = "float32" if np.dtype(dtype) == np.float32 else "float64"
    dag = hipgen.res_jac_hes(expr, pars, consts)
    src = _SHIM.format(strict_h=_STRICT_H, real="float" if dn == "float32" else "double",
                       model=hipgen.emit_model(dag, "GenModel", dn))
    with open(_STRICT_H) as fh:
        key = hashlib.sha256((src + fh.read()).encode()).hexdigest()[:24]
    if key in _libs:
        return _libs[key]
    os.makedirs(_GEN, exist_ok=True)
    so = os.path.join(_GEN, f"gen_{key}.so")
    if not os.path.exists(so):
        cpp = os.path.join(_GEN, f"gen_{key}.cpp")
        with open(cpp, "w") as fh:
            fh.write(src)
        subprocess.run(["g++", "-O2", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off",
                        "-fPIC", "-shared", "-o", so + ".tmp", cpp], check=True)
        os.replace(so + ".tmp", so)
    L = ctypes.CDLL(so)
    L.n = len(pars)
    L.nc = len(consts)
    _libs[key] = L
    return L


def lm_run(expr, pars, consts_names, consts, data, lb, ub, params, iters, *, dtype=np.float32,
           tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0, nthreads=0):
    """The oracle's LM driver over the generated model (strict transcendentals)."""
    gl = build(expr, pars, consts_names, dtype)
    suf = _suf(dtype)
    data = _c(data, dtype)
    M, N = data.shape
    n, nc = len(pars), len(consts_names)
    cst, cpv = _consts_arg(consts, nc, M, N, dtype)
    out = {"params": _c(params, dtype).copy(), "lam": np.full(N, 5.0, dtype),
           "step_type": np.ones(N, np.int8), "f": np.zeros(N, dtype),
           "first_f": np.zeros(N, dtype), "last_f": np.zeros(N, dtype),
           "n_iter": np.zeros(N, np.int32)}
    lbc, ubc = _c(lb, dtype), _c(ub, dtype)
    L = lib()
    ct = ctypes.c_float if suf == "f" else ctypes.c_double
    P = ctypes.c_void_p
    fn = getattr(L, "oracle_lm_run_gen_" + suf)
    fn.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, P, ctypes.c_int,
                   P, P, P, P, P, P, P, P, P, P, ctypes.c_int, ct, ct, ct, ct, ct, ctypes.c_int]
    fn.restype = ctypes.c_int
    acc_p = ctypes.cast(gl.gen_accum, P)
    res_p = ctypes.cast(gl.gen_resid, P)
    rc = fn(acc_p, res_p, n, nc, M, N, _ptr(cst), cpv, _ptr(data), _ptr(lbc), _ptr(ubc),
            _ptr(out["params"]), _ptr(out["lam"]), _ptr(out["step_type"]), _ptr(out["f"]),
            _ptr(out["first_f"]), _ptr(out["last_f"]), _ptr(out["n_iter"]), iters,
            tol, mu, eta, acc, dec, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_lm_run_gen failed: {rc}")
    return out


def ssr(expr, pars, consts_names, consts, data, params, *, dtype=np.float32, nthreads=0):
    gl = build(expr, pars, consts_names, dtype)
    suf = _suf(dtype)
    data = _c(data, dtype)
    M, N = data.shape
    cst, cpv = _consts_arg(consts, len(consts_names), M, N, dtype)
    p = _c(params, dtype)
    out = np.zeros(N, dtype)
    P = ctypes.c_void_p
    fn = getattr(lib(), "oracle_ssr_gen_" + suf)
    fn.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, P, ctypes.c_int,
                   P, P, P, ctypes.c_int]
    fn.restype = ctypes.c_int
    rc = fn(ctypes.cast(gl.gen_resid, P), len(pars), len(consts_names), M, N, _ptr(cst), cpv,
            _ptr(data), _ptr(p), _ptr(out), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_ssr_gen failed: {rc}")
    return out
