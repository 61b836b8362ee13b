"""CPU oracle for the batched per-voxel LM fitter -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package; it is the checker, never the thing measured or shipped.
The product (``pycompute_amd``) never imports it and fails loudly when its HIP
library is missing.

Contents:
  * ctypes bindings of ``liboracle.so`` (``lm_oracle.cpp``): the operation-by-
    operation restatement of the reference LM iteration, FP32 and FP64.
  * numpy restatements of the reference's host-side preprocessing:
    ``bpolyfit`` (pycompute/cuda/lsq.py:22-30) and the IVIM default initial
    guess (test/ivim/test_ivim.ipynb cell 6, lines 146-165).

Pinning (see tests/test_oracle.py):
  * GMW81 solve: mtests/test_mats.m:49-56 known answer; numpy.linalg.solve on
    well-conditioned SPD matrices (GMW81 adds nothing there); GMW81 invariants
    on indefinite matrices (L D L^T = P B P^T + E with E diagonal >= 0, D > 0,
    |l_ij| sqrt(d_j) <= beta); pivot = rotation [n-1, 0, .., n-2].
  * models: closed forms checked against sympy derivatives and finite
    differences; IVIM against the reference's emitted CSE.
  * LM iteration: invariants (SSR non-increasing over accepted steps, lambda
    changes only by acc/dec, step_type codes, bounds), noise-free start at the
    truth converges in one iteration, FP32 vs FP64 agreement.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MODEL_IDS = {"ivim": 0, "t2": 1, "se": 2, "dki": 3}
MODEL_NP = {"ivim": 4, "t2": 2, "se": 3, "dki": 22}
MODEL_NC = {"ivim": 1, "t2": 1, "se": 1, "dki": 4}


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for suf, ct in (("f", ctypes.c_float), ("d", ctypes.c_double)):
            P = ctypes.c_void_p
            fn = getattr(L, "oracle_lm_run_" + suf)
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                           P, ctypes.c_int, P, P, P, P, P, P, P, P, P, P, ctypes.c_int,
                           ct, ct, ct, ct, ct, P, P, P, ctypes.c_int]
            fn.restype = ctypes.c_int
            fn = getattr(L, "oracle_ssr_" + suf)
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                           P, ctypes.c_int, P, P, P, ctypes.c_int]
            fn.restype = ctypes.c_int
            fn = getattr(L, "oracle_model_eval_" + suf)
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, P, ctypes.c_int, P, P, P, P]
            fn.restype = ctypes.c_int
            for name, args in (("oracle_gmw81_", [ctypes.c_int, P]),
                               ("oracle_diag_pivot_", [ctypes.c_int, P, P]),
                               ("oracle_gmw81_solve_", [ctypes.c_int, P, P, P])):
                fn = getattr(L, name + suf)
                fn.argtypes = args
                fn.restype = None
        _lib = L
    return _lib


def _suf(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f"
    if dtype == np.float64:
        return "d"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _consts_arg(consts, nc, M, N, dtype):
    """Uniform consts ((nc, M), (nc*M,), (nc*M, 1)) -> per_voxel=0;
    (nc*M, N) or (1, nc*M, N) -> per_voxel=1."""
    c = np.asarray(consts)
    if c.ndim == 3:  # reference (1, M, N) layout
        c = c.reshape(-1, c.shape[-1])
    if c.size == nc * M:
        return _c(c.reshape(-1), dtype), 0
    if c.shape == (nc * M, N):
        return _c(c, dtype), 1
    raise ValueError(f"consts of shape {np.shape(consts)} fit neither ({nc * M},) nor "
                     f"({nc * M}, {N})")


def lm_run(model, consts, data, lb, ub, params, iters, *, dtype=np.float32, lam=None,
           step_type=None, f=None, tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0,
           trace=False, nthreads=0):
    """Run ``iters`` reference launches of the LM iteration on every voxel.

    Returns a dict with params, lam, step_type, f (the reference's f_t),
    first_f, last_f, n_iter (iterations in which the voxel was active) and,
    when ``trace``, per-launch f/lam/step_type.
    """
    suf = _suf(dtype)
    data = _c(data, dtype)
    M, N = data.shape
    n = MODEL_NP[model]
    nc = MODEL_NC[model]
    cst, cpv = _consts_arg(consts, nc, M, N, dtype)
    out = {
        "params": _c(params, dtype).copy(),
        "lam": (np.full(N, 5.0, dtype) if lam is None else _c(lam, dtype).reshape(N).copy()),
        "step_type": (np.ones(N, np.int8) if step_type is None
                      else _c(step_type, np.int8).reshape(N).copy()),
        "f": (np.zeros(N, dtype) if f is None else _c(f, dtype).reshape(N).copy()),
        "first_f": np.zeros(N, dtype),
        "last_f": np.zeros(N, dtype),
        "n_iter": np.zeros(N, np.int32),
    }
    lbc, ubc = _c(lb, dtype), _c(ub, dtype)
    assert out["params"].shape == (n, N) and lbc.shape == (n, N) and ubc.shape == (n, N)
    tr = {}
    if trace:
        tr = {"trace_f": np.zeros((iters, N), dtype), "trace_lam": np.zeros((iters, N), dtype),
              "trace_st": np.zeros((iters, N), np.int8)}
    rc = getattr(lib(), "oracle_lm_run_" + suf)(
        MODEL_IDS[model], n, nc, M, N, _ptr(cst), cpv, _ptr(data), _ptr(lbc), _ptr(ubc),
        _ptr(out["params"]), _ptr(out["lam"]), _ptr(out["step_type"]), _ptr(out["f"]),
        _ptr(out["first_f"]), _ptr(out["last_f"]), _ptr(out["n_iter"]), iters,
        tol, mu, eta, acc, dec,
        _ptr(tr.get("trace_f")), _ptr(tr.get("trace_lam")), _ptr(tr.get("trace_st")), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_lm_run failed: {rc}")
    out.update(tr)
    return out


def ssr(model, consts, data, params, *, dtype=np.float32, nthreads=0):
    """Sum of squared residuals at ``params`` (the reference's F kernel)."""
    suf = _suf(dtype)
    data = _c(data, dtype)
    M, N = data.shape
    cst, cpv = _consts_arg(consts, MODEL_NC[model], M, N, dtype)
    p = _c(params, dtype)
    out = np.zeros(N, dtype)
    rc = getattr(lib(), "oracle_ssr_" + suf)(MODEL_IDS[model], MODEL_NP[model], MODEL_NC[model],
                                             M, N, _ptr(cst), cpv, _ptr(data), _ptr(p), _ptr(out),
                                             nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_ssr failed: {rc}")
    return out


def model_eval(model, consts, params, M, *, dtype=np.float64):
    """Model value (M,N), Jacobian (n,M,N) and packed second derivatives (nh,M,N)."""
    suf = _suf(dtype)
    p = _c(params, dtype)
    n, N = p.shape
    nh = n * (n + 1) // 2
    cst, cpv = _consts_arg(consts, MODEL_NC[model], M, N, dtype)
    m = np.zeros((M, N), dtype)
    J = np.zeros((n, M, N), dtype)
    d2 = np.zeros((nh, M, N), dtype)
    getattr(lib(), "oracle_model_eval_" + suf)(MODEL_IDS[model], M, N, _ptr(cst), cpv, _ptr(p),
                                               _ptr(m), _ptr(J), _ptr(d2))
    return m, J, d2


def gmw81(mat, dtype=np.float64):
    """In-place GMW81 of solver.py:20-115 on a copy of a full symmetric matrix."""
    a = _c(mat, dtype).copy()
    getattr(lib(), "oracle_gmw81_" + _suf(dtype))(a.shape[0], _ptr(a))
    return a


def diag_pivot(mat, dtype=np.float64):
    a = _c(mat, dtype).copy()
    perm = np.zeros(a.shape[0], np.int32)
    getattr(lib(), "oracle_diag_pivot_" + _suf(dtype))(a.shape[0], _ptr(a), _ptr(perm))
    return a, perm


def gmw81_solve(mat, rhs, dtype=np.float64):
    """Full reference solve (pivot, scale, GMW81, LDL^T solve, unscale, unpivot)."""
    a = _c(mat, dtype).copy()
    b = _c(rhs, dtype)
    x = np.zeros_like(b)
    getattr(lib(), "oracle_gmw81_solve_" + _suf(dtype))(a.shape[0], _ptr(a), _ptr(b), _ptr(x))
    return x


# ---------------------------------------------------------------------------
# numpy restatements of the caller-side preprocessing
# ---------------------------------------------------------------------------

def bpolyfit(x, y, deg, dtype=np.float64):
    """pycompute/cuda/lsq.py:22-30: batched least squares of y on [1, x, .., x^deg].

    x: (1, m, N) or (m, N) or (m,); y: (m, N). Returns (deg+1, N). The
    reference uses pinv(A) @ y, which equals the least-squares solution when
    A has full column rank; this restatement uses lstsq per voxel.
    """
    y = np.asarray(y, np.float64)
    m, N = y.shape
    x = np.asarray(x, np.float64)
    if x.ndim == 1:
        x = np.repeat(x[:, None], N, axis=1)
    elif x.ndim == 3:
        x = x.reshape(x.shape[1], x.shape[2])
    if x.shape[1] == 1:
        x = np.repeat(x, N, axis=1)
    out = np.zeros((deg + 1, N))
    for v in range(N):
        A = np.stack([x[:, v] ** i for i in range(deg + 1)], axis=1)
        out[:, v] = np.linalg.lstsq(A, y[:, v], rcond=None)[0]
    return out.astype(dtype)


def ivim_default_init(b, data, lb, ub, lo_idx, hi_idx):
    """test_ivim.ipynb cell 6 (:146-165): segmented log-linear fits.

    lo_idx/hi_idx select the low-b and high-b measurements (the reference
    uses [0:10) and [15:) of its 21-point protocol).
    """
    b = np.asarray(b, np.float64)
    y = np.log(np.asarray(data, np.float64))
    lo = bpolyfit(-b[lo_idx], y[lo_idx], 1)
    hi = bpolyfit(-b[hi_idx], y[hi_idx], 1)
    lo[0] = np.exp(lo[0])
    hi[0] = np.exp(hi[0])
    with np.errstate(all="ignore"):
        p = np.stack([lo[0], 1.0 - hi[0] / lo[0], lo[1], hi[1]])
    p = np.nan_to_num(p, nan=0.0, posinf=0.0, neginf=0.0)
    return np.clip(p, lb, ub)
