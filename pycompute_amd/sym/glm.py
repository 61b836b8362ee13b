"""Structure detection for the wide-model kernel: generalised linear models.

A model is a GLM here when f = phi(eta) with phi in {exp, identity} and eta
affine in the parameters: eta = a0(c) + sum_k A_k(c) p_k, the coefficients
depending on the consts only. Then for measurement i

    J_k = phi'(eta) A_k,   d2f/dp_j dp_k = phi''(eta) A_j A_k,

so the reference's per-entry Hessian update h_jk += d2f_jk r + J_j J_k
(pycompute/cuda/lsqnonlin.py:206-346) becomes h_jk += w A_j A_k with
w = phi'^2 + phi'' r: one FMA per packed entry once w A_j is formed. The
log-linear diffusion models (DTI, DKI) have this form. The coefficient table
is consts-only, so it is evaluated once per protocol on the host (float64,
rounded to the kernel's type) and read by the kernel with scalar loads.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import expr as E


@dataclass
class GLMForm:
    phi: str              # "exp" or "identity"
    eta: E.Node
    coef: list            # d eta / d p_k (consts-only nodes)


def glm_form(dag) -> GLMForm | None:
    """The GLM structure of a ModelDAG, or None."""
    m = dag.model
    if m.op == "fn" and m.name == "exp":
        phi, eta = "exp", m.args[0]
    else:
        phi, eta = "identity", m
    pars = set(dag.pars)
    if not (E.free_symbols(eta) & pars):
        return None
    coef = [E.diff(dag.graph, eta, p) for p in dag.pars]
    for c in coef:
        if E.free_symbols(c) & pars:
            return None
    return GLMForm(phi, eta, coef)


def table_stride(nparam: int) -> int:
    """Row length of the per-measurement table: A (n), A^2 (n), offset (1),
    padded to a multiple of 8 floats for wide scalar loads."""
    return (2 * nparam + 1 + 7) // 8 * 8


def coefficient_table(form: GLMForm, dag, consts: np.ndarray) -> np.ndarray:
    """(M, stride) float64 table for voxel-uniform consts (nconst, M): per
    measurement A_0..A_{n-1}, A_0^2..A_{n-1}^2, a0 = eta(p = 0), zero padding."""
    consts = np.asarray(consts, np.float64).reshape(len(dag.consts), -1)
    M = consts.shape[1]
    n = dag.nparam
    env = {c: consts[k] for k, c in enumerate(dag.consts)}
    env.update({p: np.zeros(M) for p in dag.pars})
    tab = np.zeros((M, table_stride(n)))
    memo: dict = {}
    for k, c in enumerate(form.coef):
        tab[:, k] = np.broadcast_to(E.evaluate(c, env, memo), (M,))
    tab[:, n:2 * n] = tab[:, :n] ** 2
    tab[:, 2 * n] = np.broadcast_to(E.evaluate(form.eta, env, memo), (M,))
    # affine check: eta at random parameters equals the table's affine form
    rng = np.random.default_rng(0)
    p = rng.uniform(-1.0, 1.0, n)
    env2 = dict(env)
    env2.update({name: np.full(M, p[k]) for k, name in enumerate(dag.pars)})
    direct = np.broadcast_to(E.evaluate(form.eta, env2, {}), (M,))
    affine = tab[:, 2 * n] + tab[:, :n] @ p
    if not np.allclose(direct, affine, rtol=1e-9, atol=1e-9 * (1 + np.abs(direct).max())):
        raise ValueError("model is not affine in its parameters inside phi")
    return tab
