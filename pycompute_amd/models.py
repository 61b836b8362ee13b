"""Built-in model specifications for the benchmark workloads.

IVIM is the reference's own model, parameter order and absolute bounds
(test/ivim/test_ivim.ipynb:128-137, 359-361). Mono-exponential T2, the
stretched exponential and the 22-parameter diffusion-kurtosis (DKI)
exp-linear model use the same expression syntax with physical bounds chosen
here (the reference has no such models).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class ModelSpec:
    name: str
    expr: str
    pars: tuple
    consts: tuple
    lb: tuple
    ub: tuple

    @property
    def nparam(self):
        return len(self.pars)

    @property
    def nconst(self):
        return len(self.consts)


IVIM = ModelSpec(
    "ivim", "S0*(f*exp(-b*D_1)+(1-f)*exp(-b*D_2))", ("S0", "f", "D_1", "D_2"), ("b",),
    (0.0, 1e-3, 1e-6, 1e-8), (2e4, 0.999, 0.1, 0.02))

T2 = ModelSpec("t2", "S0*exp(-TE/T2)", ("S0", "T2"), ("TE",), (0.0, 1.0), (2e4, 5000.0))

SE = ModelSpec("se", "S0*exp(-(b*D)**alpha)", ("S0", "D", "alpha"), ("b",),
               (0.0, 1e-6, 0.1), (2e4, 0.1, 1.0))

# DKI: log S = lnS0 - b sum_k a_k(g) D_k + b^2/6 sum_m w_m(g) K_m, with
# D = (Dxx, Dyy, Dzz, Dxy, Dxz, Dyz) and K_m = MD^2 W_m over the 15 unique
# elements of the symmetric 4th-order kurtosis tensor (multiplicities 1,4,6,12).
DKI_D = ("Dxx", "Dyy", "Dzz", "Dxy", "Dxz", "Dyz")
DKI_K = ("Kxxxx", "Kyyyy", "Kzzzz", "Kxxxy", "Kxxxz", "Kxyyy", "Kyyyz", "Kxzzz", "Kyzzz",
         "Kxxyy", "Kxxzz", "Kyyzz", "Kxxyz", "Kxyyz", "Kxyzz")
_DKI_A = ("gx*gx", "gy*gy", "gz*gz", "2*gx*gy", "2*gx*gz", "2*gy*gz")
_DKI_W = ("gx**4", "gy**4", "gz**4", "4*gx**3*gy", "4*gx**3*gz", "4*gx*gy**3", "4*gy**3*gz",
          "4*gx*gz**3", "4*gy*gz**3", "6*gx**2*gy**2", "6*gx**2*gz**2", "6*gy**2*gz**2",
          "12*gx**2*gy*gz", "12*gx*gy**2*gz", "12*gx*gy*gz**2")
DKI_EXPR = ("exp(lnS0 - b*(" + " + ".join(f"{a}*{d}" for a, d in zip(_DKI_A, DKI_D)) + ")"
            + " + b*b/6*(" + " + ".join(f"{w}*{k}" for w, k in zip(_DKI_W, DKI_K)) + "))")
# Bounds: physical ranges with margin. Diffusivities up to 5e-3 mm^2/s; the
# K terms MD^2 W stay below (3e-3)^2 * ~2.2 = 2e-5 for kurtosis up to 2, so
# +-2e-4 keeps exp(eta) finite at b = 2000 (b^2/6 * 2e-4 = 133 per unit w)
# while leaving a 10x margin.
DKI = ModelSpec(
    "dki", DKI_EXPR, ("lnS0",) + DKI_D + DKI_K, ("b", "gx", "gy", "gz"),
    (-10.0,) + (0.0, 0.0, 0.0, -5e-3, -5e-3, -5e-3) + (-2e-4,) * 15,
    (10.0,) + (5e-3, 5e-3, 5e-3, 5e-3, 5e-3, 5e-3) + (2e-4,) * 15)

MODELS = {m.name: m for m in (IVIM, T2, SE, DKI)}


def dki_design(consts: np.ndarray) -> np.ndarray:
    """(M, 22) design matrix of log S for consts (4, M) = (b, gx, gy, gz)."""
    b, gx, gy, gz = consts
    a = np.stack([gx * gx, gy * gy, gz * gz, 2 * gx * gy, 2 * gx * gz, 2 * gy * gz])
    w = np.stack([gx ** 4, gy ** 4, gz ** 4, 4 * gx ** 3 * gy, 4 * gx ** 3 * gz, 4 * gx * gy ** 3,
                  4 * gy ** 3 * gz, 4 * gx * gz ** 3, 4 * gy * gz ** 3, 6 * gx ** 2 * gy ** 2,
                  6 * gx ** 2 * gz ** 2, 6 * gy ** 2 * gz ** 2, 12 * gx ** 2 * gy * gz,
                  12 * gx * gy ** 2 * gz, 12 * gx * gy * gz ** 2])
    return np.concatenate([np.ones((1, b.size)), -b * a, (b * b / 6.0) * w]).T
