"""Writes the golden fixtures of tests/golden from the pinned CPU oracle.

    python tests/golden/make_golden.py

* kat_test_mats.json -- the reference's known-answer system
  (mtests/test_mats.m:49-56) with its float64 solution A \\ b.
* gmw81_stress.npz -- batches of ill-conditioned packed systems made by the
  recipe of test/linalg/test_gmw81.py:17-37 (10*U[0,1) entries raised to the
  8th power, diagonal x100), n = 4 and n = 22, with the oracle's float32 and
  float64 solutions of the reference solve (pivot, scaling, GMW81, LDL^T).
* ivim_c1_256.npz -- the first 256 voxels of the 10-b-value IVIM workload
  (seeded), the default initial guess, and the oracle's float32 and float64
  fit results after 100 launches.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

KAT_A = [[1.28651749e+02, -3.17334712e-01, -2.32065395e-01, -3.49286795e-02],
         [-3.17334712e-01, 4.08983994e+00, 8.86404276e-01, -6.62028670e-01],
         [-2.32065395e-01, 8.86404276e-01, 4.78870010e+00, 1.85799852e-01],
         [-3.49286795e-02, -6.62028670e-01, 1.85799852e-01, 1.95811105e+00]]
KAT_B = [3.1105268, -0.01635835, -0.35105333, -0.33321792]


def stress_systems(n, batch, rng):
    nm = n * (n + 1) // 2
    mat = 10.0 * rng.random((nm, batch)).astype(np.float32)
    rhs = 10.0 * rng.random((n, batch)).astype(np.float32)
    for _ in range(3):
        mat = (mat * mat).astype(np.float32)
        rhs = (rhs * rhs).astype(np.float32)
    k = 0
    for i in range(n):
        for j in range(i + 1):
            if i == j:
                mat[k] *= 100
            k += 1
    return mat, rhs


def full(packed, n):
    A = np.zeros((n, n))
    k = 0
    for i in range(n):
        for j in range(i + 1):
            A[i, j] = A[j, i] = packed[k]
            k += 1
    return A


def solve_all(mat, rhs, n, dtype):
    return np.stack([oracle.gmw81_solve(full(mat[:, v], n).astype(dtype), rhs[:, v].astype(dtype),
                                        dtype=dtype) for v in range(mat.shape[1])], axis=1)


def main():
    A = np.array(KAT_A)
    b = np.array(KAT_B)
    with open(os.path.join(HERE, "kat_test_mats.json"), "w") as fh:
        json.dump({"source": "mtests/test_mats.m:49-56", "A": KAT_A, "b": KAT_B,
                   "x": np.linalg.solve(A, b).tolist()}, fh, indent=1)

    rng = np.random.Generator(np.random.PCG64(81))
    out = {}
    for n, batch in ((4, 256), (22, 64)):
        mat, rhs = stress_systems(n, batch, rng)
        out[f"mat{n}"], out[f"rhs{n}"] = mat, rhs
        out[f"sol{n}_f32"] = solve_all(mat, rhs, n, np.float32)
        out[f"sol{n}_f64"] = solve_all(mat, rhs, n, np.float64)
    np.savez_compressed(os.path.join(HERE, "gmw81_stress.npz"), **out)

    from pycompute_amd import init, synth
    from pycompute_amd.models import IVIM
    w = synth.ivim_c1(4096, seed=1)
    y = np.ascontiguousarray(w["data"][:, :256])
    b = w["consts"][0]
    lo, hi = init.ivim_split(b)
    lb = np.repeat(np.array(IVIM.lb, np.float64)[:, None], 256, 1)
    ub = np.repeat(np.array(IVIM.ub, np.float64)[:, None], 256, 1)
    p0 = oracle.ivim_default_init(b, y, lb, ub, lo, hi)
    res = {"consts": w["consts"], "data": y, "lb": lb, "ub": ub, "p0": p0}
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        r = oracle.lm_run("ivim", w["consts"], y.astype(dt), lb.astype(dt), ub.astype(dt),
                          p0.astype(dt), 100, dtype=dt)
        for k in ("params", "f", "lam", "step_type"):
            res[f"{k}_{tag}"] = r[k]
        res[f"ssr_{tag}"] = oracle.ssr("ivim", w["consts"], y.astype(dt), r["params"], dtype=dt)
    np.savez_compressed(os.path.join(HERE, "ivim_c1_256.npz"), **res)


if __name__ == "__main__":
    main()
