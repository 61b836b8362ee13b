"""The committed golden fixtures (tests/golden, written by make_golden.py):
the oracle still reproduces them bit for bit, and what mathematics fixes about
them holds (known answer, exact solves of positive definite systems)."""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _full(packed, n):
    A = np.zeros((n, n))
    k = 0
    for i in range(n):
        for j in range(i + 1):
            A[i, j] = A[j, i] = packed[k]
            k += 1
    return A


def test_kat_fixture():
    with open(os.path.join(HERE, "kat_test_mats.json")) as fh:
        d = json.load(fh)
    A, b, x = np.array(d["A"]), np.array(d["b"]), np.array(d["x"])
    np.testing.assert_allclose(A @ x, b, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.gmw81_solve(A, b, dtype=np.float64), x, rtol=1e-12)
    np.testing.assert_allclose(x, [0.02397962, -0.01595351, -0.06262880, -0.16919653], rtol=1e-6)


@pytest.mark.parametrize("n", [4, 22])
def test_gmw81_stress_fixture(n):
    g = np.load(os.path.join(HERE, "gmw81_stress.npz"))
    mat, rhs = g[f"mat{n}"], g[f"rhs{n}"]
    pd = 0
    for v in range(mat.shape[1]):
        A = _full(mat[:, v].astype(np.float64), n)
        for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
            x = oracle.gmw81_solve(A.astype(dt), rhs[:, v].astype(dt), dtype=dt)
            np.testing.assert_array_equal(x, g[f"sol{n}_{tag}"][:, v])
        # where the scaled matrix is comfortably positive definite, GMW81 adds
        # nothing and the float64 solution is the exact one
        s = 1.0 / np.sqrt(np.abs(np.diag(A)))
        ev = np.linalg.eigvalsh(A * s[:, None] * s[None, :])
        if ev.min() > 1e-3:
            pd += 1
            np.testing.assert_allclose(g[f"sol{n}_f64"][:, v], np.linalg.solve(A, rhs[:, v]),
                                       rtol=1e-8, atol=0)
    if n == 4:        # the 22-dim stress systems are all indefinite after scaling
        assert pd > 0


def test_ivim_c1_fixture():
    d = np.load(os.path.join(HERE, "ivim_c1_256.npz"))
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        r = oracle.lm_run("ivim", d["consts"], d["data"].astype(dt), d["lb"].astype(dt),
                          d["ub"].astype(dt), d["p0"].astype(dt), 100, dtype=dt)
        for k in ("params", "f", "lam", "step_type"):
            np.testing.assert_array_equal(r[k], d[f"{k}_{tag}"], err_msg=f"{k} {tag}")
    # float32 tracks float64 on the final residual for nearly every voxel
    rel = np.abs(d["ssr_f32"] - d["ssr_f64"]) / d["ssr_f64"]
    assert np.mean(rel <= 1e-4) >= 0.95
    assert np.all(d["params_f64"] >= d["lb"]) and np.all(d["params_f64"] <= d["ub"])
