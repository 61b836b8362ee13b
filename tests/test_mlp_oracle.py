"""Pins of the MLP oracle (no GPU): backward against central finite
differences in float64, bf16 rounding against its definition."""
import numpy as np

from oracle import mlp


def _net(rng, I=16, H=32, O=4):
    return [rng.standard_normal((H, I)) * 0.3, rng.standard_normal(H) * 0.1,
            rng.standard_normal((H, H)) * 0.2, rng.standard_normal(H) * 0.1,
            rng.standard_normal((O, H)) * 0.2, rng.standard_normal(O) * 0.1]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((24, 16))
    T = rng.standard_normal((24, 4))
    params = _net(rng)
    loss, g, _ = mlp.forward_backward(X, T, *params, round_bf16=False)
    names = ["w1", "b1", "w2", "b2", "w3", "b3"]
    for pi, name in enumerate(names):
        P = params[pi]
        for idx in [tuple(rng.integers(0, s) for s in P.shape) for _ in range(6)]:
            h = 1e-6
            Pp = [p.copy() for p in params]
            Pm = [p.copy() for p in params]
            Pp[pi][idx] += h
            Pm[pi][idx] -= h
            lp = mlp.forward_backward(X, T, *Pp, round_bf16=False)[0]
            lm = mlp.forward_backward(X, T, *Pm, round_bf16=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(g[name][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, idx, g[name][idx], fd)


def test_gelu_grad_is_derivative():
    x = np.linspace(-6, 6, 1001)
    h = 1e-6
    np.testing.assert_allclose(mlp.gelu_grad(x), (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h),
                               atol=1e-8)


def test_bf16_rounding():
    # exactly representable values stay; halfway cases round to even
    assert mlp.bf16(np.array([1.0]))[0] == 1.0
    one_ulp = 2.0 ** -7
    assert mlp.bf16(np.array([1.0 + one_ulp / 2]))[0] == 1.0              # tie -> even
    assert mlp.bf16(np.array([1.0 + 3 * one_ulp / 2]))[0] == 1.0 + 2 * one_ulp
    x = np.random.default_rng(1).standard_normal(1000)
    r = mlp.bf16(x)
    assert np.max(np.abs(r - x) / np.abs(x)) <= 2.0 ** -8


def _gelu_poly_constants():
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "pycompute_amd", "csrc",
                            "mlp_lib.hip")).read()
    body = re.search(r"kGeluP\[6\] = \{([^}]*)\}", src).group(1)
    coef = [float(v.strip().rstrip("f")) for v in body.split(",")]
    clamp = float(re.search(r"kGeluClamp = ([0-9.]+)f", src).group(1))
    return coef, clamp


def test_inference_gelu_polynomial_bounds():
    """The inference kernels' polynomial GELU (x (1/2 + z P(z^2)), z = clamp(x
    / c, -1, 1)) against the tanh-form GELU in float64: absolute error <=
    2.8e-4 on |x| <= c, relative error <= 2.6e-4 beyond (s held at 0 or 1)."""
    coef, c = _gelu_poly_constants()
    x = np.linspace(-200.0, 200.0, 400001)
    z = np.clip(x / c, -1.0, 1.0)
    g = x * (0.5 + z * np.polyval(coef[::-1], z * z))
    err = np.abs(g - mlp.gelu(x))
    inside = np.abs(x) <= c
    assert err[inside].max() <= 2.8e-4
    assert (err[~inside] / np.abs(x[~inside])).max() <= 2.6e-4
    # the odd part vanishes at 0 and s stays within [0, 1]
    s = 0.5 + z * np.polyval(coef[::-1], z * z)
    assert s.min() >= 0.0 and s.max() <= 1.0 and abs(np.interp(0.0, x, g)) < 1e-12


def test_fp16_polynomial_forward_beats_bf16_operands():
    """End to end on a random 16-128-128-4 network, inference numerics
    simulated in numpy (fp16 operands and the kernel's fp16 polynomial GELU,
    z = 2 clamp(x / 2c + 1/2, 0, 1) - 1, fp32 accumulate) stay closer to the
    float64 forward than bf16 operands with the exact GELU."""
    coef, c = _gelu_poly_constants()
    h = np.float16

    def gelu16(a):
        x = a.astype(h)
        w = np.minimum(np.maximum(x * h(0.5 / c) + h(0.5), h(0.0)), h(1.0))
        z = w * h(2.0) - h(1.0)
        t = z * z
        p = t * h(coef[5]) + h(coef[4])
        for k in coef[3::-1]:
            p = p * t + h(k)
        return (x * (z * p + h(0.5))).astype(np.float64)

    rng = np.random.default_rng(0)

    def init(fi, fo):
        b = 2.0 * np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-b, b, (fo, fi))

    w1, w2, w3 = init(16, 128), init(128, 128), init(128, 4)
    b1, b2, b3 = rng.normal(0, 0.3, 128), rng.normal(0, 0.3, 128), rng.normal(0, 0.3, 4)
    X = rng.uniform(0.05, 1.0, (4000, 16))
    exact = mlp.gelu(mlp.gelu(X @ w1.T + b1) @ w2.T + b2) @ w3.T + b3
    q = [a.astype(h).astype(np.float64) for a in (X, w1, w2, w3)]
    f16 = gelu16(gelu16(q[0] @ q[1].T + b1) @ q[2].T + b2) @ q[3].T + b3
    r = [mlp.bf16(a) for a in (X, w1, w2, w3)]
    a1 = mlp.bf16(mlp.gelu(r[0] @ r[1].T + b1))
    bf = mlp.bf16(mlp.gelu(a1 @ r[2].T + b2)) @ r[3].T + b3
    e16, ebf = np.abs(f16 - exact), np.abs(bf - exact)
    assert np.percentile(e16, 99) < np.percentile(ebf, 99)
    assert np.median(e16) < np.median(ebf) and e16.max() < 1e-2
