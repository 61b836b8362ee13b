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
