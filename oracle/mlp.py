"""CPU oracle of the MLP initialiser -- TEST INFRASTRUCTURE ONLY.

numpy float64 forward and backward of the in -> H -> H -> out GELU network
with mean-squared-error loss, with the option of rounding every matrix-core
operand to bf16 exactly where the HIP kernels do (GEMM inputs: activations
and weights), so the GPU gradients can be compared tightly. The float64 path
without rounding is pinned against central finite differences in
tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

C = 0.7978845608028654  # sqrt(2/pi)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C * (1 + 3 * 0.044715 * x * x)


def forward_backward(X, T, w1, b1, w2, b2, w3, b3, round_bf16=True):
    """X (B, in), T (B, out). Returns loss and the gradient dict."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    B, O = T.shape
    Z1 = q(X) @ q(w1).T + b1
    H1 = gelu(Z1)
    Z2 = q(H1) @ q(w2).T + b2
    H2 = gelu(Z2)
    Y = q(H2) @ q(w3).T + b3
    D = Y - T
    loss = np.sum(D * D) / (B * O)
    dY = 2.0 * D / (B * O)
    g = {}
    g["w3"] = q(dY).T @ q(H2)
    g["b3"] = dY.sum(0)
    dZ2 = (q(dY) @ q(w3)) * gelu_grad(Z2)
    g["w2"] = q(dZ2).T @ q(H1)
    g["b2"] = dZ2.sum(0)
    dZ1 = (q(dZ2) @ q(w2)) * gelu_grad(Z1)
    g["w1"] = q(dZ1).T @ q(X)
    g["b1"] = dZ1.sum(0)
    return loss, g, Y


def forward(X, w1, b1, w2, b2, w3, b3, round_bf16=True):
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    H1 = gelu(q(X) @ q(w1).T + b1)
    H2 = gelu(q(H1) @ q(w2).T + b2)
    return q(H2) @ q(w3).T + b3
