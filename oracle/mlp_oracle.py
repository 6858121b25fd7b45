"""TEST INFRASTRUCTURE: a NumPy FP64 restatement of the surrogate MLP
(n_in -> H -> H -> n_out, exact-erf GELU, mean squared error) and its
hand-derived backward pass, used to check the GPU gradients.

Pinned independently by a central finite-difference check of its own
gradient (tests/test_surrogate_cpu.py).
"""
import numpy as np
from scipy.special import erf


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(params, x):
    W1, b1, W2, b2, W3, b3 = params
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ W3.T + b3
    return y, (z1, h1, z2, h2)


def loss_and_grad(params, x, t):
    W1, b1, W2, b2, W3, b3 = params
    y, (z1, h1, z2, h2) = forward(params, x)
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / e.size
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ W3) * gelu_grad(z2)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * gelu_grad(z1)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    return loss, [gW1, gb1, gW2, gb2, gW3, gb3]
