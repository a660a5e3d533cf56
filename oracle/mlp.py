"""CPU restatement of the learned-guess perceptron (test infrastructure only).

The reference has no learned guess (its SCF starts from PySCF's minao/SAD guess through
`pyscf.scf.hf.RHF`, cuobc/scf/hf.py:272-274), so this oracle restates the network the
product defines: y = W3^T silu(W2^T silu(W1^T x + b1) + b2) + b3, row-major activations
[batch][features], weights [in][out].  `emulate=True` rounds to bf16 (nearest even) at the
points where the HIP path stores bf16: the input, the weights, both hidden activations and
the three backward deltas; products accumulate in fp32.  `emulate=False` runs the same
algebra in float64, which `tests/test_guess_oracle.py` pins against central finite
differences of the loss.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, emulate=True):
    """params: dict W1, W2, W3 ([in][out]), b1, b2, b3.  Returns (y, cache)."""
    dt = np.float32 if emulate else np.float64
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=dt))
    W = [r(params[f'W{i}']) for i in (1, 2, 3)]
    b = [np.asarray(params[f'b{i}'], dtype=dt) for i in (1, 2, 3)]
    a0 = r(x)
    z1 = (a0 @ W[0]).astype(dt) + b[0]
    a1 = r(silu(z1))
    z2 = (a1 @ W[1]).astype(dt) + b[1]
    a2 = r(silu(z2))
    y = (a2 @ W[2]).astype(dt) + b[2]
    return y, dict(a0=a0, z1=z1, a1=a1, z2=z2, a2=a2, W=W)


def backward(cache, dy, emulate=True):
    """Gradients (dW1, dW2, dW3, db1, db2, db3) of sum(dy * y)."""
    dt = np.float32 if emulate else np.float64
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=dt))
    W = cache['W']
    dy = np.asarray(dy, dtype=dt)
    db3 = dy.sum(0)
    dz3 = r(dy)
    dW3 = cache['a2'].T @ dz3
    g2 = (dz3 @ W[2].T) * silu_grad(cache['z2'])
    db2 = g2.sum(0)
    dz2 = r(g2)
    dW2 = cache['a1'].T @ dz2
    g1 = (dz2 @ W[1].T) * silu_grad(cache['z1'])
    db1 = g1.sum(0)
    dz1 = r(g1)
    dW1 = cache['a0'].T @ dz1
    return dict(W1=dW1, W2=dW2, W3=dW3, b1=db1, b2=db2, b3=db3)


def mse(y, t, mask, scale):
    """(loss = sum mask (y - t)^2, dy = scale * mask * (y - t))."""
    d = y - t
    return float(np.sum(mask * d * d)), scale * mask * d


def adam(w, g, m1, m2, lr, b1, b2, eps, step):
    """One bias-corrected Adam step; returns (w, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    w = w - lr * (m1 * c1) / (np.sqrt(m2 * c2) + eps)
    return w, m1, m2
