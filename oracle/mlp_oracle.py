"""FP64 restatement of the device MLP (TEST INFRASTRUCTURE ONLY).

Only ``tests/`` may import this module.  The network has no counterpart in
the reference (the learned FIC initialiser and fit surrogate are additions);
what it restates is the textbook math the device kernels implement:

* layers in -> hidden x depth -> out, hidden activations
  gelu(z) = z/2 (1 + erf(z/sqrt 2)), linear output;
* flat parameter vector: per layer W [out x in] row-major, then b [out];
* loss = mean over all n x out entries of (y - t)^2;
* Adam (Kingma & Ba 2015, Algorithm 1) with bias correction;
* ``bf16=True`` rounds every GEMM operand to bfloat16 (round to nearest
  even) exactly where the device kernel does: X and W in the forward
  product, dZ and A in the weight gradient, dZ and W in the input gradient.
  Accumulation, biases and activations stay FP64 here (FP32 on the device).

Pins (tests/test_mlp_oracle.py): central finite differences of the loss,
torch.autograd in float64, and the closed-form first Adam step
(p - lr * sign(g)).
"""
from __future__ import annotations

import math

import numpy as np

_SQRT2 = math.sqrt(2.0)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / _SQRT2))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / _SQRT2)) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def bf16_round(a):
    """float64 -> nearest bfloat16 (ties to even) -> float64, via float32."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def dims(in_dim, hidden, depth, out_dim):
    return [in_dim] + [hidden] * depth + [out_dim]


def unpack(p, ds):
    """Flat parameter vector -> [(W, b)] per layer."""
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for l in range(len(ds) - 1):
        k, n = ds[l], ds[l + 1]
        W = p[o:o + n * k].reshape(n, k)
        o += n * k
        b = p[o:o + n]
        o += n
        out.append((W, b))
    assert o == p.size
    return out


def forward(p, ds, X, bf16=False):
    """Returns the output and the per-layer (input, pre-activation) cache."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    A = np.asarray(X, np.float64)
    cache = []
    layers = unpack(p, ds)
    for l, (W, b) in enumerate(layers):
        Z = r(A) @ r(W).T + b
        cache.append((A, Z))
        A = Z if l == len(layers) - 1 else gelu(Z)
    return A, cache


def loss_grad(p, ds, X, T, bf16=False):
    """Mean squared error and its gradient w.r.t. the flat parameters."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Y, cache = forward(p, ds, X, bf16)
    T = np.asarray(T, np.float64)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dZ = 2.0 * (Y - T) / n
    layers = unpack(p, ds)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        A, _ = cache[l]
        gW = r(dZ).T @ r(A)
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dZ = (r(dZ) @ r(W)) * gelu_grad(cache[l - 1][1])
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam_step(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
