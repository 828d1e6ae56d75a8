"""Device MLP (mlp.hip: MFMA GEMMs, fused GELU epilogues, split-K weight
gradients, Adam) against the FP64 restatement in oracle/mlp_oracle.py."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _data(n, in_dim, out_dim, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, in_dim)).astype(np.float32)
    T = (np.sin(X[:, :1]) + 0.5 * X[:, 1:2] ** 2 - 0.3 * X[:, 2:3] * X[:, 0:1]).astype(np.float32)
    return X, np.repeat(T, out_dim, axis=1) * np.linspace(1, 0.5, out_dim, dtype=np.float32)


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("shape", [(5, 256, 4, 1, 2048), (7, 96, 2, 2, 333)])
def test_loss_and_gradient_match_oracle(bf16, shape):
    in_dim, hid, depth, out_dim, n = shape
    print("case: mlp grad", shape, "bf16", bf16, flush=True)
    X, T = _data(n, in_dim, out_dim, 1)
    with bnm.MLP(in_dim, hid, depth, out_dim, bf16=bf16, seed=3) as net:
        p = net.get_params()
        ds = net.dims()
        loss, g = net.loss_grad(X, T)
        Y = net.predict(X)
    Yo, _ = M.forward(p, ds, X, bf16)
    lo, go = M.loss_grad(p, ds, X, T, bf16)
    # FP32 accumulation vs FP64: relative 1e-4 on the loss, 1e-3 (norm) on the
    # gradient; with bf16 operands the oracle rounds the same operands, and an
    # operand that lands on the other side of a bf16 rounding boundary moves a
    # product by 2^-8 relative, so the bounds are 1e-3 / 2e-2.
    tl, tg = (1e-3, 2e-2) if bf16 else (1e-4, 1e-3)
    assert abs(loss - lo) <= tl * lo
    assert np.linalg.norm(g - go) <= tg * np.linalg.norm(go)
    assert np.abs(Y - Yo).max() <= (2e-2 if bf16 else 1e-4) * max(1.0, np.abs(Yo).max())
    # every layer's weight and bias gradient, not only the global norm
    o = 0
    for l in range(len(ds) - 1):
        for size in (ds[l] * ds[l + 1], ds[l + 1]):
            a, b = g[o:o + size], go[o:o + size]
            assert np.linalg.norm(a - b) <= 3 * tg * np.linalg.norm(b) + 1e-7, (l, size)
            o += size


def test_first_adam_step_matches_oracle():
    print("case: mlp adam step", flush=True)
    X, T = _data(512, 4, 1, 2)
    with bnm.MLP(4, 64, 2, 1, bf16=False, seed=5) as net:
        p0 = net.get_params()
        _, g = net.loss_grad(X, T)
        net.fit(X, T, epochs=1, batch=512, lr=1e-3, seed=0)   # one full batch = one step
        p1 = net.get_params()
    ds = net.dims()
    _, go = M.loss_grad(p0, ds, X, T)
    pe, _, _ = M.adam_step(p0.astype(np.float64), go, np.zeros_like(go), np.zeros_like(go), 1, 1e-3)
    # first Adam step moves each parameter by lr * g/(|g| + eps') ~ lr * sign(g):
    # compare where the gradient is clearly non-zero
    big = np.abs(go) > 1e-3 * np.abs(go).max()
    np.testing.assert_allclose(p1[big], pe[big], atol=2e-6)
    assert np.sign(g[big]).tolist() == np.sign(go[big]).tolist()


@pytest.mark.parametrize("bf16", [False, True])
def test_training_reduces_loss(bf16):
    print("case: mlp training bf16", bf16, flush=True)
    X, T = _data(16384, 5, 2, 3)
    Xv, Tv = _data(4096, 5, 2, 4)
    with bnm.MLP(5, 256, 4, 2, bf16=bf16, seed=7) as net:
        l0, _ = net.loss_grad(Xv, Tv)
        losses = net.fit(X, T, epochs=30, batch=512, lr=1e-3, seed=1)
        l1, _ = net.loss_grad(Xv, Tv)
        # held-out loss recomputed in FP64 from the trained parameters
        lo, _ = M.loss_grad(net.get_params(), net.dims(), Xv, Tv, bf16)
    print("val loss", l0, "->", l1, "train", losses[0], "->", losses[-1])
    assert np.all(np.isfinite(losses))
    assert l1 < 0.05 * l0
    assert abs(l1 - lo) <= 0.05 * lo + 1e-6
