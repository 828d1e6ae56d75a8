"""Pins of the FP64 MLP restatement (oracle/mlp_oracle.py) that the device
network's gradients are checked against: finite differences, torch.autograd
in float64, torch's bfloat16 rounding and torch.optim.Adam."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def _net(in_dim, hid, depth, out_dim, seed):
    ds = M.dims(in_dim, hid, depth, out_dim)
    rng = np.random.default_rng(seed)
    n = sum(ds[l] * ds[l + 1] + ds[l + 1] for l in range(len(ds) - 1))
    return ds, rng.normal(0, 0.5, n)


def test_gradient_matches_central_differences():
    ds, p = _net(3, 5, 2, 2, 0)
    rng = np.random.default_rng(1)
    X, T = rng.normal(size=(7, 3)), rng.normal(size=(7, 2))
    _, g = M.loss_grad(p, ds, X, T)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_grad(p + e, ds, X, T)[0] - M.loss_grad(p - e, ds, X, T)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-7 * max(1.0, np.abs(g).max())


def test_gradient_matches_torch_autograd():
    ds, p = _net(5, 16, 3, 2, 2)
    rng = np.random.default_rng(3)
    X, T = rng.normal(size=(33, 5)), rng.normal(size=(33, 2))
    loss, g = M.loss_grad(p, ds, X, T)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    A = torch.tensor(X)
    o = 0
    for l in range(len(ds) - 1):
        k, n = ds[l], ds[l + 1]
        W = pt[o:o + n * k].reshape(n, k)
        o += n * k
        b = pt[o:o + n]
        o += n
        A = A @ W.T + b
        if l < len(ds) - 2:
            A = torch.nn.functional.gelu(A)
    lt = torch.mean((A - torch.tensor(T)) ** 2)
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-13 * abs(loss)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(0, 3, 5000), [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -(1 + 2 ** -8 + 2 ** -20), 0.0]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16_round(x), ref)
    assert M.bf16_round(np.array([1 + 2 ** -8]))[0] == 1.0           # tie -> even
    assert M.bf16_round(np.array([1 + 3 * 2 ** -9]))[0] == 1 + 2 ** -7  # above the tie -> up


def test_bf16_gradient_differs_from_fp32_but_close():
    ds, p = _net(4, 32, 2, 1, 5)
    rng = np.random.default_rng(6)
    X, T = rng.normal(size=(64, 4)), rng.normal(size=(64, 1))
    _, g = M.loss_grad(p, ds, X, T)
    _, gb = M.loss_grad(p, ds, X, T, bf16=True)
    rel = np.linalg.norm(g - gb) / np.linalg.norm(g)
    assert 1e-5 < rel < 3e-2


def test_adam_matches_torch_and_first_step_closed_form():
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=50)
    gs = [rng.normal(size=50) for _ in range(4)]
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m1, m2 = p0.copy(), np.zeros(50), np.zeros(50)
    for s, g in enumerate(gs, 1):
        p, m1, m2 = M.adam_step(p, g, m1, m2, s, 1e-3)
        pt.grad = torch.tensor(g)
        opt.step()
        np.testing.assert_allclose(p, pt.detach().numpy(), rtol=0, atol=1e-15)
        if s == 1:
            np.testing.assert_allclose(p, p0 - 1e-3 * np.sign(gs[0]), atol=1e-10)


def test_gelu_grad_is_derivative():
    z = np.linspace(-6, 6, 1001)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)
    assert M.gelu(np.array([0.0]))[0] == 0.0
