"""Pins the CPU oracle of the learned-predictor network (oracle/mlp_oracle.py).

* BF16 rounding: hand-worked ties-to-even cases and torch's own float32 ->
  bfloat16 conversion on random values;
* GELU: closed-form values, the identity gelu(z) - gelu(-z) = z, its
  derivative against central differences;
* forward and gradients (float64 mode) against torch autograd on the same
  parameters, and the gradients against central finite differences;
* the Adam step against torch.optim.Adam.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def test_bf16_round_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 0.0, 3.0e38, 1e-40], np.float32)
    r = M.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # halfway, even mantissa below
    assert r[2] == 1.0 + 2 ** -6          # halfway, even mantissa above
    assert r[3] == 1.0 + 2 ** -7          # exactly representable
    assert r[4] == -2.5 and r[5] == 0.0
    rng = np.random.default_rng(0)
    v = (rng.standard_normal(10000) * np.exp(rng.uniform(-20, 20, 10000))).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16_round(v), ref)


def test_gelu_closed_forms():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu_d(np.array([0.0]))[0] - 0.5) < 1e-15
    z = np.linspace(-6, 6, 121)
    np.testing.assert_allclose(M.gelu(z) - M.gelu(-z), z, atol=1e-14)
    assert abs(M.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    h = 1e-6
    np.testing.assert_allclose(M.gelu_d(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate="none").numpy()
    np.testing.assert_allclose(M.gelu(z), ref, rtol=1e-14, atol=1e-15)


def _tiny(seed=0, in_dim=6, width=5, hidden=2, out_dim=3, rows=7):
    rng = np.random.default_rng(seed)
    dims = M.layer_dims(in_dim, width, hidden, out_dim)
    n = sum(di * do + do for di, do in dims)
    p = rng.normal(0, 0.5, n)
    X = rng.normal(0, 1, (rows, in_dim))
    T = rng.normal(0, 1, (rows, out_dim))
    mask = (rng.random((rows, out_dim)) < 0.8).astype(np.float64)
    return dims, p, X, T, mask


def test_forward_and_grads_match_torch_autograd():
    dims, p, X, T, mask = _tiny()
    count = 5
    loss, g = M.loss_grads(p, X, T, mask, count, dims, bf16=False)
    pt = torch.tensor(p, requires_grad=True)
    h = torch.tensor(X)
    off = 0
    for l, (di, do) in enumerate(dims):
        W = pt[off: off + di * do].reshape(do, di); off += di * do
        b = pt[off: off + do]; off += do
        z = h @ W.T + b
        h = torch.nn.functional.gelu(z, approximate="none") if l + 1 < len(dims) else z
    lt = (((h - torch.tensor(T)) * torch.tensor(mask)) ** 2).sum() / count
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-13 * abs(loss)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-11, atol=1e-13)


def test_grads_match_finite_differences():
    dims, p, X, T, mask = _tiny(seed=3)
    loss, g = M.loss_grads(p, X, T, mask, 4, dims, bf16=False)
    eps = 1e-6
    for i in range(0, p.size, 3):
        pp = p.copy(); pp[i] += eps
        pm = p.copy(); pm[i] -= eps
        fd = (M.loss_grads(pp, X, T, mask, 4, dims, bf16=False)[0] - M.loss_grads(pm, X, T, mask, 4, dims, bf16=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_matches_torch():
    rng = np.random.default_rng(1)
    p = rng.normal(size=50)
    pt = torch.tensor(p, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p); v = np.zeros_like(p)
    for t in range(1, 4):
        g = rng.normal(size=50)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = M.adam_step(p, g, m, v, t, 1e-2)
    np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-14)
