"""GPU kernels of the learned-predictor network against the CPU oracle
(oracle/mlp_oracle.py in its BF16-emulating mode).

Bars: outputs within 1e-2 of the oracle's largest |y| entrywise, loss within
1e-3 relative, every parameter tensor's gradient within 2e-2 relative
(Frobenius) — the slack covers FP32 accumulation order and the BF16
rounding flips of hidden activations it causes.  A small regression problem
checks that training with the GPU Adam kernel reduces the loss 10-fold and
that one GPU Adam step equals the oracle's.
"""
import numpy as np
import pytest
import torch

import xyce_amd as X
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _data(rows, in_dim, out_dim, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (rows, in_dim)).astype(np.float32)
    t = rng.normal(0, 1, (rows, out_dim)).astype(np.float32)
    mask = (rng.random((rows, out_dim)) < 0.9).astype(np.float32)
    return x, t, mask


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda().contiguous()


@pytest.mark.parametrize("in_dim,width,hidden,out_dim,rows", [(128, 128, 2, 64, 128), (1600, 512, 3, 256, 256)])
def test_forward_and_gradients_match_oracle(gpu, in_dim, width, hidden, out_dim, rows):
    net = X.MLP(in_dim, width, hidden, out_dim, max_rows=rows, seed=7)
    # give the output layer full-size weights so every layer matters
    p = net.params()
    rng = np.random.default_rng(11)
    dims = net.dims()
    off = 0
    for di, do in dims:
        p[off: off + di * do] = rng.uniform(-1, 1, di * do) / np.sqrt(di)
        off += di * do
        p[off: off + do] = rng.normal(0, 0.1, do)
        off += do
    net.set_params(p)
    x, t, mask = _data(rows, in_dim, out_dim, 3)
    xd, td, md = _dev(x, torch.bfloat16), _dev(t), _dev(mask)
    y = torch.zeros(rows, out_dim, dtype=torch.float32, device="cuda")
    net.forward(xd, y)
    torch.cuda.synchronize()
    y_ref, _ = M.forward(p.astype(np.float64), x, dims, bf16=True)
    err = np.max(np.abs(y.cpu().numpy() - y_ref)) / np.max(np.abs(y_ref))
    print(f"forward max err {err:.2e}")
    assert err <= 1e-2
    count = int(mask.sum())
    loss, g = net.loss_backward(xd, td, md, count, want_grads=True)
    loss_ref, g_ref = M.loss_grads(p.astype(np.float64), x, t, mask, count, dims, bf16=True)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    off = 0
    for l, (di, do) in enumerate(dims):
        for name, n in (("W", di * do), ("b", do)):
            a, b = g[off: off + n], g_ref[off: off + n]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print(f"layer {l} {name}: rel grad err {rel:.2e}")
            assert rel <= 2e-2, (l, name, rel)
            off += n


def test_adam_step_matches_oracle(gpu):
    in_dim, width, hidden, out_dim, rows = 128, 128, 1, 64, 64
    net = X.MLP(in_dim, width, hidden, out_dim, max_rows=rows, seed=3)
    p0 = net.params().astype(np.float64)
    x, t, mask = _data(rows, in_dim, out_dim, 5)
    count = int(mask.sum())
    _, g = net.loss_backward(_dev(x, torch.bfloat16), _dev(t), _dev(mask), count, want_grads=True)
    net.adam(1e-3)
    net.sync()
    p1 = net.params()
    ref, _, _ = M.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(gpu):
    """Regression on a fixed random linear map of the input."""
    in_dim, width, hidden, out_dim, rows = 128, 256, 2, 64, 512
    net = X.MLP(in_dim, width, hidden, out_dim, max_rows=rows, seed=1)
    rng = np.random.default_rng(2)
    A = rng.normal(0, 1 / np.sqrt(in_dim), (out_dim, in_dim)).astype(np.float32)
    x = rng.normal(0, 1, (rows, in_dim)).astype(np.float32)
    t = x @ A.T
    mask = np.ones_like(t)
    xd, td, md = _dev(x, torch.bfloat16), _dev(t), _dev(mask)
    first = net.loss_backward(xd, td, md, rows * out_dim)
    for it in range(300):
        loss = net.loss_backward(xd, td, md, rows * out_dim)
        net.adam(2e-3)
    print(f"loss {first:.3e} -> {loss:.3e}")
    assert loss <= 0.1 * first
