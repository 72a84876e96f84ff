"""Host side of the learned Newton predictor.

The network (``xyce_amd.MLP``) is trained on step pairs that the batched
engine itself records while it runs with the polynomial predictor
(``Batch.set_predictor(1, ...)``), and then served inside a batch
(``Batch.set_predictor(2, net)``).  Training runs entirely on the GPU
through the library's forward, backward and Adam kernels; torch is used only
to hold device buffers and to draw minibatch permutations.

The polynomial predictor of the reference (OneStep::obtainPredictor,
src/TimeIntegrationPKG/N_TIA_OneStep.C:135) stays the fallback and the LTE
reference, so the network can only change how many Newton iterations a step
takes, never the step control or the tolerances.
"""
from __future__ import annotations

import time

import numpy as np

from . import MLP, PRED_FEAT, PRED_WIN, Batch


def collect_pairs(circuits, tstop, tstep, hmax, capacity, chunk=512, max_launches=2_000_000, base="poly"):
    """Runs the batch to the end with pair collection on; returns (x, t, m,
    counters) with x bf16 [r][PRED_FEAT], t/m float32 [r][PRED_WIN] on the GPU
    (targets relative to ``base``, 'poly' or 'cubic')."""
    b = Batch(circuits, tstop, tstep, hmax)
    b.set_predictor(1, pair_capacity=capacity, base=base)
    c = b.run_to_end(chunk=chunk, max_launches=max_launches)
    x, t, m = b.pairs()
    c.update(b.pred_counters())
    b.close()
    return x, t, m, c


def train(net: MLP, x, t, m, epochs=4, batch=2048, lr=1e-3, seed=0, log=None):
    """Minibatch Adam on masked MSE; returns the per-epoch mean loss and the
    loss of the zero correction (the polynomial predictor) for reference."""
    import torch
    assert batch % 64 == 0 and batch <= net.max_rows
    n = x.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    base = float(((t * m) ** 2).sum() / m.sum())
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        tot, k = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            mb = m[idx].contiguous()
            cnt = int(mb.sum().item())
            if cnt == 0:
                continue
            tot += net.loss_backward(x[idx].contiguous(), t[idx].contiguous(), mb, cnt)
            net.adam(lr)
            k += 1
        hist.append(tot / max(k, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e} (polynomial predictor {base:.4e}), {time.time() - t0:.1f} s")
    net.sync()
    return dict(loss=hist, poly_loss=base, seconds=time.time() - t0)


def run(circuits, tstop, tstep, hmax, net=None, record_max=0, chunk=512, max_launches=2_000_000, base="poly",
        deadzone=0.0):
    """Runs a batch to the end, with the network serving predictions when
    given (as corrections of ``base`` above ``deadzone`` Newton weights);
    returns (counters, batch)."""
    b = Batch(circuits, tstop, tstep, hmax, record_max=record_max)
    if net is not None:
        b.set_predictor(2, net, base=base, deadzone=deadzone)
    c = b.run_to_end(chunk=chunk, max_launches=max_launches)
    if net is not None:
        c.update(b.pred_counters())
    return c, b


def new_net(max_rows=4096, seed=5):
    """The network of the learned predictor: 1600 -> 3 x 512 GELU -> 256."""
    return MLP(PRED_FEAT, 512, 3, PRED_WIN, max_rows=max_rows, seed=seed)
