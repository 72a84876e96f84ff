"""Learned Newton predictor: trained on the GPU by the library's own kernels
on step pairs the batched engine records, served in-batch.

Checks:
* pair collection: one record per window of every accepted step after the
  history fills, features and targets finite, the mask covering exactly
  the instance's unknowns;
* serving leaves step control and tolerances unchanged: waveforms within
  RELTOL of the FP64 oracle, accepted step counts within 2 % of the run
  without the network;
* what the network gains: Newton iterations per accepted step with and
  without it, which must fall strictly on both families (ring oscillators,
  about 2.07 iterations per step with the polynomial predictor, and the
  power-grid family, about 1.10), and below what the extrapolation the
  network corrects achieves alone.
"""
import numpy as np
import pytest
import torch

import xyce_amd as X
from oracle.pyoracle import Oracle
from tests.helpers import edge_shift, waveform_error, waveform_error_tube_interp
from xyce_amd import circuits as C
from xyce_amd import predictor as P

pytestmark = pytest.mark.gpu


def _rings(n, seed0):
    rng = np.random.default_rng(seed0)
    return [C.ring(int(s), seed=seed0 + i, mc_seed=seed0 + i) for i, s in enumerate(rng.choice([11, 13, 15, 17, 19, 21], n))]


def test_pairs_are_collected(gpu):
    cs = _rings(8, 1)
    x, t, m, c = P.collect_pairs(cs, 0.5e-9, 1e-12, 1e-11, capacity=1000)
    assert c["done"] == 8 and c["failed"] == 0
    # one record per accepted step once 4 points (the operating point and 3 steps) exist
    assert x.shape[0] == c["pairs"]
    assert c["pairs"] >= c["steps"] - 8 * 4
    assert torch.isfinite(x.float()).all() and torch.isfinite(t).all()
    assert float(m.sum(1).min()) >= 13 and float(m.sum(1).max()) <= 23
    # targets are in Newton weights, clipped to +-256; the padding rows are zero
    assert float(t.abs().max()) <= 256.0
    assert float((t * (1 - m)).abs().max()) == 0.0
    # the sixth feature plane is x_n itself: the supply node sits at 1.8 V in every record
    xn = x[:, 5 * X.PRED_WIN:6 * X.PRED_WIN].float()
    assert float(((xn - 1.8).abs() <= 0.01).any(dim=1).float().min()) == 1.0


def _grids(n, seed0):
    """The power-grid family: 10x10 RC meshes, 4 supply pads, 16 PWL sinks
    at seeded nodes and times."""
    return [C.power_grid(10, 10, seed=seed0 + 3 * i, pad_pitch=5, n_sinks=16, tstop=2e-9) for i in range(n)]


def _evaluate(net, test_set, tstop, hmax, oscillator=False, check=(0, 3), base="poly"):
    """Runs the test set with and without the network; prints iterations per
    step and waveform errors.  Non-oscillating circuits: both runs within
    RELTOL*|V| + ABSTOL of the FP64 oracle at the same tolerances.
    Oscillators: edges move by a fraction of a step whenever the Newton path
    changes (the oracle at the reference's tolerances is itself ~10 mV from
    an accurate solution after 2 ns), so both runs are measured against an
    accurate oracle run (RELTOL 1e-6): the network's run may move the edges,
    and differ from it once aligned within that shift, at most twice as much
    as the reference's own answer at these tolerances or the polynomial run."""
    poly, b0 = P.run(test_set, tstop, 1e-12, hmax, record_max=40000)
    learned, b1 = P.run(test_set, tstop, 1e-12, hmax, net=net, record_max=40000, base=base)
    # the network's own share: the same serving path with a network whose
    # output is zero proposes the base extrapolation the network corrects
    zero = P.new_net(max_rows=net.max_rows)
    zero.set_params(np.zeros(zero.n_params, np.float32))
    alone, b2 = P.run(test_set, tstop, 1e-12, hmax, net=zero, record_max=0, base=base)
    n = len(test_set)
    assert poly["done"] == learned["done"] == alone["done"] == n
    ipb = poly["newton_iters"] / poly["steps"]
    ipl = learned["newton_iters"] / learned["steps"]
    ipc = alone["newton_iters"] / alone["steps"]
    print(f"Newton iterations per accepted step: polynomial {ipb:.4f}, learned {ipl:.4f} "
          f"({learned['learned_steps']} steps from the network, {learned['learned_newton_failures']} retried), "
          f"its base ({base}) alone {ipc:.4f}; accepted steps {poly['steps']} / {learned['steps']} / {alone['steps']}")
    tol = 1e-3 * 1.8 + 1e-6
    for i in check:
        o = Oracle(test_set[i])
        if oscillator:
            o.set_tolerances(1e-6, 1e-9, 1e-5, 1e-9)
        r = o.tran(tstop, 1e-12, hmax)
        assert r["status"] == 0
        if not oscillator:
            errs = {}
            for name, b in (("polynomial", b0), ("learned", b1)):
                tg, xg = b.points(i, o.n)
                errs[name] = waveform_error(tg, xg, r["t"], r["x"])
            print(f"instance {i}: waveform error vs FP64 oracle: polynomial {errs['polynomial']:.2e} V, "
                  f"learned {errs['learned']:.2e} V")
            assert errs["polynomial"] <= tol and errs["learned"] <= tol
            continue
        # oscillators: against the accurate run, the edges' time shift and,
        # aligned within that shift, the waveform error, each no more than twice
        # those of the reference's own answer at these tolerances (the oracle at
        # the default tolerances) or of the polynomial run (the last 50 ps left
        # out so that all runs cross the same number of times)
        rd = Oracle(test_set[i]).tran(tstop, 1e-12, hmax)
        cut = lambda t, x: (t[t <= tstop - 50e-12], x[t <= tstop - 50e-12])
        ta, xa = cut(r["t"], r["x"])
        res = {}
        for name, tx in (("reference", (rd["t"], rd["x"])), ("polynomial", b0.points(i, o.n)),
                         ("learned", b1.points(i, o.n))):
            tg, xg = cut(*tx)
            sh = edge_shift(tg, xg, ta, xa)
            res[name] = (sh, waveform_error_tube_interp(tg, xg, ta, xa, sh + 1e-13))
        print(f"instance {i} against the accurate oracle run: edge shift / aligned waveform error: " +
              ", ".join(f"{k} {v[0] * 1e12:.2f} ps / {v[1]:.2e} V" for k, v in res.items()))
        assert res["learned"][0] <= 2 * max(res["reference"][0], res["polynomial"][0])
        assert res["learned"][1] <= 2 * max(res["reference"][1], res["polynomial"][1])
    return poly, learned, ipb, ipl, ipc


def test_trained_predictor_on_power_grids(gpu):
    """The power-grid family: linear, so a proposal within the Newton
    weights saves the second solve of a step.  Trained on 2M step pairs
    (5800 grids over 2 ns: about 357 accepted steps each), 8 epochs."""
    x, t, m, c = P.collect_pairs(_grids(5800, 100), 2e-9, 1e-12, 0.0, capacity=420)
    assert x.shape[0] >= 2_000_000
    net = P.new_net(max_rows=2048, seed=5)
    hist = P.train(net, x, t, m, epochs=8, batch=2048, lr=1e-3)
    print(f"{x.shape[0]} pairs; loss {hist['loss']} (polynomial predictor {hist['poly_loss']:.3e})")
    assert hist["loss"][-1] < hist["poly_loss"]
    poly, learned, ipb, ipl, _ = _evaluate(net, _grids(16, 5000), 2e-9, 0.0)
    assert abs(learned["steps"] - poly["steps"]) <= 0.02 * poly["steps"] + 16
    assert ipl < ipb


def test_trained_predictor_on_rings(gpu):
    """Ring oscillators (nonlinear, about 2.07 Newton iterations per step with
    the polynomial predictor): the network corrects the cubic extrapolation,
    trained on 0.93 M step pairs of 1024 rings for 32 epochs; the iterations per step must
    fall below both the polynomial predictor's and the cubic's alone."""
    x, t, m, c = P.collect_pairs(_rings(1024, 100), 2e-9, 1e-12, 1e-11, capacity=1200, base="cubic")
    net = P.new_net(max_rows=2048, seed=5)
    hist = P.train(net, x, t, m, epochs=32, batch=2048, lr=1e-3)
    print(f"{x.shape[0]} pairs; loss {hist['loss'][-1]:.4g} (zero correction {hist['poly_loss']:.4g})")
    assert hist["loss"][-1] < hist["poly_loss"]
    poly, learned, ipb, ipl, ipc = _evaluate(net, _rings(16, 900), 2e-9, 1e-11, oscillator=True, check=(0, 7),
                                             base="cubic")
    assert abs(learned["steps"] - poly["steps"]) <= 0.02 * poly["steps"] + 16
    assert ipl < ipc < ipb
