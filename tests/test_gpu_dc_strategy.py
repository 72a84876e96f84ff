"""DC operating point strategy of the batched engine (Newton from initJct,
then gmin stepping, then source stepping, as Interface::spiceStrategy runs
them) against the oracle's restatement of the same strategy.

* small gate arrays (80 gates): whether plain Newton from initJct
  converges on them is decided by rounding (perturbing gmin or the
  temperature in the fifth digit flips it in the oracle), so the stage is
  printed, not compared; the operating point equals the oracle's within the
  DC Newton tolerance (1e-3 |V| + 1 uV), every gate output sits at the logic
  level of its inputs, and the transient that follows has the oracle's
  outcome (waveforms within RELTOL where it completes), the oracle following
  the engine's path (two valid paths end up to twice the DC tolerance apart
  at leakage-held nodes).  Started at gmin stepping they report that stage
  and reach the same point.  Larger arrays,
  which need gmin stepping robustly, are checked on the single-circuit
  engine (test_gpu_transient_engine.py);
* each stage started on its own (.OPTIONS NONLIN CONTINUATION=GMIN|SOURCE)
  reaches the operating point plain Newton finds.
"""
import numpy as np
import pytest

import xyce_amd as X
from oracle.pyoracle import Oracle
from tests.helpers import waveform_error
from tests.test_oracle_devices import _gate_logic
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu


def _dc_tol(x, x0):
    return bool(np.all(np.abs(x - x0) <= 1e-3 * np.abs(x0) + 1e-6))


@pytest.mark.parametrize("first", [None, "gmin"])
def test_gate_arrays_reach_the_oracle_operating_point(gpu, first):
    cs = [C.gate_dag(80, seed=8, frac_nand=0.8), C.gate_dag(80, seed=48, frac_nand=0.8), C.ring(11, seed=3),
          C.inverter_chain(3, TOX=2e-8, CJ=2e-4)]
    tstop = 0.3e-9
    b = X.Batch(cs, tstop, 1e-12, 1e-11, record_max=4000)
    if first:
        b.dc_strategy(first)
    b.run_to_end(chunk=256)
    for i, c in enumerate(cs):
        # the oracle along the same path: a DC point is fixed only to the DC
        # Newton tolerance, and on these arrays two valid paths (plain Newton,
        # gmin stepping) end up to twice that bar apart at leakage-held nodes
        stage = b.dc_stage(i)
        if first:
            assert stage == first
        else:
            assert stage in ("newton", "gmin")
        o = Oracle(c)
        if stage != "newton":
            o.set_dc_first_stage(stage)
        st, x0, *_ = o.dcop()
        assert st > 0 and o.dc_strategy() == stage
        print(f"instance {i}: DC stage {stage}, plain-strategy oracle {Oracle(c).dc_strategy() if i < 2 else 'newton'}")
        t_g, x_g = b.points(i, o.n)
        assert t_g[0] == 0.0 and _dc_tol(x_g[0], x0), np.max(np.abs(x_g[0] - x0))
        if i < 2:
            gid = o.pattern()["node_gid"]
            for k, lv in _gate_logic(c).items():
                assert abs(x_g[0, gid[k]] - 1.8 * lv) <= 1e-3
        # the transient from there: the oracle's outcome
        st_g, _ = b.instance(i, o.n)
        ro = Oracle(c)
        if stage != "newton":
            ro.set_dc_first_stage(stage)
        r = ro.tran(tstop, 1e-12, 1e-11)
        print(f"instance {i}: n={o.n}; transient gpu mode {st_g['mode']:.0f} steps {st_g['steps']:.0f}, "
              f"oracle status {r['status']} steps {r['steps']}")
        assert (st_g["mode"] == 2) == (r["status"] == 0)
        assert abs(st_g["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
        if r["status"] == 0:
            e = waveform_error(t_g[t_g <= 0.1e-9], x_g[t_g <= 0.1e-9], r["t"], r["x"])
            assert e <= 1e-3 * 1.8 + 1e-6, e


@pytest.mark.parametrize("stage", ["gmin", "source"])
def test_each_stage_reaches_the_newton_point(gpu, stage):
    cs = [C.ring(11, seed=3), C.inverter_chain(3, TOX=2e-8, CJ=2e-4)]
    b = X.Batch(cs, 0.05e-9, 1e-12, 1e-11, record_max=1000)
    b.dc_strategy(stage)
    cnt = b.run_to_end(chunk=256)
    assert cnt["done"] == len(cs) and cnt["failed"] == 0, cnt
    for i, c in enumerate(cs):
        o = Oracle(c)
        st, x0, *_ = o.dcop()
        assert st > 0 and o.dc_strategy() == "newton"
        assert b.dc_stage(i) == stage
        t_g, x_g = b.points(i, o.n)
        assert _dc_tol(x_g[0], x0), np.max(np.abs(x_g[0] - x0))
    # the strategy is fixed once a launch has run
    assert X.lib().xg_batch_dc_strategy(b.h, 0) < 0
