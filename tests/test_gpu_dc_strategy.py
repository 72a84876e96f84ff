"""DC operating point strategy of the batched engine (Newton from initJct,
then gmin stepping, then source stepping, as Interface::spiceStrategy runs
them) against the oracle's restatement of the same strategy.

* gate arrays where plain Newton fails need gmin stepping in both; the
  engine reports the same stage, its operating point equals the oracle's
  within the DC Newton tolerance (1e-3 |V| + 1 uV), every gate output sits
  at the logic level of its inputs, and the transient that follows has the
  oracle's outcome (waveforms within RELTOL where it completes);
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


def test_gate_arrays_need_and_get_gmin_stepping(gpu):
    cs = [C.gate_dag(80, seed=8, frac_nand=0.8), C.gate_dag(80, seed=48, frac_nand=0.8), C.ring(11, seed=3),
          C.inverter_chain(3, TOX=2e-8, CJ=2e-4)]
    tstop = 0.3e-9
    b = X.Batch(cs, tstop, 1e-12, 1e-11, record_max=4000)
    b.run_to_end(chunk=256)
    assert [b.dc_stage(i) for i in range(len(cs))] == ["gmin", "gmin", "newton", "newton"]
    for i, c in enumerate(cs):
        o = Oracle(c)
        st, x0, *_ = o.dcop()
        assert st > 0 and b.dc_stage(i) == o.dc_strategy()
        t_g, x_g = b.points(i, o.n)
        assert t_g[0] == 0.0 and _dc_tol(x_g[0], x0), np.max(np.abs(x_g[0] - x0))
        if i < 2:
            gid = o.pattern()["node_gid"]
            for k, lv in _gate_logic(c).items():
                assert abs(x_g[0, gid[k]] - 1.8 * lv) <= 1e-3
        # the transient from there: the same outcome as the oracle's (these
        # gate arrays stall within a few ps in both: their NAND stacks leave
        # nodes held only by gmin and no capacitance)
        st_g, _ = b.instance(i, o.n)
        r = Oracle(c).tran(tstop, 1e-12, 1e-11)
        print(f"instance {i}: n={o.n}, DC {b.dc_stage(i)}; transient gpu mode {st_g['mode']:.0f} "
              f"steps {st_g['steps']:.0f}, oracle status {r['status']} steps {r['steps']}")
        assert (st_g["mode"] == 2) == (r["status"] == 0)
        assert abs(st_g["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
        if r["status"] == 0:
            assert waveform_error(t_g, x_g, r["t"], r["x"]) <= 1e-3 * 1.8 + 1e-6


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
