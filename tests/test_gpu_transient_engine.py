"""The device-resident transient of one circuit (``xg_tran_*``) against the
oracle.

The engine restates the oracle's driver (DC operating point with its
strategy, DampedNewton, OneStep trapezoid with breakpoints and step control)
with every vector in device memory and the GPU's loads and LU.  Against the
all-CPU oracle on circuits it can run:

* a ring oscillator and an inverter chain with capacitances (nonlinear,
  limiting on), both LU methods: step and Newton counts within 2 %, node
  voltages within RELTOL |V| + ABSTOL while the step sequences agree;
* a power grid (linear, PWL sinks): the same step sequence and node voltages
  within 1e-9 V, and the summed node equations (charge balance) met at every
  accepted step, recorded through the engine's probe rows;
* a gate array whose DC point needs the DC strategy: the same strategy stage
  and operating point;
* the same gate array with UIC (no operating point, the run starts from
  x = 0): the oracle's step and Newton counts and its early waveform.
"""
import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C
from oracle.pyoracle import Oracle
from tests.helpers import grid_charge_balance, grid_probe, waveform_error

pytestmark = pytest.mark.gpu


def _early(tg, xg, r, t_end):
    mg, mo = tg <= t_end, r["t"] <= t_end
    return waveform_error(tg[mg], xg[mg], r["t"][mo], r["x"][mo])


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("make,tstop,hmax", [
    (lambda: C.ring(11, seed=4, mc_seed=4), 2e-9, 1e-11),
    (lambda: C.inverter_chain(5, TOX=2e-8, CJ=2e-4, CGSO=3e-10, CGDO=3e-10, CGBO=2e-10), 3e-9, 0.0),
], ids=["ring11", "chain5_caps"])
def test_engine_matches_oracle_nonlinear(gpu, make, tstop, hmax, method):
    c = make()
    o = Oracle(c)
    r = o.tran(tstop, 1e-12, hmax)
    assert r["status"] == 0
    e = X.Transient(c, tstop, 1e-12, hmax, lu_method=method)
    e.set_record(len(r["t"]) + 4000)
    st = e.run()
    assert st["finished"] == 1 and st["status"] == 0, st
    tg, og, xg = e.points(st["points"])
    print(f"{st}; oracle steps {r['steps']} Newton {r['newton_iters']}")
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    assert abs(st["newton_iters"] - r["newton_iters"]) <= 0.02 * r["newton_iters"] + 2
    # the DC point and the first stretch (same step sequence)
    np.testing.assert_allclose(xg[0], r["x"][0], rtol=0, atol=1e-9)
    n_same = 0
    while n_same < min(len(tg), len(r["t"])) and abs(tg[n_same] - r["t"][n_same]) <= 1e-12 * r["t"][n_same] + 1e-24:
        n_same += 1
    assert n_same >= 20
    err = np.max(np.abs(xg[:n_same] - r["x"][:n_same]))
    assert err <= 1e-3 * 1.8 + 1e-6, err
    assert tg[-1] == pytest.approx(tstop, rel=1e-12)


@pytest.mark.parametrize("method", [1, 2])
def test_engine_power_grid_matches_oracle_and_balances_charge(gpu, method):
    c = C.power_grid(12, 12, seed=5, pad_pitch=6, n_sinks=16, tstop=0.4e-9)
    o = Oracle(c)
    w = grid_probe(c, o, extra_nodes=[1, 2])
    o.set_probe(w)
    r = o.tran(0.4e-9, 1e-12, 1e-12)
    assert r["status"] == 0
    e = X.Transient(c, 0.4e-9, 1e-12, 1e-12, lu_method=method)
    e.set_record(len(r["t"]) + 100, w)
    st = e.run()
    assert st["finished"] == 1 and st["status"] == 0, st
    tg, og, yg = e.points(st["points"])
    assert st["steps"] == r["steps"] and len(tg) == len(r["t"])
    np.testing.assert_array_equal(og, r["order"])
    np.testing.assert_allclose(tg, r["t"], rtol=1e-13, atol=1e-24)
    np.testing.assert_allclose(yg[:, 3:], r["x"][:, 3:], rtol=0, atol=1e-9)
    res = dict(t=tg, x=yg, order=og)
    assert grid_charge_balance(c, o, res, 1e-12, 0.4e-9) <= 1.0


def test_engine_gate_array_dc_strategy(gpu):
    c = C.gate_dag(400, seed=1)
    o = Oracle(c)
    r = o.tran(0.3e-9, 1e-12, 1e-11)
    assert r["status"] == 0
    e = X.Transient(c, 0.3e-9, 1e-12, 1e-11)
    e.set_record(len(r["t"]) + 2000)
    st = e.run()
    assert st["finished"] == 1 and st["status"] == 0, st
    tg, og, xg = e.points(st["points"])
    print(f"gate array 400: {st}; oracle DC stage {o.dc_strategy()}, steps {r['steps']}")
    assert ("newton", "gmin", "source")[st["dc_stage"]] == o.dc_strategy()
    np.testing.assert_allclose(xg[0], r["x"][0], rtol=0, atol=1e-6)
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    assert _early(tg, xg, r, 0.1e-9) <= 1e-3 * 1.8 + 1e-6


def test_engine_gate_array_uic(gpu):
    c = C.gate_dag(400, seed=1)
    o = Oracle(c)
    o.set_uic()
    r = o.tran(0.3e-9, 1e-12, 1e-11)
    assert r["status"] == 0
    e = X.Transient(c, 0.3e-9, 1e-12, 1e-11)
    e.set_uic()
    e.set_record(len(r["t"]) + 2000)
    st = e.run()
    assert st["finished"] == 1 and st["status"] == 0, st
    tg, og, xg = e.points(st["points"])
    print(f"gate array 400, UIC: {st}; oracle steps {r['steps']}, Newton {r['newton_iters']}")
    assert st["dc_stage"] == -1 and st["dc_iters"] == 0
    assert np.all(xg[0] == 0.0)
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    assert abs(st["newton_iters"] - r["newton_iters"]) <= 0.02 * r["newton_iters"] + 2
    assert _early(tg, xg, r, 0.1e-9) <= 1e-3 * 1.8 + 1e-6
