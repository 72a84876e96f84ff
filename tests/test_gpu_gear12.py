"""Gear12 time integration (.OPTIONS TIMEINT METHOD=GEAR) in the batched
engine against the oracle's Gear12 (variable-step BDF1/BDF2, pinned on the
CPU by the BDF interpolation condition and RC closed forms).

Bars as for the trapezoid: step counts within 2 % (+2), node voltages within
RELTOL |V| + ABSTOL (1e-3 * 1.8 V + 1 uV) of the oracle's run.  A Gear run
and a trapezoid run of the same circuit differ (the method is switched), and
the method cannot change after the first launch.
"""
import numpy as np
import pytest

import xyce_amd as X
from oracle.pyoracle import Oracle
from tests.helpers import waveform_error
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu

CAPS = dict(TOX=2e-8, CJ=2e-4, CJSW=1e-9, CGSO=1e-10, CGDO=1e-10, CGBO=1e-10)
TOL = 1e-3 * 1.8 + 1e-6


def _oracle_gear(c, tstop, hmax):
    o = Oracle(c)
    o.set_method("gear")
    r = o.tran(tstop, 1e-12, hmax)
    assert r["status"] == 0
    return o, r


@pytest.mark.parametrize("make,tstop,hmax", [
    (lambda: C.ring(5, seed=5), 2e-9, 1e-11),
    (lambda: C.ring(11, seed=11), 2e-9, 1e-11),
    (lambda: C.inverter_chain(4, **CAPS), 4e-9, 1e-11),
    (lambda: C.power_grid(6, 6, seed=21, pad_pitch=3, n_sinks=4, tstop=1e-9), 1e-9, 0.0),
])
def test_gear_batch_matches_oracle(gpu, make, tstop, hmax):
    c = make()
    b = X.Batch([c], tstop, 1e-12, hmax, record_max=20000, method="gear")
    cnt = b.run_to_end(chunk=512)
    assert cnt["done"] == 1 and cnt["failed"] == 0, cnt
    o, r = _oracle_gear(c, tstop, hmax)
    t_g, x_g = b.points(0, o.n)
    st, _ = b.instance(0, o.n)
    err = waveform_error(t_g, x_g, r["t"], r["x"])
    print(f"gear: steps gpu {st['steps']} oracle {r['steps']}, rejects oracle {r['rejects']}, "
          f"max waveform diff {err:.3e} V")
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    assert err <= TOL, err


def test_gear_mixed_batch_and_method_switch(gpu):
    """Monte Carlo rings of uneven size under Gear: each instance equals its
    own single-instance run; the trapezoid run of the same batch differs;
    the method is fixed once a launch has run."""
    cs = [C.ring(s, seed=40 + s, mc_seed=40 + s) for s in (11, 31, 7)]
    tstop = 1e-9
    b = X.Batch(cs, tstop, 1e-12, 1e-11, record_max=4000, method="gear")
    cnt = b.run_to_end(chunk=256)
    assert cnt["done"] == 3 and cnt["failed"] == 0, cnt
    bt = X.Batch(cs, tstop, 1e-12, 1e-11, record_max=4000)
    bt.run_to_end(chunk=256)
    for i, c in enumerate(cs):
        n = X.host_pattern(c)["n"]
        b1 = X.Batch([c], tstop, 1e-12, 1e-11, record_max=4000, method="gear")
        b1.run_to_end(chunk=256)
        t_a, x_a = b.points(i, n)
        t_b, x_b = b1.points(0, n)
        assert np.array_equal(t_a, t_b) and np.array_equal(x_a, x_b)
        t_t, x_t = bt.points(i, n)
        assert len(t_t) != len(t_a) or not np.array_equal(x_t, x_a)
    assert X.lib().xg_batch_set_method(b.h, 0) < 0


def test_gear_netlist_option(gpu):
    """.OPTIONS TIMEINT METHOD=GEAR in a netlist runs the Gear integrator."""
    from xyce_amd.netlist import parse_netlist, run_transient
    body = """inverter
V1 vdd 0 DC 1.8
V2 in 0 PULSE(0 1.8 0.2n 50p 50p 1n 3n)
M1 out in 0 0 nch W=2u L=0.5u
M2 out in vdd vdd pch W=4u L=0.5u
C1 out 0 5f
.model nch NMOS LEVEL=1 VTO=0.7 KP=110u
.model pch PMOS LEVEL=1 VTO=-0.7 KP=50u
.OPTIONS TIMEINT METHOD=GEAR
.TRAN 1p 3n 0 10p
.PRINT TRAN V(out)
.END
"""
    t, cols = run_transient(body)
    c, an = parse_netlist(body)
    assert an["method"] == "gear"
    o, r = _oracle_gear(c, 3e-9, 10e-12)
    col = o.pattern()["node_gid"][an["nodes"]["OUT"]]
    err = waveform_error(t, cols["V(OUT)"][:, None], r["t"], r["x"][:, [col]])
    assert err <= TOL, err
