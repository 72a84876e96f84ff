"""Netlist subset reader and .PRN writer (host side, no GPU).

* SPICE numbers with scale suffixes;
* a hand-written netlist (continuations, comments, units, .MODEL after its
  devices) builds the devices it lists, and its DC operating point is the
  closed-form one;
* generated circuits written as Xyce netlists (``Circuit.to_spice``) and read
  back give the same devices, models and sources node for node (by name), and
  the same oracle DC operating point;
* unsupported input raises NetlistError naming the line;
* .PRN files round-trip.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from xyce_amd import circuits as C
from xyce_amd.netlist import NetlistError, parse_netlist, read_prn, value, write_prn


def test_spice_numbers():
    assert value("1.5k") == 1500.0
    assert value("2u") == 2e-6
    assert math.isclose(value("10fF"), 1e-14)
    assert value("3MEG") == 3e6
    assert value("1e-12") == 1e-12
    assert math.isclose(value("5mil"), 5 * 25.4e-6)
    assert value("1m") == 1e-3
    assert value("-.5") == -0.5
    with pytest.raises(NetlistError):
        value("abc")


HAND = """divider and inverter
* a comment line
V1 vdd 0 DC 1.8
R1 vdd mid 1k ; end-of-line comment
R2 mid 0
+ 2k
M1 out mid 0 0 nch W=2u L=0.5u
M2 out mid vdd vdd pch W=4u
+ L=0.5u
C1 out 0 5f
I1 0 out PULSE(0 10u 1n 10p 10p 1n 4n)
.model nch NMOS (LEVEL=1 VTO=0.7 KP=110u)
.model pch PMOS LEVEL=1 VTO=-0.7 KP=50u
.TRAN 1p 5n 0 10p
.PRINT TRAN V(out) V(mid)
.END
"""


def test_hand_written_netlist():
    c, an = parse_netlist(HAND)
    assert an["title"] == "divider and inverter"
    assert an["tran"] == (1e-12, 5e-9, 0.0, 10e-12)
    assert an["print"] == ["out", "mid"]
    assert [t for t, _ in c.order] == [C.VSRC, C.RES, C.RES, C.MOS1, C.MOS1, C.CAP, C.ISRC]
    assert c.res == [(1, 2, 1000.0), (2, 0, 2000.0)]
    assert len(c.models) == 2 and c.models[c.mos_model[0]][0] == 1 and c.models[c.mos_model[1]][0] == -1
    assert c.mos_inst[1][0] == 0.5e-6 and c.mos_inst[1][1] == 4e-6
    assert c.waves[c.isrc[0][2]][0] == C.W_PULSE
    o = Oracle(c)
    st, x, *_ = o.dcop()
    assert st > 0
    gid = o.pattern()["node_gid"]
    assert abs(x[gid[an["nodes"]["MID"]]] - 1.2) <= 1e-9     # divider 1.8 * 2k / 3k (gate draws nothing)


@pytest.mark.parametrize("maker", [lambda: C.ring(11, seed=3, mc_seed=3),
                                   lambda: C.inverter_chain(3, TOX=2e-8, CJ=2e-4, CGSO=1e-10),
                                   lambda: C.power_grid(4, 4, seed=2, pad_pitch=2, n_sinks=3, tstop=1e-9),
                                   lambda: C.gate_dag(40, seed=1)])
def test_generated_netlists_round_trip(maker):
    c = maker()
    c2, an = parse_netlist(c.to_spice(tran=(1e-12, 1e-9)))
    assert an["tran"][:2] == (1e-12, 1e-9)
    name = lambda circ, i: "0" if i == 0 else circ.names.get(i, f"n{i}").upper()
    assert [t for t, _ in c.order] == [t for t, _ in c2.order]
    for (t, i), (_, j) in zip(c.order, c2.order):
        if t == C.MOS1:
            assert [name(c, v) for v in c.mos_nodes[i]] == [name(c2, v) for v in c2.mos_nodes[j]]
            m1 = c.models[c.mos_model[i]].copy()
            inst = c.mos_inst[i]
            if not math.isnan(inst[9]):
                m1[1] = inst[9]
            if not math.isnan(inst[10]):
                m1[2] = inst[10]
            np.testing.assert_array_equal(np.nan_to_num(m1, nan=-7.0),
                                          np.nan_to_num(c2.models[c2.mos_model[j]], nan=-7.0))
            np.testing.assert_array_equal(np.nan_to_num(inst[:9], nan=-7.0), np.nan_to_num(c2.mos_inst[j][:9], nan=-7.0))
        else:
            a = {C.RES: c.res, C.CAP: c.cap, C.VSRC: c.vsrc, C.ISRC: c.isrc}[t][i]
            b = {C.RES: c2.res, C.CAP: c2.cap, C.VSRC: c2.vsrc, C.ISRC: c2.isrc}[t][j]
            assert name(c, a[0]) == name(c2, b[0]) and name(c, a[1]) == name(c2, b[1])
            if t in (C.RES, C.CAP):
                assert a[2] == b[2]
            else:
                wa, wb = c.waves[a[2]], c2.waves[b[2]]
                assert wa[0] == wb[0]
                np.testing.assert_array_equal(np.nan_to_num(wa[1], nan=-7.0), np.nan_to_num(wb[1], nan=-7.0))
    # same operating point, node by node
    o1, o2 = Oracle(c), Oracle(c2)
    s1, x1, *_ = o1.dcop()
    s2, x2, *_ = o2.dcop()
    assert s1 > 0 and s2 > 0
    g1, g2 = o1.pattern()["node_gid"], o2.pattern()["node_gid"]
    by_name = {name(c2, k): g2[k] for k in range(1, c2.n_nodes)}
    for k in range(1, c.n_nodes):
        assert abs(x1[g1[k]] - x2[by_name[name(c, k)]]) <= 1e-9


@pytest.mark.parametrize("bad", ["t\nQ1 a b c qmod\n", "t\n.model q NPN\n", "t\n.model m NMOS LEVEL=2\n",
                                 "t\nM1 a b c d nomodel\n.END\n", "t\nV1 a 0 SIN(0 1 1k)\n", "t\n.AC DEC 10 1 1k\n"])
def test_unsupported_input_is_rejected(bad):
    with pytest.raises(NetlistError):
        parse_netlist(bad)


def test_prn_round_trip(tmp_path):
    t = np.linspace(0, 1e-9, 11)
    cols = {"V(OUT)": np.sin(t * 1e9), "V(MID)": np.cos(t * 1e9)}
    p = tmp_path / "x.prn"
    write_prn(p, t, cols)
    text = p.read_text()
    assert text.splitlines()[0].split() == ["Index", "TIME", "V(OUT)", "V(MID)"]
    assert text.rstrip().endswith("End of Xyce(TM) Simulation")
    hdr, data = read_prn(p)
    assert hdr == ["Index", "TIME", "V(OUT)", "V(MID)"]
    np.testing.assert_array_equal(data[:, 0], np.arange(11))
    np.testing.assert_allclose(data[:, 1], t, rtol=1e-8, atol=1e-30)
    np.testing.assert_allclose(data[:, 2], cols["V(OUT)"], rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("line,method", [(".OPTIONS TIMEINT METHOD=GEAR", "gear"), (".options timeint method=8", "gear"),
                                         (".OPTIONS TIMEINT METHOD=TRAP", "trap"), (".OPTIONS TIMEINT RELTOL=1e-3", "trap"),
                                         (".OPTIONS DEVICE GMIN=1e-12", "trap")])
def test_integration_method_option(line, method):
    _, an = parse_netlist("t\nR1 a 0 1k\nV1 a 0 DC 1\n" + line + "\n.TRAN 1p 1n\n.END\n")
    assert an["method"] == method


def test_unknown_integration_method_is_rejected():
    with pytest.raises(NetlistError):
        parse_netlist("t\nR1 a 0 1k\n.OPTIONS TIMEINT METHOD=BDF5\n.END\n")


@pytest.mark.parametrize("line", [".OPTIONS DEVICE VOLTLIM=0", ".OPTIONS DEVICE GMIN=1e-9", ".OPTIONS TIMEINT RELTOL=1e-4",
                                  ".OPTIONS NONLIN MAXSTEP=50", ".OPTIONS LINSOL TYPE=SUPERLU", ".OPTIONS OUTPUT INITIAL_INTERVAL=1p",
                                  ".OPTIONS"])
def test_options_the_path_would_not_honour_are_refused(line):
    with pytest.raises(NetlistError):
        parse_netlist("t\nR1 a 0 1k\nV1 a 0 DC 1\n" + line + "\n.TRAN 1p 1n\n.END\n")


def test_default_options_may_be_restated():
    _, an = parse_netlist("t\nR1 a 0 1k\nV1 a 0 DC 1\n.OPTIONS DEVICE VOLTLIM=1 GMIN=1e-12\n"
                          ".OPTIONS LINSOL TYPE=klu\n.OPTIONS TIMEINT METHOD=GEAR ABSTOL=1u RELTOL=1e-3\n.TRAN 1p 1n\n.END\n")
    assert an["method"] == "gear"


@pytest.mark.parametrize("tail", ["OFF", "L=0.5u W=2u OFF", "IC=1,2,0"])
def test_mosfet_off_and_initial_conditions_are_refused(tail):
    text = ("t\nVDD vdd 0 DC 1.8\nM1 out in 0 0 NM " + tail + "\n.MODEL NM NMOS (LEVEL=1 VTO=0.7 KP=110u)\n"
            "R1 vdd out 1k\nV2 in 0 DC 1\n.TRAN 1p 1n\n.END\n")
    with pytest.raises(NetlistError, match="OFF|IC"):
        parse_netlist(text)
