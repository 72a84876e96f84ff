"""Pins for three oracle functions that the load-level pins leave open (no GPU).

* Voltage-limiter correction vectors (``dFdxdVp``, ``dQdxdVp``;
  N_DEV_MOSFET1.C:4473-4541).  When the limiters move a junction voltage away
  from the one the solution vector implies, the device adds the linearised
  change of its currents and charges, J * (v_limited - v_original), so that
  the Newton residual describes the solution vector's own voltages.  The test
  rebuilds that product from the device's own dF/dx and dQ/dx stamps (which
  the finite-difference pins fix) and a node-voltage perturbation that carries
  the limiter's moves.  Two features of the reference are stated explicitly:
  the junction conductances enter without gmin, and the bulk row of the charge
  correction carries ``gcgb * (vbd - vbd_orig)`` where the stamp would give
  ``gcbd * (vbd - vbd_orig)`` (N_DEV_MOSFET1.C:4524-4527).
* Meyer back-averaging (N_DEV_MOSFET1.C:3304-3352).  The gate charges advance
  by (C(v_new)/2 + C(v_old)/2 + overlap) * dv: over a gate-voltage ramp every
  increment is the trapezoid integral of the Meyer capacitance, which the
  textbook pins on ``qmeyer`` fix, and the accumulated charge is the integral
  of the capacitance over the ramp.
* DampedNewton's convergence rule (N_NLS_DampedNewton.C:1191-1387,
  N_NLS_NLParams.C:101-111).  Through the driver hooks the test sees every
  load and solve of a transient; from the solution vectors and updates alone it
  decides, with the weights of ``updateWeights_`` (:297-337), after which
  iteration each Newton solve must stop (no device limited, max |dx_i| / w_i
  <= 0.33), and the oracle must stop exactly there.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, qmeyer
from xyce_amd import circuits as C

EPSOX = 3.9 * 8.854214871e-12        # N_DEV_MOSFET1.C:3962
GMIN = 1e-12


def _mos_circuit(mtype, vto, **mp):
    c = C.Circuit()
    nd, ng, ns, nb = c.node("d"), c.node("g"), c.node("s"), c.node("b")
    m = c.add_model(C.model_card(mtype, VTO=vto, KP=110e-6, **mp))
    c.add_mos(nd, ng, ns, nb, m, C.inst_card(W=2e-6, L=0.5e-6, AD=4e-12, AS=4e-12, PD=8e-6, PS=8e-6))
    for node in (nd, ng, ns, nb):
        c.add_vsrc(node, 0, c.add_wave(C.W_DC, [0.0]))
    return c, (nd, ng, ns, nb)


def _dense(o, p, vals):
    A = np.zeros((o.n, o.n))
    rp, ci = p["rowptr"], p["colind"]
    for r in range(o.n):
        for k in range(rp[r], rp[r + 1]):
            A[r, ci[k]] += vals[k]
    return A


# ---------------------------------------------------------------- limiter vectors
LIMIT_CASES = [
    # dtype, node voltages (d, g, s, b), previous store (vbs, vgs, vds), label
    (+1, (2.6, 3.2, 0.0, 1.1), (0.0, 0.2, 0.5), "nmos forward mode, source junction forward"),
    (+1, (-2.2, 3.71, 0.0, 0.23), (-0.36, 1.52, -0.24), "nmos reverse mode, drain junction forward"),
    (+1, (-1.3, 2.9, 0.0, 0.1), (0.0, 0.3, -1.0), "nmos reverse mode, drain junction limiter alone"),
    (-1, (-2.6, -3.2, 0.0, -1.1), (0.0, 0.2, 0.5), "pmos forward mode"),
    (-1, (2.2, -3.71, 0.0, -0.23), (-0.36, 1.52, -0.24), "pmos reverse mode"),
]


@pytest.mark.parametrize("dtype,volts,old,label", LIMIT_CASES, ids=[c[3] for c in LIMIT_CASES])
def test_limiter_vectors_are_the_stamps_times_the_limiters_moves(dtype, volts, old, label):
    vto = 0.7 * dtype
    c, nodes = _mos_circuit(dtype, vto, TOX=2e-8, CGSO=3e-10, CGDO=3e-10, CGBO=2e-10, CJ=2e-4, CJSW=1e-9,
                            MJ=0.45, GAMMA=0.4, LAMBDA=0.03)
    o = Oracle(c, gmin=GMIN)
    p = o.pattern()
    gid = [int(p["node_gid"][nd]) for nd in nodes]
    x = np.zeros(o.n)
    for g, v in zip(gid, volts):
        x[g] = v
    sta0, sto0 = (int(v) for v in p["mos_stasto"][0])
    sto = np.zeros(o.nsto)
    vbs0, vgs0, vds0 = old
    sto[sto0:sto0 + 6] = [vbs0 - vds0, vbs0, vgs0, vds0, vto, 0.0]     # the store holds dtype * Von
    sta = np.zeros(o.nsta)
    r = o.load(x, sta, sto, sta_next=sta, sto_next=sto, dcop=0, tran=1, newton_iter=1, order=1, h=1e-11)
    assert r["limited"] == 1, "the case must make a junction limiter act"

    # the limiter's moves, in the device's own (dtype-signed) branch voltages
    vd, vg, vs, vb = volts
    orig = dict(vbs=dtype * (vb - vs), vgs=dtype * (vg - vs), vds=dtype * (vd - vs))
    sn = r["sto_next"]
    lim = dict(vbd=sn[sto0 + 0], vbs=sn[sto0 + 1], vgs=sn[sto0 + 2], vds=sn[sto0 + 3])
    d_vbs, d_vgs, d_vds = (lim[k] - orig[k] for k in ("vbs", "vgs", "vds"))
    d_vbd = lim["vbd"] - (orig["vbs"] - orig["vds"])
    assert d_vbd == pytest.approx(d_vbs - d_vds, abs=1e-15)
    assert max(abs(d_vbs), abs(d_vbd)) > 1e-3
    # node perturbation with the source as reference: node voltages move by dtype * branch moves
    dV = np.zeros(o.n)
    dV[gid[0]], dV[gid[1]], dV[gid[3]] = dtype * d_vds, dtype * d_vgs, dtype * d_vbs

    JF = _dense(o, p, r["dFdx"])
    JQ = _dense(o, p, r["dQdx"])
    d, g, s, b = gid
    # F rows: the stamp's product, with the junctions' gmin taken out
    expect_F = JF @ dV
    expect_F[b] -= dtype * GMIN * (d_vbd + d_vbs)
    expect_F[d] += dtype * GMIN * d_vbd
    expect_F[s] += dtype * GMIN * d_vbs
    mag_F = np.abs(JF) @ np.abs(dV)
    for row in (b, d, s):
        assert abs(r["dFdxdVp"][row] - expect_F[row]) <= 1e-12 * mag_F[row] + 1e-30, (label, row)
    assert r["dFdxdVp"][g] == 0.0
    # Q rows: the stamp's product; the bulk row uses the gate-bulk capacitance
    # on the bulk-drain move where the stamp has the bulk-drain capacitance
    gcgb = -JQ[g, b]
    gcbd = -JQ[d, b]
    assert gcgb > 0 and gcbd > 0 and abs(gcgb - gcbd) > 1e-3 * max(gcgb, gcbd)
    expect_Q = JQ @ dV
    expect_Q[b] += dtype * (gcgb - gcbd) * d_vbd
    mag_Q = np.abs(JQ) @ np.abs(dV) + abs(gcgb - gcbd) * abs(d_vbd)
    for row in (g, b, d, s):
        assert abs(r["dQdxdVp"][row] - expect_Q[row]) <= 1e-12 * mag_Q[row] + 1e-40, (label, row)
    # the anomaly is real: the pure stamp product misses the bulk row by that term
    assert abs(r["dQdxdVp"][b] - (JQ @ dV)[b]) > 1e-6 * mag_Q[b]
    # without limiting the vectors vanish (up to the rounding of the branch
    # voltages' re-derivation, e.g. vds = vgs - (vgs - vds), which the
    # reference's exact comparison with the original voltages also sees)
    sto2 = sto.copy()
    sto2[sto0:sto0 + 6] = [orig["vbs"] - orig["vds"], orig["vbs"], orig["vgs"], orig["vds"], vto, 0.0]
    r2 = o.load(x, sta, sto2, sta_next=sta, sto_next=sto2, dcop=0, tran=1, newton_iter=1, order=1, h=1e-11)
    assert r2["limited"] == 0
    vmax = max(abs(v) for v in volts)
    assert np.max(np.abs(r2["dFdxdVp"])) <= 4e-16 * vmax * np.max(np.abs(JF))
    assert np.max(np.abs(r2["dQdxdVp"])) <= 4e-16 * vmax * np.max(np.abs(JQ))


# ---------------------------------------------------------------- Meyer back-averaging
def _meyer_full(vgs, vgd, vgb, vto, phi, cox, ovl):
    """Full Meyer capacitances (2 x qmeyer's half values) plus overlaps, for
    a device with GAMMA = 0 (Von = VTO) in forward mode."""
    vdsat = max(vgs - vto, 0.0)
    half = qmeyer(np.array([[vgs, vgd, vgb, vto, vdsat, phi, cox]]))[0]
    return 2 * half + np.asarray(ovl)


def test_meyer_charges_are_trapezoid_integrals_of_the_capacitance():
    vto, phi, tox, w, l = 0.7, 0.6, 2e-8, 2e-6, 0.5e-6
    cgso, cgdo, cgbo = 3e-10, 2e-10, 1e-10
    c, nodes = _mos_circuit(+1, vto, TOX=tox, CGSO=cgso, CGDO=cgdo, CGBO=cgbo, PHI=phi)
    o = Oracle(c, gmin=GMIN)
    p = o.pattern()
    gid = [int(p["node_gid"][nd]) for nd in nodes]
    sta0, sto0 = (int(v) for v in p["mos_stasto"][0])
    cox = EPSOX / tox * l * w
    ovl = (cgso * w, cgdo * w, cgbo * l)
    vd = 1.0
    ramp = np.linspace(-0.5, 1.9, 121)        # cutoff, saturation and linear regions
    x = np.zeros(o.n)
    x[gid[0]] = vd
    x[gid[1]] = ramp[0]
    sta = np.zeros(o.nsta)
    sto = np.zeros(o.nsto)
    r = o.load(x, sta, sto, dcop=1, tran=0, newton_iter=0)
    sta, sto = r["sta_next"], r["sto_next"]
    q = [sta[sta0:sta0 + 3].copy()]
    caps = [_meyer_full(ramp[0], ramp[0] - vd, ramp[0], vto, phi, cox, ovl)]
    # the operating point's charges are the capacitances times the voltages
    np.testing.assert_allclose(q[0], caps[0] * [ramp[0], ramp[0] - vd, ramp[0]], rtol=1e-13)
    for vg in ramp[1:]:
        x[gid[1]] = vg
        r = o.load(x, sta, sto, sta_next=sta, sto_next=sto, dcop=0, tran=1, newton_iter=0, order=1, h=1e-11)
        assert r["limited"] == 0
        sta, sto = r["sta_next"], r["sto_next"]
        q.append(sta[sta0:sta0 + 3].copy())
        caps.append(_meyer_full(vg, vg - vd, vg, vto, phi, cox, ovl))
    q, caps = np.array(q), np.array(caps)
    dv = np.diff(ramp)[:, None]             # vgs, vgd and vgb all move with the gate
    trap = 0.5 * (caps[1:] + caps[:-1]) * dv
    np.testing.assert_allclose(np.diff(q, axis=0), trap, rtol=1e-12, atol=1e-12 * np.max(np.abs(trap)))
    # the capacitances do change along the ramp, so a one-sided rule would differ
    right = caps[1:] * dv
    assert np.max(np.abs(right - trap)) > 1e-3 * np.max(np.abs(trap))
    # the accumulated charge is the integral of the capacitance over the ramp
    fine = np.linspace(ramp[0], ramp[-1], 24001)
    cf = np.array([_meyer_full(v, v - vd, v, vto, phi, cox, ovl) for v in fine])
    integral = np.trapezoid(cf, fine, axis=0)
    np.testing.assert_allclose(q[-1] - q[0], integral, rtol=2e-4)
    assert np.all(np.abs(right.sum(axis=0) - integral) > 2e-3 * np.abs(integral))


# ---------------------------------------------------------------- Newton convergence rule
def _record_run(c, tstop, tstep, hmax):
    """A transient whose loads and solves go through the driver hooks to a
    second oracle; returns the run and the per-call log."""
    inner = Oracle(c)
    drv = Oracle(c)
    n, nnz, nsta, nsto = inner.n, inner.nnz, inner.nsta, inner.nsto
    arr = lambda ptr, k: np.ctypeslib.as_array(ptr, shape=(max(k, 1),))
    p = inner.pattern()
    log = []

    def load(ctx, ssi, ssd, x, sc, oc, sn, on, F, Q, B, dFv, dQv, dFdx, dQdx):
        si, sd = arr(ssi, 7), arr(ssd, 4)
        xv = arr(x, n)[:n].copy()
        out = inner.load(xv, arr(sc, nsta)[:nsta].copy(), arr(oc, nsto)[:nsto].copy(),
                         sta_next=arr(sn, nsta)[:nsta].copy(), sto_next=arr(on, nsto)[:nsto].copy(),
                         dcop=int(si[0]), tran=int(si[1]), init_jct=int(si[2]), newton_iter=int(si[4]),
                         order=int(si[5]), h=float(sd[0]), time=float(sd[1]), tstep=float(sd[2]), tstop=float(sd[3]),
                         tranop=int(si[6]))
        arr(sn, nsta)[:nsta] = out["sta_next"][:nsta]
        arr(on, nsto)[:nsto] = out["sto_next"][:nsto]
        for name, ptr in (("F", F), ("Q", Q), ("B", B), ("dFdxdVp", dFv), ("dQdxdVp", dQv)):
            arr(ptr, n)[:n] = out[name]
        arr(dFdx, nnz)[:nnz] = out["dFdx"]
        arr(dQdx, nnz)[:nnz] = out["dQdx"]
        log.append(("load", int(si[0]), int(si[4]), float(sd[1]), xv, int(out["limited"])))
        return int(out["limited"])

    from oracle.pyoracle import lib as olib, _p

    def solve(ctx, nn, rp, ci, vals, b, x):
        xs = np.zeros(nn)
        st = olib().orc_lu_solve_sparse(nn, _p(p["rowptr"]), _p(p["colind"]),
                                        _p(arr(vals, nnz)[:nnz].copy()), _p(arr(b, nn)[:nn].copy()), _p(xs))
        arr(x, nn)[:nn] = xs
        log.append(("solve", xs.copy(), arr(b, nn)[:nn].copy()))
        return int(st)

    drv.set_hooks(load, solve)
    r = drv.tran(tstop, tstep, hmax)
    return r, log


def test_newton_stops_where_the_weighted_update_rule_says():
    c = C.ring(7, seed=3)
    r, log = _record_run(c, 0.6e-9, 1e-12, 1e-11)
    assert r["status"] == 0
    # split the transient's loads into Newton solves: a solve starts with a
    # load at newtonIter 0 and alternates solve, load
    solves, cur = [], None
    for e in log:
        if e[0] == "load" and e[1] == 1:
            continue                     # DC operating point
        if e[0] == "load" and e[2] == 0:
            cur = dict(time=e[3], x=[e[4]], lim=[e[5]], dx=[], rhs=[])
            solves.append(cur)
        elif e[0] == "load":
            cur["x"].append(e[4]); cur["lim"].append(e[5])
        elif cur is not None:
            cur["dx"].append(e[1]); cur["rhs"].append(e[2])
    # the accepted points, to find each solve's last accepted solution
    t_acc, x_acc = r["t"], r["x"]
    in_window, decided = 0, 0
    for k, s in enumerate(solves):
        h_idx = np.searchsorted(t_acc, s["time"] * (1 - 1e-12)) - 1
        xcurr = x_acc[h_idx]
        w = 1e-2 * np.maximum(np.abs(xcurr), np.abs(s["x"][0])) + 1e-6
        stop = None
        for it in range(1, len(s["x"])):
            wmax = np.max(np.abs(s["dx"][it - 1]) / w)
            if s["lim"][it]:
                continue
            if 0.33 < wmax <= 0.9:
                in_window += 1
            if wmax <= 0.33:
                stop = it
                break
        # every solve of this run ends by the update rule (none fails or stalls)
        assert stop is not None, (k, s["time"], len(s["x"]) - 1)
        decided += 1
        assert len(s["x"]) - 1 == stop, (k, s["time"], stop, len(s["x"]) - 1)
        # the residual test is met wherever the update test stops the solve
        if stop < len(s["rhs"]):
            assert np.max(np.abs(s["rhs"][stop])) <= 1e-2
    assert decided >= 100
    # iterations whose update lies between 0.33 and 0.9 of the weights exist,
    # so a different tolerance would change the iteration counts
    assert in_window >= 10
    assert sum(len(s["x"]) - 1 for s in solves) == r["newton_iters"]


# ---------------------------------------------------------------- tranop
def test_tranop_stamps_the_operating_point_capacitances():
    """At the operating point of a .TRAN (tranopFlag, N_DEV_SolverState.C:334)
    the reference stamps the gate capacitances into dQdx
    (N_DEV_MOSFET1.C:4829-4836); at a plain operating point they are zero.
    In saturation the DC Meyer values are the textbook ones: Cgs = 2/3 Cox W L
    plus overlap, Cgd and Cgb the overlaps alone."""
    vto, phi, tox, w, l = 0.7, 0.6, 2e-8, 2e-6, 0.5e-6
    cgso, cgdo, cgbo = 3e-10, 2e-10, 1e-10
    c, nodes = _mos_circuit(+1, vto, TOX=tox, CGSO=cgso, CGDO=cgdo, CGBO=cgbo, PHI=phi)
    o = Oracle(c, gmin=GMIN)
    p = o.pattern()
    d, g, s, b = (int(p["node_gid"][nd]) for nd in nodes)
    x = np.zeros(o.n)
    x[d], x[g] = 1.0, 1.5             # vgs - vto = 0.8 < vds: saturation
    sta, sto = np.zeros(o.nsta), np.zeros(o.nsto)
    cox = EPSOX / tox * w * l
    cgs, cgd, cgb = 2.0 / 3.0 * cox + cgso * w, cgdo * w, cgbo * l
    for tranop in (0, 1):
        r = o.load(x, sta, sto, dcop=1, tran=0, newton_iter=0, tranop=tranop)
        A = _dense(o, p, r["dQdx"])
        if not tranop:
            assert np.all(A == 0.0)
            continue
        np.testing.assert_allclose([A[g, g], A[g, s], A[g, d], A[g, b]], [cgs + cgd + cgb, -cgs, -cgd, -cgb],
                                   rtol=1e-12)
