"""Shared helpers for the parity tests."""
import numpy as np


def entry_tol(a, b, scale, rel=1e-12, abs_floor=0.0):
    """max |a-b| / (rel * scale + abs_floor) — the per-entry parity measure.

    ``scale`` is the magnitude the entry was summed from (sum of |terms| or
    max(|a|,|b|)); a value <= 1 passes."""
    d = np.abs(np.asarray(a) - np.asarray(b))
    den = rel * np.asarray(scale) + abs_floor
    den = np.where(den > 0, den, 1e-300)
    return float(np.max(d / den)) if d.size else 0.0


def random_ring_state(o, rng, vdd=1.8):
    """A plausible non-equilibrium state of a ring oscillator: node voltages
    in [0, VDD], VDD node at VDD, branch current small, store/state near it."""
    x = rng.uniform(0.0, vdd, size=o.n)
    return x


def waveform_error(t_a, x_a, t_b, x_b):
    """max over the common time range of |x_a - interp(x_b)| per node."""
    t0 = max(t_a[0], t_b[0]); t1 = min(t_a[-1], t_b[-1])
    m = (t_a >= t0) & (t_a <= t1)
    err = 0.0
    for j in range(x_a.shape[1]):
        xb = np.interp(t_a[m], t_b, x_b[:, j])
        err = max(err, float(np.max(np.abs(x_a[m, j] - xb))))
    return err


def waveform_error_tube(t_a, x_a, t_b, x_b, dt, n_shifts=17):
    """Like waveform_error, but each point of run a may match run b anywhere
    within +-dt in time (the comparison SPICE regression tools make for
    oscillators, whose edges shift by a fraction of a step when the Newton
    path differs within its tolerances); the shifts tried are ``n_shifts``
    evenly spaced values in [-dt, dt]."""
    t0 = max(t_a[0], t_b[0]) + dt; t1 = min(t_a[-1], t_b[-1]) - dt
    m = (t_a >= t0) & (t_a <= t1)
    err = 0.0
    shifts = np.linspace(-dt, dt, n_shifts)
    for j in range(x_a.shape[1]):
        d = np.full(int(m.sum()), np.inf)
        for s in shifts:
            d = np.minimum(d, np.abs(x_a[m, j] - np.interp(t_a[m] + s, t_b, x_b[:, j])))
        err = max(err, float(np.max(d)) if d.size else 0.0)
    return err


def _quad_interp(tb, xb, ta):
    """Quadratic (three-point Lagrange) interpolation of the rows of xb at ta,
    through the sample before and the two after each interval start."""
    idx = np.clip(np.searchsorted(tb, ta), 1, len(tb) - 2)
    t0, t1, t2 = tb[idx - 1], tb[idx], tb[idx + 1]
    l0 = (ta - t1) * (ta - t2) / ((t0 - t1) * (t0 - t2))
    l1 = (ta - t0) * (ta - t2) / ((t1 - t0) * (t1 - t2))
    l2 = (ta - t0) * (ta - t1) / ((t2 - t0) * (t2 - t1))
    return l0[:, None] * xb[idx - 1] + l1[:, None] * xb[idx] + l2[:, None] * xb[idx + 1]


def waveform_error_tube_interp(t_a, x_a, t_b, x_b, dt, n_shifts=129):
    """Largest distance of run a's points from run b's waveform when each
    point may match b anywhere within +-dt in time, not counting the
    interpolation error of b's piecewise-linear waveform between its own
    samples: at each time that error is estimated by the distance between
    b's linear and quadratic interpolants, and subtracted.  Two runs with
    different step sequences can only be compared through an interpolant; a
    run compared with itself at shifted sample times gives about 0."""
    t0 = max(t_a[0], t_b[1]) + dt
    t1 = min(t_a[-1], t_b[-3]) - dt
    m = (t_a >= t0) & (t_a <= t1)
    ta, a = t_a[m], x_a[m]
    best = np.full(a.shape, np.inf)
    for s in np.linspace(-dt, dt, n_shifts) if dt > 0 else (0.0,):
        lin = np.stack([np.interp(ta + s, t_b, x_b[:, j]) for j in range(x_b.shape[1])], 1)
        q = _quad_interp(t_b, x_b, ta + s)
        best = np.minimum(best, np.maximum(0.0, np.abs(a - lin) - np.abs(q - lin)))
    return float(best.max()) if best.size else 0.0


def edge_times(t, v, level=0.9):
    """Times at which ``v`` crosses ``level`` (linear interpolation between
    samples) and the direction of each crossing (+1 rising, -1 falling)."""
    s = np.asarray(v) - level
    k = np.nonzero((s[:-1] < 0) != (s[1:] < 0))[0]
    tc = t[k] - s[k] * (t[k + 1] - t[k]) / (s[k + 1] - s[k])
    return tc, np.sign(s[k + 1])


def edge_shift(t_a, x_a, t_b, x_b, level=0.9):
    """Largest time shift between corresponding crossings of ``level`` in two
    runs, over every column; the runs must cross the same number of times in
    the same directions (else +inf)."""
    worst = 0.0
    for j in range(x_a.shape[1]):
        ea, da = edge_times(t_a, x_a[:, j], level)
        eb, db = edge_times(t_b, x_b[:, j], level)
        if len(ea) != len(eb) or not np.array_equal(da, db):
            return float("inf")
        if len(ea):
            worst = max(worst, float(np.max(np.abs(ea - eb))))
    return worst


def grid_probe(c, o, extra_nodes=()):
    """Probe rows (for ``Oracle.set_probe``) of a network of resistors,
    grounded capacitors, grounded voltage sources and current sources:
    row 0 sums the node charges (sum of C_i v_i), row 1 the node rows of F
    (resistor currents to ground plus voltage-source branch currents: the
    currents between grid nodes cancel), row 2 the magnitude those rows are
    summed from (sum over node rows of sum_j |G_ij| v_j, node voltages being
    positive here), then one row per node in ``extra_nodes`` (its voltage)."""
    p = o.pattern()
    gid = p["node_gid"]
    n = o.n
    assert all(b == 0 for _, b, _ in c.vsrc), "voltage sources must have their negative terminal at ground"
    node_cols = set(int(gid[k]) for k in range(1, c.n_nodes))
    branch_cols = [j for j in range(n) if j not in node_cols]
    assert len(branch_cols) == len(c.vsrc)
    w = np.zeros((3 + len(extra_nodes), n))
    for a, b, cap in c.cap:
        if b == 0 and a != 0:
            w[0, gid[a]] += cap
        elif a == 0 and b != 0:
            w[0, gid[b]] += cap
    for a, b, r in c.res:
        for u, v in ((a, b), (b, a)):
            if u == 0:
                continue
            w[2, gid[u]] += 1.0 / r if v == 0 else 2.0 / r
            if v == 0:
                w[1, gid[u]] += 1.0 / r
    w[1, branch_cols] = 1.0
    for k, nd in enumerate(extra_nodes):
        w[3 + k, gid[nd]] = 1.0
    return w


def _pwl_source_sums(c, t):
    """(S_B, sum |I|) at times t for a circuit whose current sources are all
    PWL and grounded on one side: the node-row sum of the sources' B entries
    (-I for a current drawn from a node to ground) and the magnitude it is
    summed from; the waveforms evaluated by linear interpolation between
    their points, held constant outside them (the PWL definition)."""
    from xyce_amd import circuits as C
    sb = np.zeros(len(t))
    ssum = np.zeros(len(t))
    for a, b, wv in c.isrc:
        wt, p = c.waves[wv]
        assert wt == C.W_PWL
        cur = np.interp(t, np.asarray(p[0::2]), np.asarray(p[1::2]))
        sgn = (-1.0 if a != 0 else 0.0) + (1.0 if b != 0 else 0.0)
        sb += sgn * cur
        ssum += np.abs(cur)
    return sb, ssum


def grid_charge_balance(c, o, r, tstep, tstop):
    """Largest defect of the summed node equations over the accepted steps
    of a run recorded with ``grid_probe`` rows, relative to what floating
    point allows: the node rows of the discretised equations summed give
      order 1: S_Q(n+1) - S_Q(n) + h (S_F - S_B)(n+1) = 0,
      order 2: S_Q(n+1) - S_Q(n) + h/2 [(S_F - S_B)(n+1) + (S_F - S_B)(n)] = 0,
    with S_B(t) the sources' node-row sum (-I of each current source drawn
    from a node to ground).  The allowance per step is h (1e-9 sum|I| +
    64 eps S_abs) + 1e-9 |dS_Q| + 4 eps |S_Q|: the rows are solved to round-off
    of the conductance terms they sum (S_abs), and the probes are compensated
    sums, exact to a few units of their last place.  A converged linear network stays
    within it (value <= 1); a wrong load, solve or integration coefficient
    does not.  With ``o`` None the source sum is evaluated in numpy (PWL
    sources only), for runs too long for one oracle call per point."""
    t, y, order = r["t"], r["x"], r["order"]
    sq, sf, sabs = y[:, 0], y[:, 1], y[:, 2]
    if o is None:
        sb, ssum = _pwl_source_sums(c, t)
    else:
        sb = np.zeros(len(t))
        ssum = np.zeros(len(t))
        for a, b, wv in c.isrc:
            cur = np.array([o.wave(wv, tk, tstep, tstop) for tk in t])
            sgn = (-1.0 if a != 0 else 0.0) + (1.0 if b != 0 else 0.0)
            sb += sgn * cur
            ssum += np.abs(cur)
    g = sf - sb
    eps = np.finfo(float).eps
    worst = 0.0
    for k in range(1, len(t)):
        h = t[k] - t[k - 1]
        if order[k] == 1:
            d = sq[k] - sq[k - 1] + h * g[k]
            allow = h * (1e-9 * ssum[k] + 64 * eps * sabs[k])
        else:
            d = sq[k] - sq[k - 1] + 0.5 * h * (g[k] + g[k - 1])
            allow = h * (1e-9 * max(ssum[k], ssum[k - 1]) + 64 * eps * max(sabs[k], sabs[k - 1]))
        allow += 1e-9 * abs(sq[k] - sq[k - 1]) + 4 * eps * abs(sq[k]) + 1e-300
        worst = max(worst, abs(d) / allow)
    return worst


# Level-1 model cards of the device zoo: capacitances on (Meyer, overlap,
# bottom and sidewall junctions), each depletion-charge form of
# updateIntermediateVars: MJ = MJSW = 0.5 (the square-root path), MJ = MJSW
# (one power), MJ != MJSW (two powers; the default MJSW 0.33)
ZOO_CAPS = dict(TOX=2e-8, CJ=2e-4, CJSW=1e-9, CGSO=3e-10, CGDO=3e-10, CGBO=2e-10, GAMMA=0.4, LAMBDA=0.03)
ZOO_CARDS = ({}, dict(MJ=0.5, MJSW=0.5), dict(MJ=0.4, MJSW=0.4), dict(MJ=0.45, MJSW=0.3))


def device_zoo(n_per=25):
    """NMOS and PMOS devices of every ZOO_CARDS card, each with its own four
    terminal nodes (1 kOhm to ground each), so that a load at arbitrary node
    voltages puts devices in every operating branch."""
    from xyce_amd import circuits as C
    c = C.Circuit()
    cards = []
    for extra in ZOO_CARDS:
        for t, base in ((+1, C.NMOS_RING), (-1, C.PMOS_RING)):
            cards.append(c.add_model(C.model_card(t, **base, **{**ZOO_CAPS, **extra})))
    for m in cards:
        for _ in range(n_per):
            k = len(c.mos_model)
            nd = [c.node(f"m{k}_{j}") for j in "dgsb"]
            c.add_mos(nd[0], nd[1], nd[2], nd[3], m,
                      C.inst_card(W=2e-6, L=0.5e-6, AD=4e-12, AS=4e-12, PD=8e-6, PS=8e-6))
            for j in nd:
                c.add_res(j, 0, 1e3)
    return c


def zoo_states(o, seed):
    """A DC load at random node voltages supplies the store/state history;
    returns (x, sta, sto) with x that history's voltages plus noise."""
    rng = np.random.default_rng(seed)
    xp = rng.uniform(-0.5, 2.3, size=o.n)
    r0 = o.load(xp, np.zeros(o.nsta), np.zeros(o.nsto), dcop=1, tran=0, init_jct=0, newton_iter=1, order=1)
    return xp + rng.normal(0.0, 0.3, size=o.n), r0["sta_next"].copy(), r0["sto_next"].copy()


def branch_classes(c, mos_stasto, sto, tdep=0.4):
    """Operating branches the devices of a load went through, read from the
    store values it wrote (limited vbd, vbs, vgs, vds and von, as
    updateIntermediateVars uses them): (type, mode, cutoff|sat|lin) and
    (bs|bd, type, reverse|forward below FC*PB|forward beyond FC*PB)."""
    out = set()
    for d in range(len(c.mos_model)):
        s0 = mos_stasto[d, 1]
        vbd, vbs, vgs, vds, von_s = sto[s0:s0 + 5]
        t = int(c.models[c.mos_model[d]][0])
        mode = 1 if vds >= 0 else -1
        vgst = (vgs if mode == 1 else vgs - vds) - t * von_s
        out.add(("m", t, mode, "cutoff" if vgst <= 0 else ("sat" if vgst <= vds * mode else "lin")))
        for nm, v in (("bs", vbs), ("bd", vbd)):
            out.add((nm, t, "rev" if v <= 0 else ("fwd<fc" if v < tdep else "fwd>=fc")))
    return out


ALL_BRANCHES = ({("m", t, m, r) for t in (1, -1) for m in (1, -1) for r in ("cutoff", "sat", "lin")}
                | {(j, t, b) for t in (1, -1) for j in ("bs", "bd") for b in ("rev", "fwd<fc", "fwd>=fc")})
