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


def waveform_error_tube(t_a, x_a, t_b, x_b, dt):
    """Like waveform_error, but each point of run a may match run b anywhere
    within +-dt in time (the comparison SPICE regression tools make for
    oscillators, whose edges shift by a fraction of a step when the Newton
    path differs within its tolerances)."""
    t0 = max(t_a[0], t_b[0]) + dt; t1 = min(t_a[-1], t_b[-1]) - dt
    m = (t_a >= t0) & (t_a <= t1)
    err = 0.0
    shifts = np.linspace(-dt, dt, 17)
    for j in range(x_a.shape[1]):
        d = np.full(int(m.sum()), np.inf)
        for s in shifts:
            d = np.minimum(d, np.abs(x_a[m, j] - np.interp(t_a[m] + s, t_b, x_b[:, j])))
        err = max(err, float(np.max(d)) if d.size else 0.0)
    return err


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
    does not."""
    t, y, order = r["t"], r["x"], r["order"]
    sq, sf, sabs = y[:, 0], y[:, 1], y[:, 2]
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
