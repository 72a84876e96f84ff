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
