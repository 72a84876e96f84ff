"""xg_lu_solve_noalloc: the solve that allocates nothing, against xg_lu_solve_dev.

The entry point has no counterpart in the reference; it serves callers that
cannot allocate device memory.  Its work vectors live in one static work area
per GPU, shared by every call, so two calls in a row (the second on a smaller
system, which must not see the first one's leftovers) must each give the bits
xg_lu_solve_dev gives with the handle's own buffers.  Both factorisation
methods are covered; a NaN right-hand side gives the status and a zero
solution; FP32 factors are refused before any device work.
"""
import ctypes

import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _jacobian(c, seed):
    """A transient Jacobian and right-hand side of circuit c (oracle loads)."""
    o = Oracle(c)
    st, x0, sta, sto, Q, F, B = o.dcop()
    assert st > 0
    rng = np.random.default_rng(seed)
    x = x0 + rng.normal(0.0, 0.05, size=o.n)
    r = o.load(x, sta, sto, dcop=0, tran=1, newton_iter=1, order=1, h=1e-12)
    rhs, J = o.combine(r, np.asarray(Q), np.zeros(o.n))
    p = o.pattern()
    return o.n, p["rowptr"], p["colind"], J, rhs


def _solve_both(n, rp, ci, J, b, method):
    import torch
    dev = torch.device("cuda", 0)
    s = X.DirectSolver(n, rp, ci, J, method=method)
    assert s.numeric(J) == 0
    assert s.info()["method"] == method
    L = X.lib()
    bd = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    x_ref = torch.zeros(n, dtype=torch.float64, device=dev)
    x_na = torch.full((n,), 5.0, dtype=torch.float64, device=dev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    assert L.xg_lu_solve_dev(s.h, P(bd), P(x_ref), 0) == 0
    assert L.xg_lu_solve_noalloc(s.h, P(bd), P(x_na)) == 0
    return s, x_ref.cpu().numpy(), x_na.cpu().numpy()


@pytest.mark.parametrize("method", [1, 2])
def test_noalloc_solve_equals_the_allocating_solve(gpu, method):
    big = C.ring(201, seed=7, mc_seed=7) if method == 1 else C.power_grid(48, 48, seed=3, pad_pitch=16, n_sinks=32)
    small = C.ring(11, seed=8, mc_seed=8) if method == 1 else C.power_grid(12, 12, seed=4, pad_pitch=6, n_sinks=8)
    n1, rp1, ci1, J1, b1 = _jacobian(big, 1)
    n2, rp2, ci2, J2, b2 = _jacobian(small, 2)
    assert n2 < n1
    s1, ref1, na1 = _solve_both(n1, rp1, ci1, J1, b1, method)
    assert np.array_equal(na1, ref1)
    A1 = np.zeros((n1, n1))
    for r in range(n1):
        A1[r, ci1[rp1[r]:rp1[r + 1]]] += J1[rp1[r]:rp1[r + 1]]
    assert np.max(np.abs(A1 @ na1 - b1)) <= 1e-12 * np.max(np.abs(b1)) * 1e3
    # the second call, on a smaller system, right after the first
    s2, ref2, na2 = _solve_both(n2, rp2, ci2, J2, b2, method)
    assert np.array_equal(na2, ref2)
    # and the first handle again: nothing of the second call leaks into it
    import torch
    L = X.lib()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    bd = torch.from_numpy(b1).to("cuda")
    xd = torch.zeros(n1, dtype=torch.float64, device="cuda")
    assert L.xg_lu_solve_noalloc(s1.h, P(bd), P(xd)) == 0
    assert np.array_equal(xd.cpu().numpy(), ref1)


def test_noalloc_solve_nan_and_refusals(gpu):
    import torch
    n, rp, ci, J, b = _jacobian(C.ring(21, seed=9, mc_seed=9), 3)
    s = X.DirectSolver(n, rp, ci, J, method=1)
    assert s.numeric(J) == 0
    L = X.lib()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    bd = torch.from_numpy(b).to("cuda")
    bd[4] = float("nan")
    xd = torch.ones(n, dtype=torch.float64, device="cuda")
    assert L.xg_lu_solve_noalloc(s.h, P(bd), P(xd)) == -99
    assert torch.all(xd == 0)
    # FP32 factors: refused, the output untouched
    s.set_precision(1, 1e-12, 10)
    assert s.numeric(J) == 0
    xd.fill_(3.0)
    bd = torch.from_numpy(b).to("cuda")
    assert L.xg_lu_solve_noalloc(s.h, P(bd), P(xd)) == -1
    assert torch.all(xd == 3.0)
    assert L.xg_lu_solve_noalloc(None, P(bd), P(xd)) == -1
