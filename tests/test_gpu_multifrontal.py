"""Multifrontal LU on the GPU (xg_lu_symbolic_ex method 2) against SuperLU
(scipy) and the level-scheduled path, through the C ABI.

Bars: relative residual max|A x - b| / max|b| <= 1e-12 (FP64 factors, and
FP32 factors with FP64 refinement); solutions equal to SuperLU's and to the
level-scheduled path's within 1e-9 relative; a numerically singular matrix
returns its pivot status and a zero solution.  The 512 x 512 power-grid
mesh (262,208 unknowns) is factored at full size and checked through its
residual, against SuperLU's solution, and for repeatability of a refactor.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import xyce_amd as X
from oracle.pyoracle import Oracle
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu


def _jacobian(c, seed, h=2e-12):
    """Transient Jacobian and a residual-like right-hand side of a circuit at a
    random state near mid-rail (oracle loads)."""
    o = Oracle(c)
    rng = np.random.default_rng(seed)
    x = 0.9 + rng.normal(0.0, 0.05, size=o.n)
    sta0, sto0 = np.zeros(max(o.nsta, 1)), np.zeros(max(o.nsto, 1))
    ref = o.load(x, sta0, sto0, h=h, order=2)
    rhs, J = o.combine(ref, ref["Q"], ref["F"] - ref["B"])
    p = o.pattern()
    return o.n, p["rowptr"], p["colind"], J, rhs


def _check(n, rp, ci, J, b, tol=1e-9):
    A = sp.csr_matrix((J, ci, rp), shape=(n, n))
    xs = spla.spsolve(A.tocsc(), b)
    mf = X.DirectSolver(n, rp, ci, method=2)
    assert mf.info()["method"] == 2
    st, x = mf.doSolve(J, b)
    assert st == 0
    res = np.max(np.abs(A @ x - b)) / np.max(np.abs(b))
    err = np.max(np.abs(x - xs)) / np.max(np.abs(xs))
    print(f"n={n} {mf.info()}: residual {res:.2e}, vs SuperLU {err:.2e}")
    assert res <= 1e-12
    assert err <= tol
    return A, x


@pytest.mark.parametrize("name", ["ring101", "grid48", "dag3000"])
def test_multifrontal_matches_superlu_and_level_path(gpu, name):
    c = {"ring101": lambda: C.ring(101, seed=1),
         "grid48": lambda: C.power_grid(48, 48, seed=2, pad_pitch=16, n_sinks=64, tstop=1e-9),
         "dag3000": lambda: C.gate_dag(3000, seed=4)}[name]()
    n, rp, ci, J, b = _jacobian(c, 3)
    A, x = _check(n, rp, ci, J, b)
    lv = X.DirectSolver(n, rp, ci, method=1)
    st, xl = lv.doSolve(J, b)
    assert st == 0
    assert np.max(np.abs(x - xl)) <= 1e-9 * np.max(np.abs(xl))


def test_multifrontal_random_matrices(gpu):
    rng = np.random.default_rng(5)
    n = 3000
    rows, cols = [], []
    for i in range(n):
        for j in rng.choice(n, 3, replace=False):
            rows += [i, j]; cols += [j, i]
    A = sp.csr_matrix((rng.normal(size=len(rows)), (rows, cols)), shape=(n, n)) + sp.eye(n) * 8.0
    A.sum_duplicates(); A.sort_indices()
    b = rng.normal(size=n)
    _check(n, A.indptr, A.indices, A.data, b)


def test_multifrontal_singular_returns_status(gpu):
    rp = np.array([0, 2, 4, 6], np.int32)
    ci = np.array([0, 1, 0, 1, 1, 2], np.int32)
    v = np.array([1.0, 2.0, 2.0, 4.0, 1.0, 1.0])
    s = X.DirectSolver(3, rp, ci, method=2)
    st, x = s.doSolve(v, np.ones(3))
    assert st > 0 and np.all(x == 0.0)


def test_multifrontal_fp32_refined(gpu):
    c = C.power_grid(128, 128, seed=2, pad_pitch=64, n_sinks=256, tstop=1e-9)
    n, rp, ci, J, b = _jacobian(c, 4, h=1e-12)
    s = X.DirectSolver(n, rp, ci, values=J, method=2)
    s.set_precision(1, ir_tol=1e-12, ir_max=10)
    st, x = s.doSolve(J, b)
    assert st == 0
    ps = s.precision_stats()
    A = sp.csr_matrix((J, ci, rp), shape=(n, n))
    res = np.max(np.abs(A @ x - b)) / np.max(np.abs(b))
    print(f"n={n}: {ps}, residual {res:.2e}")
    assert res <= 1e-12 and ps["fp64_fallbacks"] == 0 and ps["fp32_pivot_failures"] == 0


def test_power_grid_512_full_size(gpu):
    """512 x 512 mesh, 64 pads, 4096 PWL sinks: the transient Jacobian
    G + C/h (trapezoid, h = 1 ps) factored and solved at full size."""
    c = C.power_grid(512, 512, seed=2, pad_pitch=64, n_sinks=4096, tstop=2e-9)
    dm = X.DeviceMgr(c)
    n = dm.n
    p = dm.pattern()
    dm.set_solver_state(dcop=0, tran=1, newton_iter=1, order=2, h=1e-12, time=1e-10, tstep=1e-12, tstop=2e-9)
    x0 = np.full(n, 1.8)
    dm.updateState(x0, np.zeros(max(dm.nsta, 1)), np.zeros(max(dm.nsto, 1)))
    vec = dm.loadDAEVectors()
    dm.loadDAEMatrices()
    rhs, J = dm.combine(vec["Q"], vec["F"] - vec["B"])
    t0 = time.time()
    s = X.DirectSolver(n, p["rowptr"], p["colind"])
    t_sym = time.time() - t0
    info = s.info()
    assert info["method"] == 2
    t0 = time.time(); assert s.numeric(J) == 0; t_num = time.time() - t0
    t0 = time.time(); st, x = s.solve(rhs); t_sol = time.time() - t0
    assert st == 0
    A = sp.csr_matrix((J, p["colind"], p["rowptr"]), shape=(n, n))
    res = np.max(np.abs(A @ x - rhs)) / np.max(np.abs(rhs))
    print(f"n={n} {info}: symbolic {t_sym:.2f} s, numeric {t_num*1e3:.1f} ms, solve {t_sol*1e3:.1f} ms, "
          f"residual {res:.2e}")
    assert res <= 1e-12
    xs = spla.spsolve(A.tocsc(), rhs)
    assert np.max(np.abs(x - xs)) <= 1e-9 * np.max(np.abs(xs))
    # refactor with the same values: bit-identical (fixed order, no atomics)
    assert s.numeric(J) == 0
    st, x2 = s.solve(rhs)
    assert np.array_equal(x, x2)
