"""Mixed-precision LU on the GPU: FP32 factors and triangular solves with FP64
residuals and iterative refinement, against the all-FP64 answer.

Systems: transient Jacobians (trapezoid, h = 2 ps) of level-1 INV/NAND2 gate
arrays wired as a random DAG, at a perturbed DC operating point of the
oracle.  The FP64 answer is SuperLU (scipy) on the oracle's CSR values.

Bars: relative residual max|b - A x| / max|b| <= 1e-12 (the refinement
target), and per-node error |x - x64| <= 1e-9 * (RELTOL |x64| + ABSTOL) with
RELTOL 1e-3, ABSTOL 1e-6, i.e. far inside the transient tolerances.  The
per-node error is printed.  Hand-made ill-conditioned systems check the FP64
fallbacks (FP32 pivot failure, refinement that stops contracting).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import xyce_amd as X
from oracle.pyoracle import Oracle
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu


def _tran_system(c, seed, rails):
    """Transient Jacobian (trapezoid, h = 2 ps) at a chosen state: every node
    near mid-rail (all MOSFETs conducting: a well-conditioned system) or
    near the rails (stacks with both devices off leave mid nodes held only by
    gmin: an ill-conditioned one)."""
    o = Oracle(c)
    rng = np.random.default_rng(seed)
    gid = o.pattern()["node_gid"]
    x = np.zeros(o.n)
    for node in range(1, c.n_nodes):
        v = rng.choice([0.05, 1.75]) if rails else 0.9
        x[gid[node]] = v + rng.normal(0.0, 0.02)
    x[gid[1]] = 1.8                       # supply node
    sta0, sto0 = np.zeros(max(o.nsta, 1)), np.zeros(max(o.nsto, 1))
    ref = o.load(x, sta0, sto0, h=2e-12, order=2)
    rhs, J = o.combine(ref, ref["Q"], ref["F"] - ref["B"])
    return o, J, rhs


@pytest.mark.parametrize("gates", [300, 3000])
def test_fp32_refined_solve_matches_fp64(gpu, gates):
    c = C.gate_dag(gates, seed=4)
    o, J, rhs = _tran_system(c, gates, rails=False)
    p = o.pattern()
    s = X.DirectSolver(o.n, p["rowptr"], p["colind"], values=J)   # pivot order steered by the values
    s.set_precision(1, ir_tol=1e-12, ir_max=10)
    for rep in range(2):       # refactor + solve twice (the Newton loop's pattern)
        st, x = s.doSolve(J * (1.0 + 0.01 * rep), rhs)
        assert st == 0
    ps = s.precision_stats()
    assert ps["precision"] == 1 and ps["fp64_fallbacks"] == 0 and ps["fp32_pivot_failures"] == 0, ps
    A = sp.csr_matrix((J * 1.01, p["colind"], p["rowptr"]), shape=(o.n, o.n))
    x64 = spla.spsolve(A.tocsc(), rhs)
    res = np.max(np.abs(A @ x - rhs)) / np.max(np.abs(rhs))
    err = np.abs(x - x64) / (1e-3 * np.abs(x64) + 1e-6)
    print(f"n={o.n}: corrections {ps['last_corrections']}, rel residual {res:.2e}, "
          f"max per-node error {err.max():.2e} of RELTOL|x|+ABSTOL")
    assert res <= 1e-12
    assert err.max() <= 1e-9
    # the FP64 path on the same system agrees too
    s64 = X.DirectSolver(o.n, p["rowptr"], p["colind"])
    st, xd = s64.doSolve(J * 1.01, rhs)
    assert st == 0
    assert np.max(np.abs(xd - x) / (1e-3 * np.abs(x64) + 1e-6)) <= 1e-9


def test_fp32_refined_solve_on_ill_conditioned_state(gpu):
    """Near-rail state: whatever mix of FP32 refinement and FP64 fallback the
    solver picks, the answer matches the FP64 path's."""
    c = C.gate_dag(2000, seed=6)
    o, J, rhs = _tran_system(c, 9, rails=True)
    p = o.pattern()
    s64 = X.DirectSolver(o.n, p["rowptr"], p["colind"])
    st, xd = s64.doSolve(J, rhs)
    assert st == 0
    s = X.DirectSolver(o.n, p["rowptr"], p["colind"])
    s.set_precision(1, ir_tol=1e-12, ir_max=10)
    st, x = s.doSolve(J, rhs)
    assert st == 0
    ps = s.precision_stats()
    A = sp.csr_matrix((J, p["colind"], p["rowptr"]), shape=(o.n, o.n))
    r32 = np.max(np.abs(A @ x - rhs)) / np.max(np.abs(rhs))
    r64 = np.max(np.abs(A @ xd - rhs)) / np.max(np.abs(rhs))
    err = np.abs(x - xd) / (1e-3 * np.abs(xd) + 1e-6)
    print(f"n={o.n}: {ps}; residual fp32+IR {r32:.2e} fp64 {r64:.2e}; max per-node diff {err.max():.2e}")
    assert r32 <= max(1e-12, 10 * r64)
    assert err.max() <= 1e-6


def _two_by_two(delta):
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    v = np.array([1.0, 1.0, 1.0, 1.0 + delta])
    return rp, ci, v


def test_fp32_singular_pivot_falls_back(gpu):
    # 1 + 1e-9 rounds to 1 in FP32: the FP32 pivot is exactly zero
    rp, ci, v = _two_by_two(1e-9)
    b = np.array([1.0, 2.0])
    s = X.DirectSolver(2, rp, ci)
    s.set_precision(1)
    st, x = s.doSolve(v, b)
    assert st == 0
    ps = s.precision_stats()
    assert ps["fp32_pivot_failures"] >= 1
    xn = np.linalg.solve(np.array([[1.0, 1.0], [1.0, 1.0 + 1e-9]]), b)
    np.testing.assert_allclose(x, xn, rtol=1e-6)


def test_refinement_that_stalls_falls_back(gpu):
    """cond ~ 1e8: the FP32 factors' error times the condition number is
    near 1 and FP64 itself cannot reach 1e-12, so the residual stops
    decreasing and the solve finishes with the FP64 factors."""
    A = np.array([[1.0, 1.0 / 3.0], [1.0, 1.0 / 3.0 + 5e-8]])
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    b = np.array([1.0, 2.0])
    s = X.DirectSolver(2, rp, ci)
    s.set_precision(1, ir_tol=1e-12, ir_max=10)
    st, x = s.doSolve(A.ravel(), b)
    assert st == 0
    ps = s.precision_stats()
    assert ps["fp32_pivot_failures"] == 0 and ps["fp64_fallbacks"] == 1, ps
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-6)


@pytest.mark.parametrize("frac_nand", [0.5, 1.0])
def test_gate_array_full_size_fp32_refined(gpu, frac_nand):
    """250k gates, half INV and half NAND2 (749,642 level-1 MOSFETs, 375,323
    unknowns) or all NAND2 (1,000,000 MOSFETs, 500,502 unknowns): the
    transient Jacobian from the GPU device loads (checked entrywise against
    the oracle's), factored by the multifrontal path in FP32 with FP64
    refinement; error per node against SuperLU's FP64 solution."""
    c = C.gate_dag(250_000, seed=4, frac_nand=frac_nand)
    o = Oracle(c)
    rng = np.random.default_rng(11)
    x = 0.9 + rng.normal(0.0, 0.05, size=o.n)
    sta0, sto0 = np.zeros(max(o.nsta, 1)), np.zeros(max(o.nsto, 1))
    ref = o.load(x, sta0, sto0, h=2e-12, order=2)
    rhs_o, J_o = o.combine(ref, ref["Q"], ref["F"] - ref["B"])
    dm = X.DeviceMgr(c)
    dm.set_solver_state(dcop=0, tran=1, newton_iter=1, order=2, h=2e-12, time=0.0, tstep=1e-12, tstop=1e-9)
    dm.updateState(x, sta0, sto0)
    vec = dm.loadDAEVectors()
    dm.loadDAEMatrices()
    rhs, J = dm.combine(ref["Q"], ref["F"] - ref["B"])
    p = o.pattern()
    # entries within 1e-12 of their magnitude; an entry summed from n_t device
    # terms (the supply rows collect one from every PMOS: n_t ~ 375k) within
    # the rounding bound 2 n_t eps of the sum, which the reference's own
    # sequential sum meets no better (the GPU sums long entries by a
    # fixed-order workgroup reduction)
    fs, qs = p["mos_fslot"].ravel(), p["mos_qslot"].ravel()
    n_t = np.bincount(fs[fs >= 0], minlength=len(J)) + np.bincount(qs[qs >= 0], minlength=len(J))
    tol = np.maximum(1e-12, 2 * n_t * np.finfo(float).eps)
    scale = np.maximum(np.abs(J_o), 1e-12 * np.max(np.abs(J_o)))
    assert np.all(np.abs(J - J_o) <= tol * scale), np.max(np.abs(J - J_o) / (tol * scale))
    s = X.DirectSolver(o.n, p["rowptr"], p["colind"], values=J)
    assert s.info()["method"] == 2
    s.set_precision(1, ir_tol=1e-12, ir_max=10)
    st, xs = s.doSolve(J, rhs)
    assert st == 0
    ps = s.precision_stats()
    A = sp.csr_matrix((J, p["colind"], p["rowptr"]), shape=(o.n, o.n))
    # the residual in extended precision: the supply rows sum ~375k products,
    # whose FP64 sequential sum alone can be 1e-12 of the right-hand side off
    prod = J.astype(np.longdouble) * xs.astype(np.longdouble)[p["colind"]]
    ax = np.add.reduceat(prod, p["rowptr"][:-1])
    ax[np.diff(p["rowptr"]) == 0] = 0
    res = float(np.max(np.abs(ax - rhs.astype(np.longdouble))) / np.max(np.abs(rhs)))
    x64 = spla.spsolve(A.tocsc(), rhs)
    err = np.abs(xs - x64) / (1e-3 * np.abs(x64) + 1e-6)
    gid = p["node_gid"]
    outs = np.array([gid[nd] for nd in range(1, c.n_nodes) if gid[nd] >= 0])
    print(f"n={o.n}, {len(c.mos_nodes)} MOSFETs, {s.info()}: {ps}; residual {res:.2e}; per-node error vs FP64 "
          f"(units of RELTOL|x|+ABSTOL): max {err[outs].max():.2e}, median {np.median(err[outs]):.2e}")
    assert res <= 1e-12
    assert err.max() <= 1e-6
