"""Edge cases of the boundary and the batched engine.

* a NaN in the right-hand side gives XG_E_NAN (-99) and a zero solution
  (AmesosSolver::doSolve, N_LAS_AmesosSolver.C:415-424) with either LU
  method and with FP32 factors;
* bad arguments are rejected with XG_E_ARG before any device work
  (transpose solves on the multifrontal method; empty batches;
  an instance whose class does not fit a workgroup's LDS; launch
  configurations out of range);
* an instance that fails (its step size collapses) stops on its own: the
  other instances of the batch follow exactly the trajectory they follow
  alone.
"""
import ctypes

import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu


def _small_system(n=40, seed=3):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 4.0 + rng.random()
        for j in rng.choice(n, 3, replace=False):
            if j != i:
                A[i, j] = rng.normal()
    rp, ci, v = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += list(nz); v += list(A[i, nz]); rp.append(len(ci))
    return n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(v), A


@pytest.mark.parametrize("method,fp32", [(1, False), (2, False), (1, True), (2, True)])
def test_nan_right_hand_side_gives_status_and_zero_solution(gpu, method, fp32):
    n, rp, ci, v, A = _small_system()
    s = X.DirectSolver(n, rp, ci, method=method)
    if fp32:
        s.set_precision(1, 1e-12, 10)
    assert s.numeric(v) == 0
    b = np.ones(n)
    st, x = s.solve(b)
    assert st == 0 and np.max(np.abs(A @ x - b)) <= 1e-10
    b[7] = np.nan
    st, x = s.solve(b)
    assert st == -99
    assert np.all(x == 0.0)
    # the handle stays usable
    st, x = s.solve(np.ones(n))
    assert st == 0 and np.max(np.abs(A @ x - 1.0)) <= 1e-10


def test_bad_arguments_are_rejected(gpu):
    n, rp, ci, v, A = _small_system()
    L = X.lib()
    P = ctypes.c_void_p
    h = P()
    # n <= 0, null pattern
    assert L.xg_lu_symbolic(0, rp.ctypes.data_as(P), ci.ctypes.data_as(P), None, ctypes.byref(h)) < 0
    assert L.xg_lu_symbolic(n, None, ci.ctypes.data_as(P), None, ctypes.byref(h)) < 0
    # column index out of range
    bad = ci.copy()
    bad[5] = n + 3
    assert L.xg_lu_symbolic(n, rp.ctypes.data_as(P), bad.ctypes.data_as(P), None, ctypes.byref(h)) < 0
    # transpose solves: the level-scheduled method only
    s = X.DirectSolver(n, rp, ci, method=2)
    assert s.numeric(v) == 0
    b = np.ones(n)
    x = np.zeros(n)
    assert L.xg_lu_solve(s.h, b.ctypes.data_as(P), x.ctypes.data_as(P), 1) == -1
    # empty batch; an instance too large for a workgroup's LDS
    with pytest.raises(X.XgError):
        X.Batch([], 1e-9, 1e-12, 1e-11)
    with pytest.raises(X.XgError):
        X.Batch([C.gate_dag(250, seed=3)], 1e-9, 1e-12, 1e-11)
    # launch configuration out of range
    bt = X.Batch([C.ring(5, seed=1)], 1e-10, 1e-12, 1e-11)
    assert L.xg_batch_config(bt.h, 96, 0, -1, -1) == -1
    assert L.xg_batch_config(bt.h, 0, 2000, -1, -1) == -1
    assert L.xg_batch_config(bt.h, 0, 0, 3, -1) == -1
    assert L.xg_batch_config(bt.h, 0, 8, 2, 0) == 0
    t = np.zeros(4); xx = np.zeros(64)
    assert L.xg_batch_points(bt.h, 0, 4, t.ctypes.data_as(P), xx.ctypes.data_as(P)) == -1   # no recording
    # device loads of a circuit with MOSFETs need its state and store arrays
    dm = X.DeviceMgr(C.ring(3, seed=1))
    xs = np.zeros(dm.n)
    assert L.xg_update_state(dm.h, xs.ctypes.data_as(P), None, None, None, None, None) == -1


def test_failing_instance_does_not_disturb_the_others(gpu):
    """A gate array whose transient stalls (step size below the minimum, as
    in the oracle: without diffusion capacitance a NAND2's internal node
    carries no charge) sits in the same batch as ring oscillators; the rings'
    trajectories are bit-identical to their runs alone."""
    stall = C.gate_dag(80, seed=8, frac_nand=0.8, diffusion_caps=False)
    rings = [C.ring(9, seed=21, mc_seed=21), C.ring(15, seed=22, mc_seed=22)]
    tstop = 0.4e-9
    b = X.Batch([rings[0], stall, rings[1]], tstop, 1e-12, 1e-11, record_max=4000)
    cnt = b.run_to_end(chunk=256)
    assert cnt["done"] == 2 and cnt["failed"] == 1, cnt
    n_stall = X.host_pattern(stall)["n"]
    assert b.instance(1, n_stall)[0]["mode"] < 0
    for i, c in ((0, rings[0]), (2, rings[1])):
        n = X.host_pattern(c)["n"]
        b1 = X.Batch([c], tstop, 1e-12, 1e-11, record_max=4000)
        b1.run_to_end(chunk=256)
        t_a, x_a = b.points(i, n)
        t_b, x_b = b1.points(0, n)
        assert np.array_equal(t_a, t_b) and np.array_equal(x_a, x_b)
