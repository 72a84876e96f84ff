"""GPU parity: the HIP path against the CPU oracle through the C ABI.

Bars:
  * pattern, node order and stamp slots: bit-exact;
  * F, Q, B, limiter vectors, dFdx, dQdx, J, RHS entries: |gpu - oracle| <=
    1e-12 * (magnitude of the entry's terms) + 1e-30;
  * linear solves: relative residual <= 1e-12 and agreement with numpy;
  * transient waveforms: within RELTOL*|V| + ABSTOL of the oracle's
    (RELTOL 1e-3, ABSTOL 1e-6, plus the interpolation error between the two
    step sequences, bounded by the LTE tolerance).
"""
import time

import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C
from oracle.pyoracle import Oracle
from tests.helpers import entry_tol, waveform_error

pytestmark = pytest.mark.gpu


def _loaded_state(o, seed, tran=True):
    """Oracle DC point + a perturbed transient state (store/state from a
    previous load) so limiting, both MOSFET modes and all regions occur."""
    st, x0, sta0, sto0, Q, F, B = o.dcop()
    assert st > 0
    rng = np.random.default_rng(seed)
    x = x0 + rng.normal(0.0, 0.4, size=o.n)
    return x, sta0, sto0


@pytest.mark.parametrize("stages,seed", [(3, 0), (11, 1), (101, 2)])
def test_pattern_bit_exact(gpu, stages, seed):
    c = C.ring(stages, seed=seed, mc_seed=seed)
    d = X.DeviceMgr(c)
    o = Oracle(c)
    pg, po = d.pattern(), o.pattern()
    for k in ("rowptr", "colind", "node_gid", "mos_lid", "mos_fslot", "mos_qslot", "mos_stasto"):
        assert np.array_equal(pg[k], po[k]), k


@pytest.mark.parametrize("stages,seed,newton_iter,order", [(5, 3, 0, 1), (11, 4, 1, 2), (101, 5, 2, 2)])
def test_load_parity(gpu, stages, seed, newton_iter, order):
    c = C.ring(stages, seed=seed, mc_seed=seed)
    d = X.DeviceMgr(c)
    o = Oracle(c)
    x, sta0, sto0 = _loaded_state(o, seed)
    kw = dict(dcop=0, tran=1, init_jct=0, newton_iter=newton_iter, order=order, h=7e-12, time=0.3e-9,
              tstep=1e-12, tstop=5e-9)
    ref = o.load(x, sta0, sto0, sta_next=sta0, sto_next=sto0, **kw)
    d.set_solver_state(**kw)
    sta_n, sto_n, lim = d.updateState(x, sta0, sto0, sto_next=sto0)
    assert lim == ref["limited"]
    assert entry_tol(sta_n, ref["sta_next"], np.abs(ref["sta_next"]) + 1e-300) <= 1
    assert entry_tol(sto_n, ref["sto_next"], np.abs(ref["sto_next"]) + 1e-300) <= 1
    vec = d.loadDAEVectors()
    # magnitude of the terms summed into each entry: bound by the sum of |contributions|,
    # which the oracle's |value| underestimates only under cancellation; use max(|F|, 1e-6*max|F|)
    for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp"):
        a, b = vec[k], ref[k]
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1, k
    dF, dQ = d.loadDAEMatrices()
    for a, b in ((dF, ref["dFdx"]), (dQ, ref["dQdx"])):
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1
    q0 = ref["Q"] * 0.999
    q2 = ref["F"] - ref["B"]
    rhs_g, J_g = d.combine(q0, q2)
    rhs_o, J_o = o.combine(ref, q0, q2)
    s_r = np.maximum(np.abs(rhs_o), 1e-9 * np.max(np.abs(rhs_o)))
    assert entry_tol(rhs_g, rhs_o, s_r, abs_floor=1e-30) <= 1
    s_j = np.maximum(np.abs(J_o), 1e-9 * np.max(np.abs(J_o)))
    assert entry_tol(J_g, J_o, s_j, abs_floor=1e-30) <= 1


CAPS = dict(TOX=2e-8, CJ=2e-4, CJSW=1e-9, CGSO=3e-10, CGDO=3e-10, CGBO=2e-10, GAMMA=0.4, LAMBDA=0.03)


@pytest.mark.parametrize("newton_iter,order", [(0, 1), (2, 2)])
def test_load_parity_with_caps(gpu, newton_iter, order):
    """Meyer gate capacitances, junction depletion charges and overlap caps on."""
    c = C.inverter_chain(4, **CAPS)
    d = X.DeviceMgr(c)
    o = Oracle(c)
    x, sta0, sto0 = _loaded_state(o, 11)
    kw = dict(dcop=0, tran=1, init_jct=0, newton_iter=newton_iter, order=order, h=3e-12, time=0.4e-9,
              tstep=1e-12, tstop=5e-9)
    sta0 = sta0 * 1.01
    ref = o.load(x, sta0, sto0, sta_next=sta0, sto_next=sto0, **kw)
    d.set_solver_state(**kw)
    sta_n, sto_n, lim = d.updateState(x, sta0, sto0, sto_next=sto0)
    assert lim == ref["limited"]
    assert entry_tol(sta_n, ref["sta_next"], np.abs(ref["sta_next"]) + 1e-300) <= 1
    vec = d.loadDAEVectors()
    for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp"):
        a, b = vec[k], ref[k]
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1, k
    assert np.max(np.abs(ref["Q"])) > 0 and np.max(np.abs(ref["dQdx"])) > 0
    dF, dQ = d.loadDAEMatrices()
    for a, b in ((dF, ref["dFdx"]), (dQ, ref["dQdx"])):
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("dcop,init_jct,newton_iter,tranop", [(1, 1, 0, 0), (1, 0, 1, 0), (1, 1, 0, 1), (1, 0, 1, 1),
                                                               (0, 0, 0, 0), (0, 0, 2, 0)])
def test_device_zoo_load_parity(gpu, seed, dcop, init_jct, newton_iter, tranop):
    """200 independent devices over four model cards (every depletion-charge
    form) at random terminal voltages: every operating branch of
    updateIntermediateVars occurs (checked on the oracle's output, as in
    test_oracle_devices.py::test_device_zoo_reaches_every_branch), and every
    state, store, vector and matrix entry meets the 1e-12 bar."""
    from tests.helpers import ALL_BRANCHES, branch_classes, device_zoo, zoo_states
    c = device_zoo()
    d = X.DeviceMgr(c)
    o = Oracle(c)
    x, sta0, sto0 = zoo_states(o, seed)
    kw = dict(dcop=dcop, tran=1 - dcop, init_jct=init_jct, newton_iter=newton_iter, order=2, h=3e-12,
              time=0.4e-9, tstep=1e-12, tstop=5e-9, tranop=tranop)
    ref = o.load(x, sta0, sto0, sta_next=sta0, sto_next=sto0, **kw)
    if tranop:
        # a transient's operating point stamps the capacitances (the zoo has Meyer and junction caps)
        assert np.max(np.abs(ref["dQdx"])) > 0
    if not init_jct:
        assert ALL_BRANCHES <= branch_classes(c, o.pattern()["mos_stasto"], ref["sto_next"])
    d.set_solver_state(**kw)
    sta_n, sto_n, lim = d.updateState(x, sta0, sto0, sto_next=sto0)
    assert lim == ref["limited"]
    assert entry_tol(sta_n, ref["sta_next"], np.abs(ref["sta_next"]) + 1e-300) <= 1
    assert entry_tol(sto_n, ref["sto_next"], np.abs(ref["sto_next"]) + 1e-300) <= 1
    vec = d.loadDAEVectors()
    for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp"):
        a, b = vec[k], ref[k]
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1, k
    dF, dQ = d.loadDAEMatrices()
    for a, b in ((dF, ref["dFdx"]), (dQ, ref["dQdx"])):
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1


@pytest.mark.parametrize("temp", [250.0, 400.0])
def test_device_zoo_load_parity_off_nominal_temperature(gpu, temp):
    """The same 200 devices with TEMP != TNOM (updateTemperature,
    N_DEV_MOSFET1.C:3487; the oracle's temperature scaling is pinned against
    device physics in test_oracle_devices.py): every entry meets the 1e-12 bar."""
    from tests.helpers import device_zoo, zoo_states
    c = device_zoo()
    d = X.DeviceMgr(c, temp=temp)
    o = Oracle(c, temp=temp)
    x, sta0, sto0 = zoo_states(o, 1)
    kw = dict(dcop=0, tran=1, init_jct=0, newton_iter=2, order=2, h=3e-12, time=0.4e-9, tstep=1e-12, tstop=5e-9)
    ref = o.load(x, sta0, sto0, sta_next=sta0, sto_next=sto0, **kw)
    d.set_solver_state(**kw)
    sta_n, sto_n, lim = d.updateState(x, sta0, sto0, sto_next=sto0)
    assert lim == ref["limited"]
    assert entry_tol(sta_n, ref["sta_next"], np.abs(ref["sta_next"]) + 1e-300) <= 1
    assert entry_tol(sto_n, ref["sto_next"], np.abs(ref["sto_next"]) + 1e-300) <= 1
    vec = d.loadDAEVectors()
    for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp"):
        a, b = vec[k], ref[k]
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1, k
    dF, dQ = d.loadDAEMatrices()
    for a, b in ((dF, ref["dFdx"]), (dQ, ref["dQdx"])):
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        assert entry_tol(a, b, scale, abs_floor=1e-30) <= 1
    # the temperature did change the loads
    o0 = Oracle(c)
    r0 = o0.load(x, sta0, sto0, sta_next=sta0, sto_next=sto0, **kw)
    assert np.max(np.abs(r0["F"] - ref["F"])) > 1e-6 * np.max(np.abs(ref["F"]))


def test_dc_load_parity(gpu):
    c = C.ring(7, seed=9, mc_seed=9)
    d = X.DeviceMgr(c)
    o = Oracle(c)
    x = np.random.default_rng(1).uniform(0, 1.8, size=o.n)
    z_sta = np.zeros(o.nsta); z_sto = np.zeros(o.nsto)
    kw = dict(dcop=1, tran=0, init_jct=1, newton_iter=0, order=1, h=1e-12, time=0.0, tstep=1e-12, tstop=1e-9)
    ref = o.load(x, z_sta, z_sto, **kw)
    d.set_solver_state(**kw)
    sta_n, sto_n, lim = d.updateState(x, z_sta, z_sto, sto_next=z_sto)
    assert lim == ref["limited"]
    np.testing.assert_allclose(sto_n, ref["sto_next"], rtol=1e-12, atol=1e-30)
    dF, dQ = d.loadDAEMatrices()
    np.testing.assert_allclose(dF, ref["dFdx"], rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("stages", [5, 101, 201])
def test_lu_solve_parity(gpu, stages):
    c = C.ring(stages, seed=stages, mc_seed=stages)
    o = Oracle(c)
    x, sta0, sto0 = _loaded_state(o, stages)
    ref = o.load(x, sta0, sto0, h=5e-12, order=2)
    rhs, J = o.combine(ref, ref["Q"], ref["F"] - ref["B"])
    p = o.pattern()
    s = X.DirectSolver(o.n, p["rowptr"], p["colind"])
    st, xg = s.doSolve(J, rhs)
    assert st == 0
    A = np.zeros((o.n, o.n))
    for r in range(o.n):
        for k in range(p["rowptr"][r], p["rowptr"][r + 1]):
            A[r, p["colind"][k]] += J[k]
    xn = np.linalg.solve(A, rhs)
    res = np.linalg.norm(A @ xg - rhs) / np.linalg.norm(rhs)
    assert res <= 1e-12, res
    np.testing.assert_allclose(xg, xn, rtol=1e-9, atol=1e-12 * np.max(np.abs(xn)))


@pytest.mark.parametrize("stages", [5, 101])
def test_lu_transpose_solve(gpu, stages):
    """doSolve(reuse_factors, transpose) (N_LAS_Solver.h:95-107): A^T x = b
    from the factors of A, against numpy, and again after a refactor with new
    values; the multifrontal method and FP32 factors refuse it."""
    c = C.ring(stages, seed=stages, mc_seed=stages)
    o = Oracle(c)
    x, sta0, sto0 = _loaded_state(o, stages)
    ref = o.load(x, sta0, sto0, h=5e-12, order=2)
    rhs, J = o.combine(ref, ref["Q"], ref["F"] - ref["B"])
    p = o.pattern()
    s = X.DirectSolver(o.n, p["rowptr"], p["colind"], method=1)
    for vals in (J, J * (1.0 + 0.1 * np.random.default_rng(stages).uniform(size=J.size))):
        A = np.zeros((o.n, o.n))
        for r in range(o.n):
            for k in range(p["rowptr"][r], p["rowptr"][r + 1]):
                A[r, p["colind"][k]] += vals[k]
        st, xt = s.doSolve(vals, rhs, transpose=True)
        assert st == 0
        res = np.linalg.norm(A.T @ xt - rhs) / np.linalg.norm(rhs)
        assert res <= 1e-12, res
        np.testing.assert_allclose(xt, np.linalg.solve(A.T, rhs), rtol=1e-9, atol=1e-12 * np.max(np.abs(xt)))
        # the plain solve from the same factors still solves A x = b
        st, xn = s.solve(rhs)
        assert np.linalg.norm(A @ xn - rhs) / np.linalg.norm(rhs) <= 1e-12
    s2 = X.DirectSolver(o.n, p["rowptr"], p["colind"], method=2)
    assert s2.numeric(J) == 0
    with pytest.raises(X.XgError):
        s2.solve(rhs, transpose=True)


def test_lu_singular_returns_status(gpu):
    # structurally nonsingular, numerically singular 3x3
    rp = np.array([0, 2, 4, 6], np.int32)
    ci = np.array([0, 1, 0, 1, 1, 2], np.int32)
    v = np.array([1.0, 2.0, 2.0, 4.0, 1.0, 1.0])
    s = X.DirectSolver(3, rp, ci)
    st, x = s.doSolve(v, np.ones(3))
    assert st > 0 and np.all(x == 0.0)


@pytest.mark.parametrize("method", [1, 2])
def test_lu_zero_pivot_value_repivots(gpu, method):
    """A nonsingular matrix with a [[0, 1], [1, 0]] block whose zeros are
    stored: the pivot order chosen from the pattern alone puts a zero on the
    diagonal (checked on the host analysis), so xg_lu_numeric re-chooses the
    pivots from the values, as KLU re-pivots at every factorisation, and the
    solve succeeds; so does a refactor with new values."""
    rp = np.array([0, 2, 4, 5], np.int32)
    ci = np.array([0, 1, 0, 1, 2], np.int32)
    A = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    h = X.host_lu_analyse(3, rp, ci)
    assert np.any(np.diag(A[h["perm_row"]][:, h["perm_col"]]) == 0.0)
    vals = lambda M: np.array([M[r, c] for r in range(3) for c in ci[rp[r]:rp[r + 1]]])
    s = X.DirectSolver(3, rp, ci, method=method)
    b = np.array([1.0, 2.0, 3.0])
    st, x = s.doSolve(vals(A), b)
    assert st == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-13)
    A2 = A + np.diag([0.0, 1e-9, 0.5])
    st, x = s.doSolve(vals(A2), b)
    assert st == 0
    np.testing.assert_allclose(x, np.linalg.solve(A2, b), rtol=1e-12)


@pytest.mark.parametrize("stages,tstop", [(5, 2e-9), (11, 2e-9)])
def test_batch_single_transient_matches_oracle(gpu, stages, tstop):
    c = C.ring(stages, seed=stages)
    b = X.Batch([c], tstop, 1e-12, 1e-11, record_max=20000)
    cnt = b.run_to_end(chunk=512)
    assert cnt["done"] == 1 and cnt["failed"] == 0, cnt
    o = Oracle(c)
    r = o.tran(tstop, 1e-12, 1e-11)
    assert r["status"] == 0
    t_g, x_g = b.points(0, o.n)
    st, _ = b.instance(0, o.n)
    # same algorithm: same step count up to rounding-driven differences
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    err = waveform_error(t_g, x_g, r["t"], r["x"])
    print(f"steps gpu {st['steps']} oracle {r['steps']}  max waveform diff {err:.3e} V")
    assert err <= 1e-3 * 1.8 + 1e-6, err


def test_batch_with_caps_matches_oracle(gpu):
    """Driven inverter chain with every capacitance on: Meyer back-averaging,
    depletion charges, overlap caps, PULSE input breakpoints."""
    c = C.inverter_chain(4, **CAPS)
    tstop = 4e-9
    b = X.Batch([c], tstop, 1e-12, 1e-11, record_max=20000)
    cnt = b.run_to_end(chunk=512)
    assert cnt["done"] == 1 and cnt["failed"] == 0, cnt
    o = Oracle(c)
    r = o.tran(tstop, 1e-12, 1e-11)
    assert r["status"] == 0
    t_g, x_g = b.points(0, o.n)
    err = waveform_error(t_g, x_g, r["t"], r["x"])
    print(f"caps: steps gpu {len(t_g)} oracle {len(r['t'])}  max waveform diff {err:.3e}")
    assert err <= 1e-3 * 1.8 + 1e-6, err


def test_batch_uneven_instances_match_individual_runs(gpu):
    """A batch of rings of different sizes gives each instance the same
    trajectory as running it alone (instances are independent)."""
    cs = [C.ring(s, seed=100 + s, mc_seed=100 + s) for s in (11, 31, 7, 61)]
    tstop = 1e-9
    b = X.Batch(cs, tstop, 1e-12, 1e-11, record_max=4000)
    cnt = b.run_to_end(chunk=256)
    assert cnt["done"] == 4 and cnt["failed"] == 0, cnt
    for i, c in enumerate(cs):
        b1 = X.Batch([c], tstop, 1e-12, 1e-11, record_max=4000)
        b1.run_to_end(chunk=256)
        n = X.host_pattern(c)["n"]
        t_a, x_a = b.points(i, n)
        t_b, x_b = b1.points(0, n)
        assert np.array_equal(t_a, t_b) and np.array_equal(x_a, x_b)


def test_launch_forms_do_not_change_results(gpu):
    """The separate evaluation kernel (one iteration per launch) and the
    fused kernel with 1 or K Newton iterations per launch (state kept in LDS,
    the term arrays in HBM or in LDS),
    with the LDS-sized instance groups one after another or side by side on
    their own streams, give bit-identical trajectories."""
    cs = [C.ring(s, seed=7 + s, mc_seed=7 + s) for s in (9, 23, 15, 31, 11)]
    out = []
    ns = [X.host_pattern(c)["n"] for c in cs]
    for split, k, conc in ((1, 1, 0), (0, 1, 0), (0, 4, 0), (0, 4, 1), (1, 1, 1), (2, 4, 0), (2, 1, 1)):
        b = X.Batch(cs, 0.5e-9, 1e-12, 1e-11, record_max=2000)
        b.config(iters_per_launch=k, split_eval=split, concurrent_groups=conc)
        b.run_to_end(chunk=64)
        out.append(tuple(b.points(i, ns[i]) for i in range(len(cs))))
    for other in out[1:]:
        for a, c in zip(out[0], other):
            assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_workgroup_size_does_not_change_results(gpu):
    """Instances of more than 128 unknowns (151- and 201-stage rings) give
    bit-identical trajectories with workgroups of 64, 128, 256 and 512 threads,
    fused or not: the residual norm and the local error estimate are summed
    in an order that does not depend on the workgroup size."""
    cs = [C.ring(s, seed=40 + s, mc_seed=40 + s) for s in (151, 201, 61)]
    ns = [X.host_pattern(c)["n"] for c in cs]
    out = []
    for block, split, k in ((128, 0, 4), (64, 0, 4), (256, 0, 4), (512, 0, 4), (256, 1, 1), (512, 2, 4)):
        b = X.Batch(cs, 0.3e-9, 1e-12, 1e-11, record_max=2000)
        b.config(block=block, iters_per_launch=k, split_eval=split)
        cnt = b.run_to_end(chunk=64)
        assert cnt["done"] == len(cs), cnt
        out.append(tuple(b.points(i, ns[i]) for i in range(len(cs))))
    for other in out[1:]:
        for a, c in zip(out[0], other):
            assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_graph_replays_match_eager_launches(gpu):
    """Launches captured in a hipGraph and replayed (with counter reads and
    synchronisation between replays) give the same trajectory as eager
    launches, also with the LDS-sized groups side by side, where each group
    runs its launches of a graph or an advance back to back on its own
    stream."""
    cs = [C.ring(s, seed=3 + s, mc_seed=3 + s) for s in (13, 29, 9, 21)]
    ns = [X.host_pattern(c)["n"] for c in cs]
    for split, k in ((1, 1), (0, 2)):
        res = []
        for chunk, conc in ((0, 0), (16, 0), (16, 1), (0, 1)):
            b = X.Batch(cs, 0.3e-9, 1e-12, 1e-11, record_max=2000)
            b.config(split_eval=split, iters_per_launch=k, concurrent_groups=conc)
            for _ in range(40):
                b.advance(16, graph_chunk=chunk)
                b.counters()
            res.append((tuple(b.points(i, ns[i]) for i in range(len(cs))), b.counters()["newton_iters"]))
        for other in res[1:]:
            assert other[1] == res[0][1], (split, other[1], res[0][1])
            for a, c in zip(res[0][0], other[0]):
                assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_batch_power_grids_match_oracle(gpu):
    """Linear RC meshes with PWL sinks at different times: each instance
    stops at its own source breakpoints and follows the oracle."""
    cs = [C.power_grid(6, 6, seed=20 + k, pad_pitch=3, n_sinks=4, tstop=1e-9) for k in range(3)]
    b = X.Batch(cs, 1e-9, 1e-12, 0.0, record_max=4000)
    cnt = b.run_to_end(chunk=256)
    assert cnt["done"] == 3 and cnt["failed"] == 0, cnt
    for i, c in enumerate(cs):
        o = Oracle(c)
        r = o.tran(1e-9, 1e-12, 0.0)
        assert r["status"] == 0
        t_g, x_g = b.points(i, o.n)
        assert abs(len(t_g) - len(r["t"])) <= 0.02 * len(r["t"]) + 2, (len(t_g), len(r["t"]))
        err = waveform_error(t_g, x_g, r["t"], r["x"])
        assert err <= 1e-3 * 1.8 + 1e-6, err


def test_netlist_transient_matches_oracle(gpu, tmp_path):
    """A hand-written Xyce netlist read by the netlist reader, run on the GPU,
    written as .PRN, against the oracle's transient of the same circuit."""
    from xyce_amd.netlist import parse_netlist, read_prn, run_transient
    text = """inverter chain
V1 vdd 0 DC 1.8
V2 in 0 PULSE(0 1.8 0.2n 50p 50p 1n 3n)
M1 a in 0 0 nch W=2u L=0.5u
M2 a in vdd vdd pch W=4u L=0.5u
M3 out a 0 0 nch W=2u L=0.5u
M4 out a vdd vdd pch W=4u L=0.5u
C1 a 0 3f
C2 out 0 5f
.model nch NMOS LEVEL=1 VTO=0.7 KP=110u
.model pch PMOS LEVEL=1 VTO=-0.7 KP=50u
.TRAN 1p 4n 0 10p
.PRINT TRAN V(a) V(out)
.END
"""
    t, cols = run_transient(text, prn_path=tmp_path / "inv.prn")
    hdr, data = read_prn(tmp_path / "inv.prn")
    assert hdr == ["Index", "TIME", "V(A)", "V(OUT)"] and len(data) == len(t)
    c, an = parse_netlist(text)
    o = Oracle(c)
    r = o.tran(4e-9, 1e-12, 10e-12)
    gid = o.pattern()["node_gid"]
    for nm in ("A", "OUT"):
        col = gid[an["nodes"][nm]]
        err = waveform_error(t, cols[f"V({nm})"][:, None], r["t"], r["x"][:, [col]])
        assert err <= 1e-3 * 1.8 + 1e-6, (nm, err)


def test_reference_driver_drives_the_dropin(gpu):
    """The drop-in boundary end to end: the oracle's restatement of the
    reference's driver (DampedNewton, OneStep, DC operating point) calls the
    GPU device loads (xg_update_state / xg_load_vectors / xg_load_matrices)
    and the GPU LU (xg_lu_numeric / xg_lu_solve) through the C ABI at every
    Newton iteration, as Xyce's DeviceMgr and Linear::Solver would be called;
    the run reproduces the all-CPU oracle run."""
    c = C.ring(5, seed=5)
    tstop = 0.4e-9
    ref = Oracle(c).tran(tstop, 1e-12, 1e-11)
    assert ref["status"] == 0
    drv = Oracle(c)
    dm = X.DeviceMgr(c)
    p = dm.pattern()
    ls = X.DirectSolver(dm.n, p["rowptr"], p["colind"])
    n, nnz, nsta, nsto = dm.n, dm.nnz, dm.nsta, dm.nsto
    arr = lambda ptr, k: np.ctypeslib.as_array(ptr, shape=(max(k, 1),))
    errors = []
    calls = {"load": 0, "solve": 0}

    def load(ctx, ssi, ssd, x, sc, oc, sn, on, F, Q, B, dFv, dQv, dFdx, dQdx):
        try:
            si, sd = arr(ssi, 7), arr(ssd, 4)
            dm.set_solver_state(dcop=int(si[0]), tran=int(si[1]), init_jct=int(si[2]), newton_iter=int(si[4]),
                                order=int(si[5]), h=float(sd[0]), time=float(sd[1]), tstep=float(sd[2]),
                                tstop=float(sd[3]), tranop=int(si[6]))
            sta_n, sto_n, lim = dm.updateState(arr(x, n)[:n].copy(), arr(sc, nsta)[:nsta].copy(),
                                               arr(oc, nsto)[:nsto].copy(), sto_next=arr(on, nsto)[:nsto].copy())
            arr(sn, nsta)[:nsta] = sta_n
            arr(on, nsto)[:nsto] = sto_n
            v = dm.loadDAEVectors()
            for name, ptr in (("F", F), ("Q", Q), ("B", B), ("dFdxdVp", dFv), ("dQdxdVp", dQv)):
                arr(ptr, n)[:n] = v[name]
            dF, dQ = dm.loadDAEMatrices()
            arr(dFdx, nnz)[:nnz] = dF
            arr(dQdx, nnz)[:nnz] = dQ
            calls["load"] += 1
            return int(lim)
        except Exception as e:      # surfaced after the run
            errors.append(repr(e))
            return 0

    def solve(ctx, nn, rp, ci, vals, b, x):
        try:
            st, xs = ls.doSolve(arr(vals, nnz)[:nnz].copy(), arr(b, nn)[:nn].copy())
            arr(x, nn)[:nn] = xs
            calls["solve"] += 1
            return int(st)
        except Exception as e:
            errors.append(repr(e))
            return -1

    drv.set_hooks(load, solve)
    r = drv.tran(tstop, 1e-12, 1e-11)
    assert not errors, errors[:3]
    assert r["status"] == 0
    print(f"drop-in run: {calls['load']} loads, {calls['solve']} solves; steps {r['steps']} vs {ref['steps']}")
    assert calls["solve"] == r["newton_iters"] or calls["solve"] >= r["newton_iters"]
    assert abs(r["steps"] - ref["steps"]) <= 0.02 * ref["steps"] + 2
    err = waveform_error(r["t"], r["x"], ref["t"], ref["x"])
    assert err <= 1e-3 * 1.8 + 1e-6, err


def test_gate_array_loads_full_size(gpu):
    """Device loads of the 250k-gate INV/NAND2 array (749,642 level-1
    MOSFETs, 375,323 unknowns) against the oracle's, entry by entry: within
    1e-12 of each entry's magnitude, or, for an entry summed from n_t device
    terms (a supply row collects one from every PMOS), within 2 n_t eps of
    it (the reference's own sequential sum meets no better); three repeated
    loads are bit-identical (fixed summation order)."""
    c = C.gate_dag(250_000, seed=4)
    o = Oracle(c)
    rng = np.random.default_rng(11)
    x = 0.9 + rng.normal(0.0, 0.05, size=o.n)
    sta0, sto0 = np.zeros(max(o.nsta, 1)), np.zeros(max(o.nsto, 1))
    kw = dict(dcop=0, tran=1, init_jct=0, newton_iter=1, order=2, h=2e-12, time=0.0, tstep=1e-12, tstop=1e-9)
    ref = o.load(x, sta0, sto0, **kw)
    p = o.pattern()
    eps = np.finfo(float).eps
    fs, qs = p["mos_fslot"].ravel(), p["mos_qslot"].ravel()
    nt_m = np.bincount(fs[fs >= 0], minlength=o.nnz) + np.bincount(qs[qs >= 0], minlength=o.nnz)
    nt_v = np.bincount(p["mos_lid"].ravel()[p["mos_lid"].ravel() < o.n], minlength=o.n) + np.diff(p["rowptr"])
    d = X.DeviceMgr(c)
    d.set_solver_state(**kw)
    runs, walls = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        d.updateState(x, sta0, sto0)
        runs.append((d.loadDAEVectors(), d.loadDAEMatrices()))
        walls.append(time.perf_counter() - t0)
    print(f"one full load through host arrays (PCIe copies included): {1e3 * min(walls):.1f} ms")
    vec, (dF, dQ) = runs[0]
    for k in ("F", "Q", "dFdxdVp", "dQdxdVp"):
        a, b = vec[k], ref[k]
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        tol = np.maximum(1e-12, 2 * nt_v * eps) * scale + 1e-30
        assert np.all(np.abs(a - b) <= tol), (k, np.max(np.abs(a - b) / tol))
    for a, b in ((dF, ref["dFdx"]), (dQ, ref["dQdx"])):
        scale = np.maximum(np.abs(b), 1e-9 * (np.max(np.abs(b)) + 1e-300))
        tol = np.maximum(1e-12, 2 * nt_m * eps) * scale + 1e-30
        assert np.all(np.abs(a - b) <= tol), np.max(np.abs(a - b) / tol)
    for v2, (dF2, dQ2) in runs[1:]:
        assert all(np.array_equal(vec[k], v2[k]) for k in vec)
        assert np.array_equal(dF, dF2) and np.array_equal(dQ, dQ2)


def _period(t, v, t0=5e-9, level=0.9):
    """Mean interval between rising crossings of ``level`` after ``t0``."""
    m = t > t0
    t, v = t[m], v[m]
    i = np.nonzero((v[:-1] < level) & (v[1:] >= level))[0]
    tc = t[i] + (level - v[i]) * (t[i + 1] - t[i]) / (v[i + 1] - v[i])
    return float(np.mean(np.diff(tc))), len(tc)


def _accurate_edge_shift(c, tstop, cols, t_o, x_o):
    """Largest crossing shift between the oracle's run at the default
    tolerances and an accurate oracle run (RELTOL 1e-5): how far the
    reference's own answer at these tolerances is from the exact waveform."""
    from tests.helpers import edge_shift
    a = Oracle(c)
    a.set_tolerances(tia_reltol=1e-5)
    ra = a.tran(tstop, 1e-12, 10e-12)
    assert ra["status"] == 0
    return edge_shift(t_o, x_o[:, cols], ra["t"], ra["x"][:, cols])


def test_ring101_50ns_matches_oracle(gpu):
    """The reference's own CPU-runnable case: 101-stage ring oscillator,
    .TRAN 0-50 ns with a 10 ps maximum step (about 22k steps, 49k Newton
    iterations), on the GPU engine against the oracle over the whole window:

    * step and Newton counts within 2 %;
    * node voltages within RELTOL |V| + ABSTOL over the first 5 ns;
    * every VDD/2 crossing of every 10th node (72 per node) within one tenth
      of the shift between the oracle's crossings and an accurate run's
      (RELTOL 1e-5; about 23 ps by the end): an oscillator's phase is neutral,
      so last-bit differences (the device's exp and log, the LU's elimination
      order) accumulate into a phase drift, and at RELTOL 1e-3 the reference's
      own drift from the exact waveform is far larger than the two runs'
      distance;
    * with the time axis allowed to move by that crossing shift, the node
      voltages within RELTOL * VDD + ABSTOL over the whole window, not
      counting the interpolation error of the oracle's piecewise-linear
      waveform between its samples (estimated by its distance from the
      quadratic interpolant; up to 11 mV on 10 ps steps near the top of a
      swing), since the two runs' samples fall at different times once their
      step sequences part (after 3.4 ns)."""
    from tests.helpers import edge_shift, waveform_error_tube_interp
    c = C.ring(101, seed=1)
    tstop = 50e-9
    b = X.Batch([c], tstop, 1e-12, 10e-12, record_max=40000)
    cnt = b.run_to_end(chunk=1024)
    assert cnt["done"] == 1 and cnt["failed"] == 0, cnt
    o = Oracle(c)
    r = o.tran(tstop, 1e-12, 10e-12)
    assert r["status"] == 0
    t_g, x_g = b.points(0, o.n)
    st, _ = b.instance(0, o.n)
    cols = list(range(0, o.n, 10))
    early_g, early_o = t_g <= 5e-9, r["t"] <= 5e-9
    early = waveform_error(t_g[early_g], x_g[early_g][:, cols], r["t"][early_o], r["x"][early_o][:, cols])
    shift = edge_shift(t_g, x_g[:, cols], r["t"], r["x"][:, cols])
    acc_shift = _accurate_edge_shift(c, tstop, cols, r["t"], r["x"])
    dt = max(shift, 1e-15)
    tube = waveform_error_tube_interp(t_g, x_g[:, cols], r["t"], r["x"][:, cols], dt)
    col = int(o.pattern()["node_gid"][2])
    pg, ng = _period(t_g, x_g[:, col])
    po, no = _period(r["t"], r["x"][:, col])
    print(f"101-stage ring, 0-50 ns: steps gpu {st['steps']:.0f} oracle {r['steps']}, Newton iterations gpu {st['newton_iters']:.0f} "
          f"oracle {r['newton_iters']}; 0-5 ns max |dV| {early:.2e} V; crossing shift gpu-oracle "
          f"{shift * 1e12:.3f} ps, oracle-accurate {acc_shift * 1e12:.2f} ps; 0-50 ns max |dV| within that shift "
          f"{tube:.2e} V; period gpu {pg * 1e9:.6f} ns oracle {po * 1e9:.6f} ns ({ng} periods)")
    assert abs(st["steps"] - r["steps"]) <= 0.02 * r["steps"] + 2
    assert abs(st["newton_iters"] - r["newton_iters"]) <= 0.02 * r["newton_iters"] + 2
    assert early <= 1e-3 * 1.8 + 1e-6
    assert shift <= 0.1 * acc_shift
    assert tube <= 1e-3 * 1.8 + 1e-6
    assert ng == no and abs(pg - po) <= 1e-4 * po


def test_mc4096_batch_runs_to_the_end_and_matches_oracle(gpu):
    """The headline workload at full size: 4096 Monte Carlo rings (stages odd
    U[11,201], per-device VTO and KP), .TRAN 0-20 ns, hmax 10 ps, in the
    autotuned launch form that the benchmark times.  Every instance finishes.
    64 instances spread evenly over the stage range, run again as a batch of
    their own, give the same step and Newton counts and the same final
    solution bit for bit (instances are independent), as do the smallest, the
    median and the largest run alone.  Against the oracle, for all 64: node
    voltages within RELTOL |V| + ABSTOL at the same times while the two step
    sequences coincide (at least the first 100 points), the same number
    of VDD/2 crossings of every 10th node, and the oscillation period within
    1e-4 after 5 ns.  Over the whole 20 ns an oscillator's phase drifts with
    last-bit differences, and with it the step sequence; the GPU's distance
    from the oracle (step counts, crossing shift, voltages within that shift
    with the oracle's own interpolation error not counted) is held to twice
    the oracle's distance from itself when one model parameter (gmin) moves
    by one part in 10^12, over the same 64 instances (the oracle's crossings
    are several picoseconds from an accurate run's, checked on three)."""
    from concurrent.futures import ThreadPoolExecutor
    from tests.helpers import edge_shift, waveform_error_tube_interp
    cs = C.mc_ring_batch(4096, seed=3)
    tstop, tstep, hmax = 20e-9, 1e-12, 10e-12
    b = X.Batch(cs, tstop, tstep, hmax)
    tune = b.autotune()
    cnt = b.run_to_end(chunk=256)
    print(f"4096 rings, {tune}: {cnt}")
    assert cnt["done"] == 4096 and cnt["failed"] == 0, cnt
    stages = np.array([(len(c.mos_model)) // 2 for c in cs])
    order = np.argsort(stages, kind="stable")
    sample = [int(order[k]) for k in np.linspace(0, len(order) - 1, 64).round().astype(int)]
    assert stages[sample[0]] == 11 and stages[sample[-1]] == 201
    sub = X.Batch([cs[i] for i in sample], tstop, tstep, hmax, record_max=20000)
    csub = sub.run_to_end(chunk=1024)
    assert csub["done"] == 64, csub
    ns = [Oracle(cs[i]).n for i in sample]
    for k, i in enumerate(sample):
        st_b, x_b = b.instance(i, ns[k])
        st_s, x_s = sub.instance(k, ns[k])
        assert st_s["steps"] == st_b["steps"] and st_s["newton_iters"] == st_b["newton_iters"], i
        assert np.array_equal(x_s, x_b), i
    for k in (0, 32, 63):
        i = sample[k]
        b1 = X.Batch([cs[i]], tstop, tstep, hmax)
        assert b1.run_to_end(chunk=1024)["done"] == 1
        st_1, x_1 = b1.instance(0, ns[k])
        st_b, x_b = b.instance(i, ns[k])
        assert st_1["steps"] == st_b["steps"] and np.array_equal(x_1, x_b), i

    def oracle_run(k):
        o = Oracle(cs[sample[k]])
        r = o.tran(tstop, tstep, hmax)
        # the oracle against itself with gmin moved by one part in 1e12
        rp = Oracle(cs[sample[k]], gmin=1e-12 * (1 + 1e-12)).tran(tstop, tstep, hmax)
        cols = list(range(0, o.n, 10))
        sp = edge_shift(rp["t"], rp["x"][:, cols], r["t"], r["x"][:, cols])
        tp = waveform_error_tube_interp(rp["t"], rp["x"][:, cols], r["t"], r["x"][:, cols], max(sp, 1e-15), n_shifts=33)
        self_d = (abs(rp["steps"] - r["steps"]) / r["steps"], sp, tp)
        return r, int(o.pattern()["node_gid"][2]), self_d

    with ThreadPoolExecutor(16) as ex:
        runs = list(ex.map(oracle_run, range(64)))
    worst = dict(steps=0.0, early=0.0, shift=0.0, period=0.0)
    for k, i in enumerate(sample):
        r, col, _ = runs[k]
        assert r["status"] == 0
        st_1, _ = sub.instance(k, ns[k])
        t_g, x_g = sub.points(k, ns[k])
        assert t_g[-1] == pytest.approx(tstop, rel=1e-12)
        cols = list(range(0, ns[k], 10))
        # while the two step sequences coincide, the voltages at the same times
        m = min(len(t_g), len(r["t"]))
        same = np.abs(t_g[:m] - r["t"][:m]) <= 1e-12 * r["t"][:m] + 1e-24
        n_same = int(np.argmin(same)) if not same.all() else m
        early = float(np.max(np.abs(x_g[:n_same] - r["x"][:n_same])))
        shift = edge_shift(t_g, x_g[:, cols], r["t"], r["x"][:, cols])
        tube = waveform_error_tube_interp(t_g, x_g[:, cols], r["t"], r["x"][:, cols], max(shift, 1e-15), n_shifts=33)
        pg, ng = _period(t_g, x_g[:, col])
        po, no = _period(r["t"], r["x"][:, col])
        ds = abs(st_1["steps"] - r["steps"]) / r["steps"]
        worst = dict(steps=max(worst["steps"], ds), early=max(worst["early"], early),
                     shift=max(worst["shift"], shift), tube=max(worst.get("tube", 0.0), tube),
                     period=max(worst["period"], abs(pg - po) / po))
        print(f"instance {i} ({stages[i]} stages): steps gpu {st_1['steps']:.0f} oracle {r['steps']}, Newton "
              f"{st_1['newton_iters']:.0f} / {r['newton_iters']}; same steps to {r['t'][n_same - 1] * 1e9:.3f} ns, "
              f"max |dV| there {early:.2e} V; crossing shift "
              f"{shift * 1e12:.3f} ps; 0-20 ns |dV| within it {tube:.2e} V; period {pg * 1e9:.6f} / {po * 1e9:.6f} ns")
        assert early <= 1e-3 * 1.8 + 1e-6 and n_same >= 100, i
        assert np.isfinite(shift), i
        assert ng == no and abs(pg - po) <= 1e-4 * po, i
    self_w = np.max(np.array([run[2] for run in runs]), axis=0)
    print(f"worst over the 64: {worst}; the oracle against itself (gmin x (1 + 1e-12)): step counts "
          f"{self_w[0]:.2%}, crossing shift {self_w[1] * 1e12:.3f} ps, voltages {self_w[2]:.2e} V")
    assert worst["steps"] <= max(0.02, 2 * self_w[0])
    assert worst["shift"] <= max(0.2e-12, 2 * self_w[1])
    assert worst["tube"] <= max(1e-3 * 1.8 + 1e-6, 2 * self_w[2])

    def accurate(k):
        r, _, _ = runs[k]
        return _accurate_edge_shift(cs[sample[k]], tstop, list(range(0, ns[k], 10)), r["t"], r["x"])

    with ThreadPoolExecutor(3) as ex:
        accs = list(ex.map(accurate, (0, 32, 63)))
    for k, acc in zip((0, 32, 63), accs):
        print(f"instance {sample[k]}: oracle crossings {acc * 1e12:.2f} ps from an accurate run's")
        assert acc >= 3e-12


