"""The drop-in boundary as the reference's callers use it.

* Accumulation.  Xyce's loader zeroes the DAE vectors and matrices and every
  device master then adds its contributions with "+="
  (N_DEV_MOSFET1.C:4393-4411, :4587-4868; N_LOA_NonlinearEquationLoader.C:
  385-390,525-526), so a master loaded after another device type must not
  wipe that device's entries.  ``xg_load_vectors`` / ``xg_load_matrices``
  add into the caller's arrays: pre-filled arrays come back as
  prefill + contribution, bit for bit (one rounding per entry).
  ``xg_update_state`` writes only the MOSFETs' state and store entries.
* Device-pointer variants (``*_dev``): the same calls on arrays in device
  memory give the same bits as the host-array calls, so a caller that keeps
  its vectors on the GPU pays no host copies; on the 250k-gate array the full
  load (state update, vectors, matrices) is timed both ways.
"""
import ctypes
import time

import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


def _state(o, seed):
    st, x0, sta0, sto0, Q, F, B = o.dcop()
    assert st > 0
    rng = np.random.default_rng(seed)
    return x0 + rng.normal(0.0, 0.3, size=o.n), sta0, sto0


def test_loads_add_into_the_callers_arrays(gpu):
    c = C.ring(11, seed=5, mc_seed=5)
    o = Oracle(c)
    x, sta, sto = _state(o, 1)
    d = X.DeviceMgr(c)
    d.set_solver_state(dcop=0, tran=1, newton_iter=1, order=2, h=2e-12, time=1e-10)
    d.updateState(x, sta, sto)
    base_v = d.loadDAEVectors()
    base_m = d.loadDAEMatrices()
    assert all(np.any(base_v[k]) for k in ("F", "Q", "B"))
    rng = np.random.default_rng(2)
    pre_v = {k: rng.normal(0.0, 1e-3, d.n) for k in base_v}
    pre_m = (rng.normal(0.0, 1e-3, d.nnz), rng.normal(0.0, 1e-15, d.nnz))
    got_v = d.loadDAEVectors(into={k: v.copy() for k, v in pre_v.items()})
    got_m = d.loadDAEMatrices(into=(pre_m[0].copy(), pre_m[1].copy()))
    for k in base_v:
        assert np.array_equal(got_v[k], pre_v[k] + base_v[k]), k
    for g, p0, b in zip(got_m, pre_m, base_m):
        assert np.array_equal(g, p0 + b)
    # a second load adds again: loads are pure functions of the device state
    again = d.loadDAEVectors(into={k: v.copy() for k, v in got_v.items()})
    assert np.array_equal(again["F"], got_v["F"] + base_v["F"])


def test_update_state_writes_only_the_mosfet_entries(gpu):
    c = C.ring(11, seed=5, mc_seed=5)
    o = Oracle(c)
    x, sta, sto = _state(o, 3)
    d = X.DeviceMgr(c)
    p = d.pattern()
    d.set_solver_state(dcop=0, tran=1, newton_iter=1, order=1, h=1e-12, time=1e-10)
    sn_ref, on_ref, lim = d.updateState(x, sta, sto)
    own_sta = np.zeros(d.nsta, bool)
    own_sto = np.zeros(d.nsto, bool)
    for s0, o0 in p["mos_stasto"]:
        own_sta[s0:s0 + 8] = True
        own_sto[o0:o0 + 6] = True
    assert not own_sta.all()                       # the capacitors own state entries too
    sentinel = np.full(d.nsta, 7.25)
    sn = sentinel.copy()
    on = sto.copy()
    L = X.lib()
    limc = ctypes.c_int(0)
    assert L.xg_update_state(d.h, x.ctypes.data_as(P), sta.ctypes.data_as(P), sto.ctypes.data_as(P),
                             sn.ctypes.data_as(P), on.ctypes.data_as(P), ctypes.byref(limc)) == 0
    assert np.array_equal(sn[own_sta], sn_ref[own_sta])
    assert np.all(sn[~own_sta] == 7.25)
    assert np.array_equal(on, on_ref)


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("make", [lambda: C.ring(101, seed=2, mc_seed=2), lambda: C.gate_dag(3000, seed=4)],
                         ids=["ring101", "gates3000"])
def test_device_pointer_variants_equal_the_host_calls(gpu, make):
    import torch
    c = make()
    o = Oracle(c)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 1.8, o.n)
    sta = rng.normal(0.0, 1e-15, max(o.nsta, 1))
    sto = rng.uniform(-0.5, 1.8, max(o.nsto, 1))
    d = X.DeviceMgr(c)
    kw = dict(dcop=0, tran=1, newton_iter=1, order=2, h=2e-12, time=3e-10)
    d.set_solver_state(**kw)
    sn, on, lim = d.updateState(x, sta, sto, sto_next=sto)
    rng2 = np.random.default_rng(5)
    pre_v = {k: rng2.normal(0.0, 1e-3, d.n) for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp")}
    pre_m = (rng2.normal(0.0, 1e-3, d.nnz), rng2.normal(0.0, 1e-15, d.nnz))
    hv = d.loadDAEVectors(into={k: v.copy() for k, v in pre_v.items()})
    hm = d.loadDAEMatrices(into=(pre_m[0].copy(), pre_m[1].copy()))
    q0 = rng2.normal(0.0, 1e-15, d.n)
    q2 = rng2.normal(0.0, 1e-6, d.n)
    h_rhs, h_J = d.combine(q0, q2)

    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    xs, sas, sos = T(x), T(sta), T(sto)
    sn_d, on_d = torch.zeros_like(sas), T(sto)
    L = X.lib()
    limc = ctypes.c_int(0)
    torch.cuda.synchronize()
    assert L.xg_update_state_dev(d.h, _dptr(xs), _dptr(sas), _dptr(sos), _dptr(sn_d), _dptr(on_d),
                                 ctypes.byref(limc)) == 0
    assert limc.value == lim
    own = np.zeros(len(sta), bool)
    for s0, _ in d.pattern()["mos_stasto"]:
        own[s0:s0 + 8] = True
    assert np.array_equal(sn_d.cpu().numpy()[own], sn[own])
    assert np.array_equal(on_d.cpu().numpy()[: d.nsto], on)
    dv = {k: T(v) for k, v in pre_v.items()}
    assert L.xg_load_vectors_dev(d.h, *[_dptr(dv[k]) for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp")]) == 0
    dm = (T(pre_m[0]), T(pre_m[1]))
    assert L.xg_load_matrices_dev(d.h, _dptr(dm[0]), _dptr(dm[1])) == 0
    for k in dv:
        assert np.array_equal(dv[k].cpu().numpy(), hv[k]), k
    for a, b in zip(dm, hm):
        assert np.array_equal(a.cpu().numpy(), b)
    rhs_d, J_d = torch.zeros(d.n, dtype=torch.float64, device=dev), torch.zeros(d.nnz, dtype=torch.float64, device=dev)
    q0d, q2d = T(q0), T(q2)
    assert L.xg_combine_dev(d.h, _dptr(q0d), _dptr(q2d), _dptr(rhs_d), _dptr(J_d)) == 0
    assert np.array_equal(rhs_d.cpu().numpy(), h_rhs) and np.array_equal(J_d.cpu().numpy(), h_J)
    # the LU on device arrays: same factors, same solution bits
    p = d.pattern()
    s_h = X.DirectSolver(d.n, p["rowptr"], p["colind"], h_J)
    s_d = X.DirectSolver(d.n, p["rowptr"], p["colind"], h_J)
    assert s_h.numeric(h_J) == 0
    st, x_h = s_h.solve(h_rhs)
    assert st == 0
    assert L.xg_lu_numeric_dev(s_d.h, _dptr(J_d)) == 0
    xd = torch.zeros(d.n, dtype=torch.float64, device=dev)
    assert L.xg_lu_solve_dev(s_d.h, _dptr(rhs_d), _dptr(xd), 0) == 0
    assert np.array_equal(xd.cpu().numpy(), x_h)
    # a NaN right-hand side: status and a zero solution, as on the host path
    bad = rhs_d.clone()
    bad[3] = float("nan")
    assert L.xg_lu_solve_dev(s_d.h, _dptr(bad), _dptr(xd), 0) == -99
    assert torch.all(xd == 0)


def test_gate_array_full_load_host_and_device_arrays(gpu):
    """Full load of the 250k-gate array (state update, vectors, matrices):
    through host arrays the PCIe copies dominate; through device arrays only
    the kernels and the accumulation remain.  Both give the same bits."""
    import torch
    c = C.gate_dag(250_000, seed=4)
    d = X.DeviceMgr(c)
    rng = np.random.default_rng(11)
    x = 0.9 + rng.normal(0.0, 0.05, size=d.n)
    sta = np.zeros(max(d.nsta, 1))
    sto = np.zeros(max(d.nsto, 1))
    d.set_solver_state(dcop=0, tran=1, newton_iter=1, order=2, h=2e-12, time=0.0)
    L = X.lib()
    limc = ctypes.c_int(0)
    hv = [np.zeros(d.n) for _ in range(5)]
    hm = [np.zeros(d.nnz) for _ in range(2)]
    sn, on = np.zeros_like(sta), sto.copy()

    def host_load():
        for a in hv + hm:
            a[:] = 0.0
        assert L.xg_update_state(d.h, x.ctypes.data_as(P), sta.ctypes.data_as(P), sto.ctypes.data_as(P),
                                 sn.ctypes.data_as(P), on.ctypes.data_as(P), ctypes.byref(limc)) == 0
        assert L.xg_load_vectors(d.h, *[a.ctypes.data_as(P) for a in hv]) == 0
        assert L.xg_load_matrices(d.h, *[a.ctypes.data_as(P) for a in hm]) == 0

    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    xs, sas, sos, snd, ond = T(x), T(sta), T(sto), T(sta), T(sto)
    dvv = [torch.zeros(d.n, dtype=torch.float64, device=dev) for _ in range(5)]
    dmm = [torch.zeros(d.nnz, dtype=torch.float64, device=dev) for _ in range(2)]

    def dev_load():
        for a in dvv + dmm:
            a.zero_()
        torch.cuda.synchronize()
        assert L.xg_update_state_dev(d.h, _dptr(xs), _dptr(sas), _dptr(sos), _dptr(snd), _dptr(ond),
                                     ctypes.byref(limc)) == 0
        assert L.xg_load_vectors_dev(d.h, *[_dptr(a) for a in dvv]) == 0
        assert L.xg_load_matrices_dev(d.h, *[_dptr(a) for a in dmm]) == 0

    th, td = [], []
    for _ in range(4):
        t0 = time.perf_counter(); host_load(); th.append(time.perf_counter() - t0)
        t0 = time.perf_counter(); dev_load(); td.append(time.perf_counter() - t0)
    print(f"250k-gate array ({d.n_mos} MOSFETs, n={d.n}, nnz={d.nnz}): one full load through host arrays "
          f"{1e3 * min(th):.2f} ms, through device arrays {1e3 * min(td):.2f} ms")
    for a, b in zip(dvv, hv):
        assert np.array_equal(a.cpu().numpy(), b)
    for a, b in zip(dmm, hm):
        assert np.array_equal(a.cpu().numpy(), b)
    assert min(td) < min(th)
