"""xyce_amd — MI355X-native per-Newton hot path of a Xyce transient run.

Python mirror of the reference interfaces this library replaces, over the
C ABI of ``libxyce_amd.so`` (declared in ``include/xyce_amd.h``):

* :class:`DeviceMgr` — ``DeviceMgr::updateState / loadDAEVectors /
  loadDAEMatrices`` and ``allDevicesConverged``
  (src/DeviceModelPKG/Core/N_DEV_DeviceMgr.C:3857,4156,3980,5603), for the
  MOSFET level-1 / R / C / V / I devices of one circuit.
* :class:`DirectSolver` — ``Linear::Solver::doSolve(reuse_factors, transpose)``
  as the KLU wrapper implements it (src/LinearAlgebraServicesPKG/
  N_LAS_AmesosSolver.C:216-478): symbolic analysis once, numeric refactor on
  every call unless ``reuse_factors``, solve; a singular factor gives a zero
  solution and a nonzero status.
* :class:`Batch` — thousands of independent circuits, each running its own
  DC operating point and OneStep transient, one launch per Newton iteration.

There is no CPU fallback: every entry point fails loudly when the HIP library
or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import circuits  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxyce_amd.so")
_LIB = None

XG_OK = 0
ERRORS = {-1: "bad argument", -2: "HIP runtime error", -3: "circuit setup failed", -99: "NaN in solution",
          -10: "no HIP device"}

EXPORTS = [
    "xg_circuit_create", "xg_circuit_destroy", "xg_circuit_sizes", "xg_circuit_pattern", "xg_set_solver_state",
    "xg_update_state", "xg_load_vectors", "xg_load_matrices", "xg_combine",
    "xg_update_state_dev", "xg_load_vectors_dev", "xg_load_matrices_dev", "xg_combine_dev",
    "xg_lu_symbolic", "xg_lu_symbolic_ex", "xg_lu_destroy", "xg_lu_numeric", "xg_lu_solve", "xg_lu_info",
    "xg_lu_numeric_dev", "xg_lu_solve_dev", "xg_lu_solve_noalloc",
    "xg_lu_set_precision", "xg_lu_precision_stats",
    "xg_batch_create", "xg_batch_destroy", "xg_batch_advance", "xg_batch_counters", "xg_batch_instance",
    "xg_batch_points", "xg_batch_autotune", "xg_batch_config", "xg_batch_set_method", "xg_batch_dc_strategy", "xg_batch_dc_stage", "xg_batch_group_info", "xg_batch_profile", "xg_host_pattern", "xg_host_lu_analyse", "xg_host_mf_plan",
    "xg_mlp_create", "xg_mlp_destroy", "xg_mlp_params", "xg_mlp_forward", "xg_mlp_loss_backward", "xg_mlp_adam",
    "xg_mlp_sync", "xg_batch_predictor", "xg_batch_predictor_options", "xg_batch_pairs", "xg_batch_pred_counters",
    "xg_tran_create", "xg_tran_destroy", "xg_tran_set_precision", "xg_tran_set_tolerances", "xg_tran_set_record",
    "xg_tran_run", "xg_tran_points", "xg_tran_solution", "xg_tran_set_progress", "xg_tran_set_uic",
]

PRED_WIN = 256      # learned predictor: unknowns per window (XG_PRED_WIN)
PRED_FEAT = 1600    # features per window (XG_PRED_FEAT)


class XgError(RuntimeError):
    pass


class CircuitDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int),
        ("n_dev", C.c_int), ("dev_type", C.c_void_p), ("dev_idx", C.c_void_p),
        ("n_models", C.c_int), ("model_par", C.c_void_p),
        ("n_mos", C.c_int), ("mos_nodes", C.c_void_p), ("mos_model", C.c_void_p), ("mos_inst", C.c_void_p),
        ("n_res", C.c_int), ("res_nodes", C.c_void_p), ("res_val", C.c_void_p),
        ("n_cap", C.c_int), ("cap_nodes", C.c_void_p), ("cap_val", C.c_void_p),
        ("n_vsrc", C.c_int), ("vsrc_nodes", C.c_void_p), ("vsrc_wave", C.c_void_p),
        ("n_isrc", C.c_int), ("isrc_nodes", C.c_void_p), ("isrc_wave", C.c_void_p),
        ("n_wave", C.c_int), ("wave_type", C.c_void_p), ("wave_off", C.c_void_p), ("wave_data", C.c_void_p),
        ("gmin", C.c_double), ("temp", C.c_double),
    ]


class SolverState(C.Structure):
    _fields_ = [("dcop", C.c_int), ("tran", C.c_int), ("init_jct", C.c_int), ("init_fix", C.c_int),
                ("newton_iter", C.c_int), ("order", C.c_int), ("h", C.c_double), ("time", C.c_double),
                ("tstep", C.c_double), ("tstop", C.c_double), ("tranop", C.c_int)]


def lib():
    """Load libxyce_amd.so (built by ``make -C xyce_amd/csrc``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise XgError(f"HIP library missing: {LIB_PATH}; build it with make -C xyce_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None or a.size == 0 else a.ctypes.data_as(C.c_void_p)


def _check(st, what):
    if st < 0:
        raise XgError(f"{what}: {ERRORS.get(st, st)}")
    return st


def make_desc(circuit, gmin=1e-12, temp=300.15):
    """Flat arrays of a :class:`circuits.Circuit` as an ``xg_circuit_desc``.
    Returns (desc, keepalive)."""
    a = circuit.arrays()
    d = CircuitDesc()
    d.n_nodes = a["n_nodes"]
    d.n_dev = len(a["dev_type"]); d.dev_type = _p(a["dev_type"]); d.dev_idx = _p(a["dev_idx"])
    d.n_models = a["n_models"]; d.model_par = _p(a["model_par"])
    d.n_mos = a["n_mos"]; d.mos_nodes = _p(a["mos_nodes"]); d.mos_model = _p(a["mos_model"]); d.mos_inst = _p(a["mos_inst"])
    d.n_res = a["n_res"]; d.res_nodes = _p(a["res_nodes"]); d.res_val = _p(a["res_val"])
    d.n_cap = a["n_cap"]; d.cap_nodes = _p(a["cap_nodes"]); d.cap_val = _p(a["cap_val"])
    d.n_vsrc = a["n_vsrc"]; d.vsrc_nodes = _p(a["vsrc_nodes"]); d.vsrc_wave = _p(a["vsrc_wave"])
    d.n_isrc = a["n_isrc"]; d.isrc_nodes = _p(a["isrc_nodes"]); d.isrc_wave = _p(a["isrc_wave"])
    d.n_wave = a["n_wave"]; d.wave_type = _p(a["wave_type"]); d.wave_off = _p(a["wave_off"]); d.wave_data = _p(a["wave_data"])
    d.gmin = gmin; d.temp = temp
    return d, a


def host_pattern(circuit, gmin=1e-12, temp=300.15):
    """Pattern and stamp slots from the host setup (no GPU)."""
    d, keep = make_desc(circuit, gmin, temp)
    sizes = np.zeros(6, np.int32)
    _check(lib().xg_host_pattern(C.byref(d), _p(sizes), None, None, None, None, None, None, None), "xg_host_pattern")
    n, nnz, nsta, nsto, n_mos, n_nodes = (int(v) for v in sizes)
    rp = np.zeros(n + 1, np.int32); ci = np.zeros(max(nnz, 1), np.int32); ng = np.zeros(n_nodes, np.int32)
    ml = np.zeros(max(6 * n_mos, 1), np.int32); fs = np.zeros(max(17 * n_mos, 1), np.int32)
    qs = np.zeros(max(14 * n_mos, 1), np.int32); ss = np.zeros(max(2 * n_mos, 1), np.int32)
    _check(lib().xg_host_pattern(C.byref(d), _p(sizes), _p(rp), _p(ci), _p(ng), _p(ml), _p(fs), _p(qs), _p(ss)),
           "xg_host_pattern")
    return dict(n=n, nnz=nnz, nsta=nsta, nsto=nsto, rowptr=rp, colind=ci[:nnz], node_gid=ng,
                mos_lid=ml[: 6 * n_mos].reshape(-1, 6), mos_fslot=fs[: 17 * n_mos].reshape(-1, 17),
                mos_qslot=qs[: 14 * n_mos].reshape(-1, 14), mos_stasto=ss[: 2 * n_mos].reshape(-1, 2))


def host_lu_analyse(n, rowptr, colind, values=None):
    """Symbolic LU analysis on the host (no GPU): permutations and L+U pattern."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colind = np.ascontiguousarray(colind, np.int32)
    v = None if values is None else np.ascontiguousarray(values, np.float64)
    info = np.zeros(4)
    pr = np.zeros(n, np.int32); pc = np.zeros(n, np.int32); lrp = np.zeros(n + 1, np.int32)
    _check(lib().xg_host_lu_analyse(n, _p(rowptr), _p(colind), _p(v), _p(info), _p(pr), _p(pc), _p(lrp), None, 0),
           "xg_host_lu_analyse")
    nlu = int(info[0])
    lci = np.zeros(nlu, np.int32)
    _check(lib().xg_host_lu_analyse(n, _p(rowptr), _p(colind), _p(v), _p(info), _p(pr), _p(pc), _p(lrp), _p(lci), nlu),
           "xg_host_lu_analyse")
    return dict(nnz_lu=nlu, factor_levels=int(info[1]), solve_levels=int(info[2]), flops=float(info[3]),
                perm_row=pr, perm_col=pc, lu_rowptr=lrp, lu_colind=lci)


class DeviceMgr:
    """One circuit's devices on the GPU (the drop-in behind DeviceMgr)."""

    def __init__(self, circuit, gmin=1e-12, temp=300.15):
        self._desc, self._keep = make_desc(circuit, gmin, temp)
        h = C.c_void_p()
        _check(lib().xg_circuit_create(C.byref(self._desc), C.byref(h)), "xg_circuit_create")
        self.h = h
        s = np.zeros(6, np.int32)
        lib().xg_circuit_sizes(self.h, _p(s))
        self.n, self.nnz, self.nsta, self.nsto, self.n_mos, self.n_nodes = (int(v) for v in s)

    def close(self):
        if getattr(self, "h", None):
            lib().xg_circuit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        ng = np.zeros(self.n_nodes, np.int32)
        ml = np.zeros(max(6 * self.n_mos, 1), np.int32)
        fs = np.zeros(max(17 * self.n_mos, 1), np.int32)
        qs = np.zeros(max(14 * self.n_mos, 1), np.int32)
        ss = np.zeros(max(2 * self.n_mos, 1), np.int32)
        _check(lib().xg_circuit_pattern(self.h, _p(rp), _p(ci), _p(ng), _p(ml), _p(fs), _p(qs), _p(ss)), "pattern")
        return dict(rowptr=rp, colind=ci[: self.nnz], node_gid=ng, mos_lid=ml[: 6 * self.n_mos].reshape(-1, 6),
                    mos_fslot=fs[: 17 * self.n_mos].reshape(-1, 17), mos_qslot=qs[: 14 * self.n_mos].reshape(-1, 14),
                    mos_stasto=ss[: 2 * self.n_mos].reshape(-1, 2))

    def set_solver_state(self, *, dcop=0, tran=1, init_jct=0, newton_iter=1, order=1, h=1e-11, time=0.0,
                         tstep=1e-12, tstop=1e-9, tranop=0):
        ss = SolverState(dcop, tran, init_jct, 0, newton_iter, order, h, time, tstep, tstop, tranop)
        _check(lib().xg_set_solver_state(self.h, C.byref(ss)), "set_solver_state")
        self._ss = ss

    def updateState(self, x, sta_cur, sto_cur, sto_next=None):
        """Returns (sta_next, sto_next, limited)."""
        x = np.ascontiguousarray(x, np.float64)
        sc = np.ascontiguousarray(sta_cur, np.float64)
        oc = np.ascontiguousarray(sto_cur, np.float64)
        sn = np.zeros(max(self.nsta, 1))
        on = np.array(oc if sto_next is None else sto_next, np.float64)
        if on.size == 0:
            on = np.zeros(1)
        lim = C.c_int(0)
        _check(lib().xg_update_state(self.h, _p(x), _p(sc), _p(oc), _p(sn), _p(on), C.byref(lim)), "updateState")
        return sn[: self.nsta], on[: self.nsto], int(lim.value)

    def loadDAEVectors(self, into=None):
        """This circuit's F, Q, B and limiter vectors.  ``into``: a dict of
        the five arrays to add them to (the reference's "+=" into vectors
        that other devices share); by default zero arrays."""
        names = ("F", "Q", "B", "dFdxdVp", "dQdxdVp")
        out = [np.zeros(self.n) for _ in names] if into is None else [into[k] for k in names]
        assert all(v.dtype == np.float64 and v.flags.c_contiguous and v.size == self.n for v in out)
        _check(lib().xg_load_vectors(self.h, *[_p(v) for v in out]), "loadDAEVectors")
        return dict(zip(names, out))

    def loadDAEMatrices(self, into=None):
        """dFdx and dQdx CSR values, added to ``into`` = (dF, dQ) if given."""
        dF, dQ = (np.zeros(self.nnz), np.zeros(self.nnz)) if into is None else into
        assert all(v.dtype == np.float64 and v.flags.c_contiguous and v.size == self.nnz for v in (dF, dQ))
        _check(lib().xg_load_matrices(self.h, _p(dF), _p(dQ)), "loadDAEMatrices")
        return dF, dQ

    def combine(self, q0, q2=None):
        rhs = np.zeros(self.n)
        J = np.zeros(self.nnz)
        q0 = None if q0 is None else np.ascontiguousarray(q0, np.float64)
        q2 = None if q2 is None else np.ascontiguousarray(q2, np.float64)
        _check(lib().xg_combine(self.h, _p(q0), _p(q2), _p(rhs), _p(J)), "combine")
        return rhs, J


class DirectSolver:
    """KLU-like direct solver on the GPU (symbolic once, refactor every call)."""

    def __init__(self, n, rowptr, colind, values=None, method=0):
        """method: 0 automatic, 1 level-scheduled entry LU, 2 multifrontal."""
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colind = np.ascontiguousarray(colind, np.int32)
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        h = C.c_void_p()
        _check(lib().xg_lu_symbolic_ex(self.n, _p(self.rowptr), _p(self.colind), _p(v), int(method), C.byref(h)),
               "xg_lu_symbolic")
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().xg_lu_destroy(self.h)
        except Exception:
            pass

    def info(self):
        i = np.zeros(6)
        lib().xg_lu_info(self.h, _p(i))
        return dict(nnz_lu=int(i[0]), factor_levels=int(i[1]), solve_levels=int(i[2]), flops=float(i[3]),
                    method=int(i[4]), supernodes=int(i[5]))

    def set_precision(self, precision, ir_tol=1e-12, ir_max=10):
        """0 = FP64 factors; 1 = FP32 factors with FP64 iterative refinement."""
        _check(lib().xg_lu_set_precision(self.h, int(precision), C.c_double(ir_tol), int(ir_max)),
               "xg_lu_set_precision")

    def precision_stats(self):
        s = np.zeros(7)
        _check(lib().xg_lu_precision_stats(self.h, _p(s)), "xg_lu_precision_stats")
        return dict(last_corrections=int(s[0]), last_rel_residual=float(s[1]), solves=int(s[2]),
                    corrections=int(s[3]), fp64_fallbacks=int(s[4]), fp32_pivot_failures=int(s[5]),
                    precision=int(s[6]))

    def numeric(self, values):
        v = np.ascontiguousarray(values, np.float64)
        return _check(lib().xg_lu_numeric(self.h, _p(v)), "xg_lu_numeric")

    def solve(self, b, transpose=False):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        st = lib().xg_lu_solve(self.h, _p(b), _p(x), int(bool(transpose)))
        if st == -99:
            return st, x
        _check(st, "xg_lu_solve")
        return st, x

    def doSolve(self, values, b, reuse_factors=False, transpose=False):
        """Linear::Solver::doSolve: returns (status, x); status != 0 on failure;
        transpose solves A^T x = b (level-scheduled method, FP64 factors)."""
        if not reuse_factors:
            st = self.numeric(values)
            if st > 0:
                return st, np.zeros(self.n)
        return self.solve(b, transpose)


class Batch:
    """Batched independent circuits (one workgroup each) on the GPU."""

    def __init__(self, circuits_list, tstop, tstep, hmax, record_max=0, gmin=1e-12, temp=300.15, method="trap"):
        """``method``: 'trap' (OneStep trapezoid, the default) or 'gear'
        (Gear12, .OPTIONS TIMEINT METHOD=GEAR)."""
        self._keep = []
        descs = (CircuitDesc * len(circuits_list))()
        for i, c in enumerate(circuits_list):
            d, k = make_desc(c, gmin, temp)
            descs[i] = d
            self._keep.append(k)
        self._descs = descs
        h = C.c_void_p()
        _check(lib().xg_batch_create(len(circuits_list), descs, C.c_double(tstop), C.c_double(tstep),
                                     C.c_double(hmax), C.c_int(record_max), C.byref(h)), "xg_batch_create")
        self.h = h
        self.n_inst = len(circuits_list)
        self.record_max = record_max
        m = {"trap": 0, "gear": 1}[method]
        if m:
            _check(lib().xg_batch_set_method(self.h, C.c_int(m)), "xg_batch_set_method")

    def close(self):
        if getattr(self, "h", None):
            lib().xg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, launches, graph_chunk=0):
        _check(lib().xg_batch_advance(self.h, C.c_int(launches), C.c_int(graph_chunk)), "xg_batch_advance")

    def counters(self):
        c = np.zeros(9)
        _check(lib().xg_batch_counters(self.h, _p(c)), "xg_batch_counters")
        keys = ("newton_iters", "loads", "steps", "rejects", "done", "failed", "launches", "seconds", "dense_solves")
        return dict(zip(keys, c.tolist()))

    def instance(self, i, n):
        ts = np.zeros(8)
        x = np.zeros(n)
        _check(lib().xg_batch_instance(self.h, C.c_int(i), _p(ts), _p(x)), "xg_batch_instance")
        keys = ("t", "h", "order", "mode", "newton_iters", "steps", "rejects", "loads")
        return dict(zip(keys, ts.tolist())), x

    def dc_strategy(self, first_stage):
        """Start every instance's DC operating point at 'newton' (the
        reference's order: Newton, gmin stepping, source stepping), 'gmin' or
        'source' stepping; before the first launch."""
        st = {"newton": 0, "gmin": 1, "source": 2}[first_stage]
        _check(lib().xg_batch_dc_strategy(self.h, C.c_int(st)), "xg_batch_dc_strategy")

    def dc_stage(self, i):
        """The DC strategy stage that produced instance i's operating point."""
        v = C.c_int(0)
        _check(lib().xg_batch_dc_stage(self.h, C.c_int(i), C.byref(v)), "xg_batch_dc_stage")
        return ("newton", "gmin", "source")[v.value]

    def points(self, i, n):
        t = np.zeros(self.record_max)
        x = np.zeros(self.record_max * n)
        cnt = lib().xg_batch_points(self.h, C.c_int(i), C.c_int(self.record_max), _p(t), _p(x))
        _check(cnt, "xg_batch_points")
        return t[:cnt], x[: cnt * n].reshape(cnt, n)

    def run_to_end(self, chunk=256, max_launches=2_000_000):
        total = 0
        while total < max_launches:
            self.advance(chunk, graph_chunk=0)
            total += chunk
            c = self.counters()
            if c["done"] + c["failed"] >= self.n_inst:
                break
        return self.counters()

    def set_predictor(self, mode, net=None, pair_capacity=0, base="poly", deadzone=0.0):
        """Learned predictor: mode 0 off, 1 collect training pairs, 2 serve `net` (an MLP).
        The network corrects ``base``: 'poly' (the polynomial predictor) or 'cubic'
        (the cubic through the last four points); corrections of at most
        ``deadzone`` Newton weights are not applied."""
        self._net = net
        _check(lib().xg_batch_predictor_options(self.h, {"poly": 0, "cubic": 1}[base], C.c_double(deadzone)),
               "xg_batch_predictor_options")
        _check(lib().xg_batch_predictor(self.h, int(mode), net.h if net is not None else None, int(pair_capacity)),
               "xg_batch_predictor")

    def pairs(self, max_records=None):
        """Collected training pairs as CUDA tensors (x bf16 [r][PRED_FEAT], t, m float32 [r][PRED_WIN])."""
        import torch
        cnt = C.c_longlong()
        _check(lib().xg_batch_pairs(self.h, 0, None, None, None, C.byref(cnt)), "xg_batch_pairs")
        n = cnt.value if max_records is None else min(cnt.value, int(max_records))
        x = torch.zeros((n, PRED_FEAT), dtype=torch.bfloat16, device="cuda")
        t = torch.zeros((n, PRED_WIN), dtype=torch.float32, device="cuda")
        m = torch.zeros((n, PRED_WIN), dtype=torch.float32, device="cuda")
        if n:
            _check(lib().xg_batch_pairs(self.h, C.c_longlong(n), C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                        C.c_void_p(m.data_ptr()), C.byref(cnt)), "xg_batch_pairs")
        return x, t, m

    def pred_counters(self):
        c = np.zeros(5)
        _check(lib().xg_batch_pred_counters(self.h, _p(c)), "xg_batch_pred_counters")
        return dict(pairs=int(c[0]), learned_steps=int(c[1]), learned_newton_failures=int(c[2]),
                    windows=int(c[3]), mode=int(c[4]))

    def config(self, block=0, iters_per_launch=0, split_eval=-1, concurrent_groups=-1):
        """block: workgroup size (0 keep); split_eval 1: separate evaluation kernel, one iteration per
        launch; 0: evaluation fused, iters_per_launch iterations per launch; concurrent_groups 1: the
        LDS-sized instance groups side by side on their own streams, each running the launches of one
        advance (or captured graph) back to back; -1 keep."""
        _check(lib().xg_batch_config(self.h, C.c_int(block), C.c_int(iters_per_launch), C.c_int(split_eval),
                                     C.c_int(concurrent_groups)), "xg_batch_config")

    PHASES = ("x_load", "device_eval", "vector_gather_residual", "norms", "newton_decision", "jacobian_gather",
              "lu_refactor", "pivot_check", "triangular_solves", "update", "step_control")

    def profile(self, enable=None):
        """enable=True/False starts/stops the per-phase timers; returns the
        accumulated microseconds per phase of one instance."""
        us = np.zeros(16)
        flag = -1 if enable is None else (1 if enable else 0)
        _check(lib().xg_batch_profile(self.h, C.c_int(flag), _p(us)), "xg_batch_profile")
        out = {k: float(us[i]) for i, k in enumerate(self.PHASES)}
        out["iterations"] = int(us[15])
        return out

    def autotune(self):
        r = np.zeros(10)
        _check(lib().xg_batch_autotune(self.h, _p(r)), "xg_batch_autotune")
        return dict(block=int(r[0]), order_mode=int(r[1]), us_per_iteration=float(r[2]), split_eval=int(r[3]),
                    iters_per_launch=int(r[4]), group_blocks=[int(v) for v in r[5:9] if v > 0],
                    concurrent_groups=int(r[9]), group_terms_in_lds=[g["terms_in_lds"] for g in self.groups()])

    def groups(self):
        """Launch groups: instances, workgroup size, term arrays in LDS, LDS
        bytes, LDS bytes with the term arrays in LDS."""
        out = []
        info = np.zeros(5, np.int32)
        g = 0
        while True:
            cnt = lib().xg_batch_group_info(self.h, C.c_int(g), _p(info))
            _check(min(cnt, 0), "xg_batch_group_info")
            if g >= cnt:
                return out
            out.append(dict(instances=int(info[0]), block=int(info[1]), terms_in_lds=int(info[2]), lds_bytes=int(info[3]),
                            lds_bytes_with_terms=int(info[4])))
            g += 1


class MLP:
    """Learned Newton-predictor network on the GPU: BF16 weights and
    activations, FP32 accumulation on the matrix cores, hand-written forward,
    backward and Adam kernels (xg_mlp_*).  Device buffers are passed as torch
    CUDA tensors (bf16 inputs, float32 targets/masks/outputs)."""

    def __init__(self, in_dim, width=512, hidden_layers=3, out_dim=256, max_rows=4096, seed=5):
        self.in_dim, self.width, self.hidden, self.out_dim = in_dim, width, hidden_layers, out_dim
        self.max_rows = max_rows
        h = C.c_void_p()
        _check(lib().xg_mlp_create(in_dim, width, hidden_layers, out_dim, max_rows, C.c_ulonglong(seed),
                                   C.byref(h)), "xg_mlp_create")
        self.h = h
        cnt = C.c_longlong()
        _check(lib().xg_mlp_params(self.h, 0, None, C.byref(cnt)), "xg_mlp_params")
        self.n_params = cnt.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().xg_mlp_destroy(self.h)
        except Exception:
            pass

    def dims(self):
        return [(self.in_dim if l == 0 else self.width, self.width if l < self.hidden else self.out_dim)
                for l in range(self.hidden + 1)]

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib().xg_mlp_params(self.h, 0, _p(p), None), "xg_mlp_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(lib().xg_mlp_params(self.h, 1, _p(p), None), "xg_mlp_params")

    @staticmethod
    def _dp(t):
        assert t.is_cuda and t.is_contiguous()
        return C.c_void_p(t.data_ptr())

    def forward(self, x, y):
        """y[rows][out] (float32, cuda) = net(x[rows][in] (bfloat16, cuda))."""
        rows = x.shape[0]
        _check(lib().xg_mlp_forward(self.h, rows, self._dp(x), self._dp(y)), "xg_mlp_forward")

    def loss_backward(self, x, t, mask, count, want_grads=False):
        rows = x.shape[0]
        loss = C.c_double()
        g = np.zeros(self.n_params, np.float32) if want_grads else None
        _check(lib().xg_mlp_loss_backward(self.h, rows, self._dp(x), self._dp(t), self._dp(mask), int(count),
                                          C.byref(loss), _p(g)), "xg_mlp_loss_backward")
        return (loss.value, g) if want_grads else loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(lib().xg_mlp_adam(self.h, C.c_float(lr), C.c_float(beta1), C.c_float(beta2), C.c_float(eps)),
               "xg_mlp_adam")

    def sync(self):
        _check(lib().xg_mlp_sync(self.h), "xg_mlp_sync")


class Transient:
    """One circuit's DC operating point and .TRAN on the GPU, every vector of
    the Newton and time-integration loop resident in device memory
    (``Transient::doLoopProcess`` with OneStep, DampedNewton and the DC
    strategy; src/AnalysisPKG/N_ANP_Transient.C:1184)."""

    STATS = ("steps", "rejects", "newton_iters", "loads", "t", "status", "finished", "dc_iters", "dc_stage",
             "seconds", "singular", "ir_corrections", "fp64_fallbacks", "points")

    def __init__(self, circuit, tstop, tstep, hmax=0.0, lu_method=0, gmin=1e-12, temp=300.15):
        """``lu_method``: 0 automatic, 1 level-scheduled, 2 multifrontal."""
        self._desc, self._keep = make_desc(circuit, gmin, temp)
        h = C.c_void_p()
        _check(lib().xg_tran_create(C.byref(self._desc), C.c_double(tstop), C.c_double(tstep), C.c_double(hmax),
                                    int(lu_method), C.byref(h)), "xg_tran_create")
        self.h = h
        self.n = host_pattern(circuit, gmin, temp)["n"]
        self._k = 0

    def close(self):
        if getattr(self, "h", None):
            lib().xg_tran_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision, ir_tol=1e-12, ir_max=10):
        _check(lib().xg_tran_set_precision(self.h, int(precision), C.c_double(ir_tol), int(ir_max)),
               "xg_tran_set_precision")

    def set_tolerances(self, reltol=1e-3, abstol=1e-6, nl_reltol=1e-2, nl_abstol=1e-6):
        _check(lib().xg_tran_set_tolerances(self.h, C.c_double(reltol), C.c_double(abstol), C.c_double(nl_reltol),
                                            C.c_double(nl_abstol)), "xg_tran_set_tolerances")

    def set_uic(self, on=True):
        """Skip the operating point (.TRAN ... UIC): the run starts from
        x = 0 with the devices' state of one load there."""
        _check(lib().xg_tran_set_uic(self.h, int(bool(on))), "xg_tran_set_uic")

    def set_progress(self, on=True):
        """A line on stderr per Newton solve of the DC operating point."""
        _check(lib().xg_tran_set_progress(self.h, int(bool(on))), "xg_tran_set_progress")

    def set_record(self, record_max, w=None, sel=None):
        """Record, at up to ``record_max`` points, the products of the rows
        of ``w`` (k x n) with the solution followed by the unknowns ``sel``,
        or the whole solution when both are None."""
        k = 0
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, np.float64).reshape(-1, self.n)
            k = w.shape[0]
            self._w = w
            wp = _p(w)
        ns = 0
        sp = None
        if sel is not None:
            sel = np.ascontiguousarray(sel, np.int32)
            ns = sel.size
            self._sel = sel
            sp = _p(sel)
        self._k = k + ns
        _check(lib().xg_tran_set_record(self.h, k, wp, ns, sp, int(record_max)), "xg_tran_set_record")

    def run(self, t_until=np.inf, max_steps=2**62):
        s = np.zeros(14)
        _check(lib().xg_tran_run(self.h, C.c_double(min(t_until, 1e300)), C.c_longlong(int(max_steps)), _p(s)),
               "xg_tran_run")
        out = dict(zip(self.STATS, (float(v) for v in s)))
        for k in ("steps", "rejects", "newton_iters", "loads", "status", "finished", "dc_iters", "dc_stage",
                  "singular", "ir_corrections", "fp64_fallbacks", "points"):
            out[k] = int(out[k])
        return out

    def points(self, max_points):
        width = self._k or self.n
        t = np.zeros(max(max_points, 1))
        o = np.zeros(max(max_points, 1), np.int32)
        y = np.zeros(max(max_points * width, 1))
        m = _check(lib().xg_tran_points(self.h, int(max_points), _p(t), _p(o), _p(y)), "xg_tran_points")
        return t[:m], o[:m], y[: m * width].reshape(m, width)

    def solution(self):
        x = np.zeros(self.n)
        _check(lib().xg_tran_solution(self.h, _p(x)), "xg_tran_solution")
        return x


def host_mf_plan(n, rowptr, colind, values=None):
    """Multifrontal plan statistics computed on the host (no GPU)."""
    info = np.zeros(6)
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    v = None if values is None else np.ascontiguousarray(values, np.float64)
    _check(lib().xg_host_mf_plan(int(n), _p(rp), _p(ci), _p(v), _p(info)), "xg_host_mf_plan")
    return dict(supernodes=int(info[0]), levels=int(info[1]), front_elements=int(info[2]), flops=float(info[3]),
                max_front=int(info[4]), nnz_lu=int(info[5]))
