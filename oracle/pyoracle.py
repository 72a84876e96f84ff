"""ctypes wrapper of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``).  Only the tests,
the smoke check and the CPU-baseline leg of the benchmark use this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_create.restype = P
        L.orc_destroy.argtypes = [P]
        L.orc_info.argtypes = [P, P]
        L.orc_pattern.argtypes = [P] * 8
        L.orc_load.argtypes = [P] * 15
        L.orc_load.restype = C.c_int
        L.orc_combine.argtypes = [P] * 7
        L.orc_lu_solve.argtypes = [C.c_int, P, P, P, P, P]
        L.orc_lu_solve.restype = C.c_int
        L.orc_wave.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_wave.restype = C.c_double
        L.orc_dcop.argtypes = [P] * 7
        L.orc_dcop.restype = C.c_int
        L.orc_tran.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, P]
        L.orc_tran.restype = C.c_int
        L.orc_tran_points.argtypes = [P, P, P]
        L.orc_tran_orders.argtypes = [P, P]
        L.orc_set_probe.argtypes = [P, C.c_int, P]
        L.orc_set_probe.restype = C.c_int
        L.orc_set_dc_first_stage.argtypes = [P, C.c_int]
        L.orc_set_dc_first_stage.restype = C.c_int
        L.orc_dc_strategy.argtypes = [P]
        L.orc_dc_strategy.restype = C.c_int
        L.orc_set_method.argtypes = [P, C.c_int]
        L.orc_set_method.restype = C.c_int
        L.orc_set_uic.argtypes = [P, C.c_int]
        L.orc_set_uic.restype = C.c_int
        for nm in ("orc_limvds", "orc_fetlim", "orc_qmeyer", "orc_pnjlim"):
            getattr(L, nm).restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class Oracle:
    """One circuit in the CPU oracle."""

    def __init__(self, circuit, gmin=1e-12, temp=300.15):
        a = circuit.arrays()
        self._keep = a
        L = lib()
        h = L.orc_create(
            C.c_int(a["n_nodes"]), C.c_int(len(a["dev_type"])), _p(a["dev_type"]), _p(a["dev_idx"]),
            C.c_int(a["n_models"]), _p(a["model_par"]),
            C.c_int(a["n_mos"]), _p(a["mos_nodes"]), _p(a["mos_model"]), _p(a["mos_inst"]),
            C.c_int(a["n_res"]), _p(a["res_nodes"]), _p(a["res_val"]),
            C.c_int(a["n_cap"]), _p(a["cap_nodes"]), _p(a["cap_val"]),
            C.c_int(a["n_vsrc"]), _p(a["vsrc_nodes"]), _p(a["vsrc_wave"]),
            C.c_int(a["n_isrc"]), _p(a["isrc_nodes"]), _p(a["isrc_wave"]),
            C.c_int(a["n_wave"]), _p(a["wave_type"]), _p(a["wave_off"]), _p(a["wave_data"]),
            C.c_double(gmin), C.c_double(temp))
        if not h:
            raise RuntimeError("orc_create failed")
        self.h = C.c_void_p(h)
        info = np.zeros(5, np.int32)
        L.orc_info(self.h, _p(info))
        self.n, self.nnz, self.nsta, self.nsto, self.n_mos = (int(v) for v in info)
        self.n_nodes = a["n_nodes"]

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
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
        lib().orc_pattern(self.h, _p(rp), _p(ci), _p(ng), _p(ml), _p(fs), _p(qs), _p(ss))
        return dict(rowptr=rp, colind=ci[: self.nnz], node_gid=ng,
                    mos_lid=ml[: 6 * self.n_mos].reshape(-1, 6),
                    mos_fslot=fs[: 17 * self.n_mos].reshape(-1, 17),
                    mos_qslot=qs[: 14 * self.n_mos].reshape(-1, 14),
                    mos_stasto=ss[: 2 * self.n_mos].reshape(-1, 2))

    def load(self, x, sta_cur, sto_cur, sta_next=None, sto_next=None, *, dcop=0, tran=1, init_jct=0,
             newton_iter=1, order=1, h=1e-11, time=0.0, tstep=1e-12, tstop=1e-9, tranop=0):
        ssi = np.array([dcop, tran, init_jct, 0, newton_iter, order, tranop], np.int32)
        ssd = np.array([h, time, tstep, tstop], np.float64)
        x = np.ascontiguousarray(x, np.float64)
        sta_next = np.array(sta_cur if sta_next is None else sta_next, np.float64)
        sto_next = np.array(sto_cur if sto_next is None else sto_next, np.float64)
        out = {k: np.zeros(self.n) for k in ("F", "Q", "B", "dFdxdVp", "dQdxdVp")}
        dF = np.zeros(self.nnz)
        dQ = np.zeros(self.nnz)
        sc = np.ascontiguousarray(sta_cur, np.float64)
        oc = np.ascontiguousarray(sto_cur, np.float64)
        lim = lib().orc_load(self.h, _p(ssi), _p(ssd), _p(x), _p(sc), _p(oc), _p(sta_next), _p(sto_next),
                             _p(out["F"]), _p(out["Q"]), _p(out["B"]), _p(out["dFdxdVp"]), _p(out["dQdxdVp"]),
                             _p(dF), _p(dQ))
        out.update(dFdx=dF, dQdx=dQ, sta_next=sta_next, sto_next=sto_next, limited=lim, _ssi=ssi, _ssd=ssd)
        return out

    def combine(self, loaded, q0, q2):
        rhs = np.zeros(self.n)
        J = np.zeros(self.nnz)
        q0 = np.ascontiguousarray(q0, np.float64)
        q2 = np.ascontiguousarray(q2, np.float64)
        lib().orc_combine(self.h, _p(loaded["_ssi"]), _p(loaded["_ssd"]), _p(q0), _p(q2), _p(rhs), _p(J))
        return rhs, J

    def wave(self, w, t, tstep=1e-12, tstop=1e-9, bptol=1e-18):
        return lib().orc_wave(self.h, w, t, tstep, tstop, bptol)

    def dcop(self):
        x = np.zeros(self.n)
        sta = np.zeros(max(self.nsta, 1))
        sto = np.zeros(max(self.nsto, 1))
        Q, F, B = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        st = lib().orc_dcop(self.h, _p(x), _p(sta), _p(sto), _p(Q), _p(F), _p(B))
        return st, x, sta[: self.nsta], sto[: self.nsto], Q, F, B

    LOAD_HOOK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int), *([C.POINTER(C.c_double)] * 13))
    SOLVE_HOOK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))

    def set_hooks(self, load=None, solve=None):
        """Let this oracle's Newton / time-integration loop (the restatement of
        the reference's driver) call ``load`` and ``solve`` (Python callables
        with orc_load / orc_lu_solve arguments) instead of its own loads and LU."""
        self._hooks = (self.LOAD_HOOK(load) if load else self.LOAD_HOOK(), self.SOLVE_HOOK(solve) if solve else self.SOLVE_HOOK())
        lib().orc_set_hooks(self.h, self._hooks[0], self._hooks[1], None)

    def set_tolerances(self, tia_reltol=1e-3, tia_abstol=1e-6, nl_reltol=1e-2, nl_abstol=1e-6):
        lib().orc_set_tolerances(self.h, C.c_double(tia_reltol), C.c_double(tia_abstol), C.c_double(nl_reltol),
                                 C.c_double(nl_abstol))

    def dc_strategy(self):
        """Strategy of the last DC operating point: 'newton', 'gmin' or
        'source' (stepping)."""
        k = lib().orc_dc_strategy(self.h)
        return ("newton", "gmin", "source")[k] if k >= 0 else "none"

    def set_uic(self, on=True):
        """Later transients skip the operating point (.TRAN ... UIC): they
        start from x = 0 with the devices' state of one load there."""
        assert lib().orc_set_uic(self.h, int(bool(on))) == 0

    def set_dc_first_stage(self, stage):
        """Start the DC strategy at 'newton' (the reference's order), 'gmin'
        or 'source' stepping."""
        assert lib().orc_set_dc_first_stage(self.h, {"newton": 0, "gmin": 1, "source": 2}[stage]) == 0

    def set_method(self, method):
        """Integration method of later transients: 'trap' (OneStep, the
        default) or 'gear' (Gear12)."""
        m = {"trap": 0, "gear": 1}[method]
        assert lib().orc_set_method(self.h, m) == 0

    def set_probe(self, w=None):
        """Record, per accepted point, the products of the rows of ``w``
        (k x n) with the solution instead of the whole solution; None
        restores whole solutions."""
        if w is None:
            self._probe_k = 0
            lib().orc_set_probe(self.h, 0, None)
            return
        w = np.ascontiguousarray(w, np.float64).reshape(-1, self.n)
        self._probe_k = w.shape[0]
        lib().orc_set_probe(self.h, w.shape[0], _p(w))

    def tran(self, tstop, tstep, hmax, record=True, t_record_stop=None):
        stats = np.zeros(7)
        st = lib().orc_tran(self.h, tstop, tstep, hmax, 1 if record else 0,
                            tstop if t_record_stop is None else t_record_stop, _p(stats))
        npts = int(stats[6])
        width = getattr(self, "_probe_k", 0) or self.n
        t = np.zeros(max(npts, 1))
        x = np.zeros(max(npts * width, 1))
        order = np.zeros(max(npts, 1), np.int32)
        lib().orc_tran_points(self.h, _p(t), _p(x))
        lib().orc_tran_orders(self.h, _p(order))
        return dict(status=st, steps=int(stats[0]), rejects=int(stats[1]), newton_iters=int(stats[2]),
                    loads=int(stats[3]), seconds=float(stats[4]), t=t[:npts], x=x[: npts * width].reshape(npts, width),
                    order=order[:npts])


def lu_solve(n, rowptr, colind, vals, b):
    x = np.zeros(n)
    st = lib().orc_lu_solve(n, _p(np.ascontiguousarray(rowptr, np.int32)), _p(np.ascontiguousarray(colind, np.int32)),
                            _p(np.ascontiguousarray(vals, np.float64)), _p(np.ascontiguousarray(b, np.float64)), _p(x))
    return st, x


def limvds(vnew, vold):
    vnew = np.ascontiguousarray(vnew, np.float64); vold = np.ascontiguousarray(vold, np.float64)
    out = np.zeros_like(vnew)
    lib().orc_limvds(len(vnew), _p(vnew), _p(vold), _p(out))
    return out


def fetlim(vnew, vold, vto):
    vnew, vold, vto = (np.ascontiguousarray(a, np.float64) for a in (vnew, vold, vto))
    out = np.zeros_like(vnew)
    lib().orc_fetlim(len(vnew), _p(vnew), _p(vold), _p(vto), _p(out))
    return out


def pnjlim(vnew, vold, vt, vcrit):
    vnew, vold, vt, vcrit = (np.ascontiguousarray(a, np.float64) for a in (vnew, vold, vt, vcrit))
    out = np.zeros_like(vnew)
    chk = np.zeros(len(vnew), np.int32)
    lib().orc_pnjlim(len(vnew), _p(vnew), _p(vold), _p(vt), _p(vcrit), _p(out), _p(chk))
    return out, chk


def qmeyer(args):
    """args: (N, 7) of vgs vgd vgb von vdsat phi cox -> (N, 3) capgs capgd capgb."""
    a = np.ascontiguousarray(args, np.float64)
    out = np.zeros((a.shape[0], 3))
    lib().orc_qmeyer(a.shape[0], _p(a), _p(out))
    return out
