"""Mixed precision on the 1M-MOSFET-class gate array: a Newton / transient run
with FP32 factors and FP64 iterative refinement against the all-FP64 run.

250k INV/NAND2 gates (749,642 MOSFETs, 375,323 unknowns), the device-resident
transient engine, .TRAN 0-0.15 ns UIC: the power-up of the array from 0 V
(with no operating point: the array's DC point is the end of a logic chain
some thousands of gates deep, which DC Newton and its gmin and source
stepping do not reach within the reference's iteration limits), OneStep
trapezoid steps, every Newton iteration factoring the Jacobian.  The
same run is made twice, with FP64 factors and with FP32 factors refined in
FP64 to a 1e-12 relative residual (device evaluation, residual assembly and
the Newton update stay FP64).  Reported per output node: the largest
difference over the run between the two, in units of the waveform tolerance
RELTOL |V| + ABSTOL (RELTOL 1e-3, ABSTOL 1e-6), on 4096 sampled gate outputs at
every accepted step and on every unknown at the end.  The FP64 engine itself
is checked against the CPU oracle on smaller arrays in
test_gpu_transient_engine.py.
"""
import time

import numpy as np
import pytest

import xyce_amd as X
from xyce_amd import circuits as C

pytestmark = pytest.mark.gpu


def _run(c, sel, precision):
    e = X.Transient(c, 0.15e-9, 1e-12, 1e-11)
    if precision:
        e.set_precision(1, 1e-12, 10)
    e.set_uic()
    e.set_record(4000, sel=sel)
    t0 = time.perf_counter()
    t_next = 0.0
    while True:
        # 25 ps at a time (a progress line each)
        t_next = min(0.15e-9, t_next + 25e-12)
        st = e.run(t_until=t_next)
        print(f"  {'FP32' if precision else 'FP64'} factors: t = {st['t'] * 1e12:.1f} ps, {st['steps']} steps, "
              f"{st['newton_iters']} Newton iterations, DC stage {st['dc_stage']}, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        if st["status"] != 0 or st["finished"]:
            break
    wall = time.perf_counter() - t0
    t, o, y = e.points(st["points"])
    return st, t, y, e.solution(), wall


def test_gate_array_fp32_factors_against_fp64_run(gpu):
    c = C.gate_dag(250_000, seed=4)
    p = X.host_pattern(c)
    outs = np.array([p["node_gid"][k] for k, name in c.names.items() if name.startswith("g")])
    rng = np.random.default_rng(0)
    sel = np.sort(rng.choice(outs, 4096, replace=False)).astype(np.int32)
    s64, t64, y64, x64, w64 = _run(c, sel, 0)
    s32, t32, y32, x32, w32 = _run(c, sel, 1)
    print(f"FP64 factors: {s64}, wall {w64:.1f} s")
    print(f"FP32 factors + FP64 refinement: {s32}, wall {w32:.1f} s")
    for s in (s64, s32):
        assert s["finished"] == 1 and s["status"] == 0, s
    assert s32["steps"] == s64["steps"] and s32["newton_iters"] == s64["newton_iters"]
    np.testing.assert_array_equal(t32, t64)
    tol = 1e-3 * np.abs(y64) + 1e-6
    per_node = np.max(np.abs(y32 - y64) / tol, axis=0)
    fin = np.abs(x32 - x64) / (1e-3 * np.abs(x64) + 1e-6)
    print(f"per sampled output node, max over {len(t64)} points of |V32 - V64| / (RELTOL|V| + ABSTOL): "
          f"max {per_node.max():.2e}, median {np.median(per_node):.2e}; every unknown at 0.15 ns: max {fin.max():.2e}; "
          f"{s32['ir_corrections']} refinement corrections, {s32['fp64_fallbacks']} FP64 fallbacks")
    assert per_node.max() <= 1e-3
    assert fin.max() <= 1e-3
    # the outputs did switch inside the window
    assert np.max(np.ptp(y64, axis=0)) > 1.0
