"""Benchmark: Newton steps/s of the batched Monte Carlo ring-oscillator engine.

Workload: 4096 CMOS ring oscillators, stage counts odd and uniform on
[11, 201], per-device Gaussian VTO (sigma 30 mV) and KP (sigma 5 %) from a
per-instance seed, .TRAN 0-20 ns with a 10 ps maximum step, FP64 throughout.
A "step" is one launch of the batch: K batched Newton iterations back to back
(K chosen by the autotuner), in each of which every active instance evaluates
its MOSFETs, stamps, refactors and solves (or, when its previous update
converged, runs its step control).  ``value`` counts the Newton iterations
(eval + stamp + LU + solve) actually completed per second over the timed
launches, all instances together, inputs resident in HBM.  The default
warm-up and timed launches cover the same stretch of the run (Newton
iterations 4800 to 11200 of every instance, all of them still integrating)
whatever K is.

Also reported: the same engine on one instance of the 101-stage ring
(single-instance FP64 GPU path) and the ratio, the roofline of the batched
kernel, and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(circuits_list):
    """Bytes one batched Newton iteration must move (all instances active):
    per MOSFET 3 FP64 parameters + class id + stamp mask (36 B), the 6 store
    values at t_n and at the last iterate and the 8 state values at t_n read
    (160 B), 14 store/state values written (112 B); per unknown x read and
    written, F/Q/B written, two charge histories read (56 B); per Jacobian
    entry the two linear-part values read (16 B)."""
    import xyce_amd as X
    total = 0
    cache = {}
    for c in circuits_list:
        key = (len(c.mos_nodes), c.n_nodes)
        if key not in cache:
            hp = X.host_pattern(c)
            cache[key] = (len(c.mos_nodes), hp["n"], hp["nnz"])
        nm, n, nnz = cache[key]
        total += nm * (36 + 160 + 112) + n * 56 + nnz * 16
    return total


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(circuits_list, sample, tstop, tstep, hmax, threads):
    import ctypes as C
    from oracle.pyoracle import Oracle, lib
    orcs = [Oracle(c) for c in circuits_list[:sample]]
    L = lib()
    L.orc_tran_many.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]
    arr = (C.c_void_p * len(orcs))(*[o.h.value for o in orcs])
    out = np.zeros(4)
    L.orc_tran_many(len(orcs), arr, tstop, tstep, hmax, threads, out.ctypes.data_as(C.c_void_p))
    return dict(newton_iters=out[0], seconds=out[1], steps=out[2], failures=out[3])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the same stretch of the run (Newton iterations 4800-11200 of
    # every instance, where every instance is still integrating) whatever
    # iterations per launch the autotuner picks
    ap.add_argument("--steps", type=int, default=None, help="timed launches (default 6400 // iterations per launch)")
    ap.add_argument("--warmup", type=int, default=None, help="warm-up launches (default 4800 // iterations per launch)")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--cpu-tstop", type=float, default=20e-9)
    ap.add_argument("--form", choices=("auto", "split", "fused"), default="auto",
                    help="launch form: autotuned, separate evaluation kernel, or fused kernel")
    ap.add_argument("--iters", type=int, default=8, help="Newton iterations per launch of the fused form")
    ap.add_argument("--block", type=int, default=0,
                    help="workgroup size; with --form split|fused skips the autotuner (instances by LDS need)")
    ap.add_argument("--profile", action="store_true",
                    help="batch leg only (no single-instance run, no CPU baseline): for rocprofv3 runs")
    ap.add_argument("--phases", action="store_true",
                    help="batch leg only, with the engine's per-phase device timers on over the timed launches "
                         "(one instance per launch group): prints microseconds per Newton iteration per phase")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (the reference kernel for rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the batched engine shards "
              "by instance with no exchange, so N GPUs give N independent runs", file=sys.stderr)
        sys.exit(2)

    import torch
    import xyce_amd as X
    from xyce_amd import circuits as C

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 10
        print(json.dumps({"copy_bytes": nbytes, "read_plus_write_GBps": round(2 * nbytes / dt / 1e9, 1)}))
        return
    tstop, tstep, hmax = 20e-9, 1e-12, 10e-12

    t0 = time.time()
    ckts = C.mc_ring_batch(args.instances, seed=3)
    log(f"generated {len(ckts)} circuits in {time.time() - t0:.1f}s")
    t0 = time.time()
    b = X.Batch(ckts, tstop, tstep, hmax)
    log(f"batch setup {time.time() - t0:.1f}s")
    if args.form != "auto" and args.block:
        b.config(block=args.block, iters_per_launch=args.iters, split_eval=1 if args.form == "split" else 0)
        tune = dict(block=args.block, order_mode=1, split_eval=int(args.form == "split"),
                    iters_per_launch=1 if args.form == "split" else args.iters, autotuned=False)
    else:
        b.config(iters_per_launch=args.iters, split_eval=0)
        tune = b.autotune()
        if args.form != "auto":
            b.config(iters_per_launch=args.iters, split_eval=1 if args.form == "split" else 0)
    log(f"launch form: {tune} (form {args.form})")
    log(f"launch groups: {b.groups()}")
    if args.form == "auto":
        k_iter = tune["iters_per_launch"]
    else:
        k_iter = 1 if args.form == "split" else args.iters
    steps = args.steps if args.steps is not None else max(1, 6400 // k_iter)
    warmup = args.warmup if args.warmup is not None else max(1, 4800 // k_iter)
    b.advance(warmup, graph_chunk=50)
    if args.phases:
        b.profile(True)
    c0 = b.counters()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    chunk = 50 if steps % 50 == 0 else 0
    b.advance(steps, graph_chunk=chunk)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    c1 = b.counters()
    dev_s = c1["seconds"]
    iters = c1["newton_iters"] - c0["newton_iters"]
    loads = c1["loads"] - c0["loads"]
    value = iters / wall
    log(f"timed {steps} launches: wall {wall*1e3:.2f} ms, events {dev_s*1e3:.2f} ms, "
        f"{iters:.0f} Newton iterations, {loads:.0f} loads, done {c1['done']}, failed {c1['failed']}, "
        f"dense {c1['dense_solves'] - c0['dense_solves']:.0f}")

    if args.phases:
        ph = b.profile(False)
        it = max(ph.pop("iterations"), 1)
        per = {k: round(v / it, 3) for k, v in ph.items()}
        print(json.dumps({"phases_us_per_iteration": per, "iterations_timed": it,
                          "sum_us": round(sum(per.values()), 3), "launches": steps, "tune": tune}), flush=True)
        return
    if args.profile:
        print(json.dumps({"profile_run": True, "launches": steps, "newton_iterations": int(iters),
                          "events_ms": round(dev_s * 1e3, 3), "tune": tune}), flush=True)
        return

    # single-instance FP64 GPU path: the same engine on the 101-stage ring
    r1 = C.ring(101, seed=1)
    b1 = X.Batch([r1], 50e-9, tstep, hmax)
    b1.config(iters_per_launch=args.iters, split_eval=0)
    tune1 = b1.autotune()
    log(f"single-instance autotune: {tune1}")
    k1 = tune1["iters_per_launch"]
    steps1 = args.steps if args.steps is not None else max(1, 6400 // k1)
    warmup1 = args.warmup if args.warmup is not None else max(1, 4800 // k1)
    chunk1 = 50 if steps1 % 50 == 0 else 0
    b1.advance(warmup1, graph_chunk=50)
    s0 = b1.counters()
    torch.cuda.synchronize()
    w1 = time.perf_counter()
    b1.advance(steps1, graph_chunk=chunk1)
    torch.cuda.synchronize()
    wall1 = time.perf_counter() - w1
    s1 = b1.counters()
    single = (s1["newton_iters"] - s0["newton_iters"]) / wall1
    log(f"single instance: {single:.1f} Newton steps/s ({wall1 / steps1 * 1e6:.1f} us/launch)")

    byts = algorithmic_bytes(ckts) * k_iter
    us_launch = dev_s / steps * 1e6
    achieved = byts / (dev_s / steps) / 1e9
    traffic = None
    tfile = os.path.join(HERE, "profiles", "traffic_k_batch_step.json")
    if os.path.exists(tfile):
        # HBM bytes per batched iteration from the rocprofv3 FETCH_SIZE /
        # WRITE_SIZE passes (scaled by the copy-kernel calibration), recorded
        # with the command that made them; times this launch's iterations
        with open(tfile) as f:
            per_it = json.load(f).get("hbm_bytes_per_iteration")
        traffic = int(per_it * k_iter) if per_it else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "stored rocprofv3 FETCH_SIZE + WRITE_SIZE per batched iteration, copy-kernel "
                                  "scaled (profiles/k_batch_step_mc4096_traffic.md), times this launch's iterations"
                if traffic else None,
                "kernel": "k_batch_eval + k_batch_step" if k_iter == 1 and tune["split_eval"] else "k_batch_step",
                "bytes_per_launch": int(byts), "us_per_launch": round(us_launch, 2),
                "iterations_per_launch": k_iter}

    cpu = None
    single_cpu = None
    if not args.no_cpu:
        # one GPU's share of the host: 16 threads (the benchmark host runs one
        # GPU's work per 16 cores, whatever the machine's thread count)
        host_threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = min(16, host_threads)
        r = cpu_baseline(ckts, args.cpu_sample, args.cpu_tstop, tstep, hmax, threads)
        cpu = {"value": round(r["newton_iters"] / r["seconds"], 1), "unit": "Newton steps/s", "cores": threads,
               "kind": "port", "cpu_model": cpu_model(), "host_threads": host_threads,
               "sample": f"first {args.cpu_sample} instances of the batch, .TRAN 0-{args.cpu_tstop * 1e9:g} ns each, "
                         f"one instance per thread at a time ({r['newton_iters']:.0f} Newton iterations in "
                         f"{r['seconds']:.1f} s)"}
        log(f"cpu baseline: {cpu}")
        # the reference's own serial case: the 101-stage ring, 0-50 ns, one core
        from oracle.pyoracle import Oracle
        o1 = Oracle(r1)
        t1 = time.perf_counter()
        ro = o1.tran(50e-9, tstep, hmax, record=False)
        dt1 = time.perf_counter() - t1
        single_cpu = {"value": round(ro["newton_iters"] / dt1, 1), "unit": "Newton steps/s", "cores": 1,
                      "kind": "port", "sample": f"101-stage ring, .TRAN 0-50 ns, hmax 10 ps: {ro['newton_iters']} "
                                                f"Newton iterations in {dt1:.2f} s"}
        log(f"single-instance cpu: {single_cpu}")

    out = {
        "metric": "Newton steps/sec (eval+stamp+LU) per batch; x vs plain FP64 GPU at RELTOL",
        "value": round(value, 1),
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "x_vs_single_instance_gpu": round(value / single, 2) if single > 0 else None,
        "single_instance_gpu": round(single, 1),
        "single_instance_cpu": single_cpu,
        "dtype": "f64",
        "data": "synthetic (seeded Monte Carlo ring oscillators)",
        "config": {"workload": f"{args.instances} Monte Carlo CMOS ring oscillators, stages odd U[11,201], "
                               "level-1 MOSFETs, VTO sigma 30 mV, KP sigma 5%, .TRAN 0-20ns hmax 10ps, FP64",
                   "instances": args.instances, "newton_iterations_timed": int(iters),
                   "autotune": tune, "parallelism": "one workgroup per instance"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
