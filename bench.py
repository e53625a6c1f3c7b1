#!/usr/bin/env python3
"""Benchmark of the MI355X per-time-step engine.

Workload (default): the 3D plane-wave / 40-sphere model of
maxwellfdtd_amd.workloads.plane_wave_spheres_3d -- 256^3 cells, 10-cell CPML on every
face, TFSF plane wave, FP32 fields -- stepped W (warm-up) + K (timed) time steps.  A
"step" is one full time step of the model: H half, E half, TFSF, integral variables.

Prints one JSON line with whole-job throughput (Gcell-updates/s), the HBM roofline of
the dominant kernel (E half-step, hipEvent-timed on the engine's stream), the CPU
baseline (the oracle, a C restatement of the reference algorithm, on the host cores for
a bounded number of steps of the same model) and the relative L2 error of the GPU
fields against that FP64 run at the same step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Gcell-updates/s + % HBM peak; batched p50/p99 latency vs plain, same error"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-integrals", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=16)
    ap.add_argument("--zchunk", type=int, default=16)
    ap.add_argument("--ty", type=int, default=4)
    ap.add_argument("--kform", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU z-slab sharding is "
                                   "a design only (see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)

    from maxwellfdtd_amd import _lib
    from maxwellfdtd_amd.models import Handle
    from maxwellfdtd_amd.workloads import bytes_per_cell_update, plane_wave_spheres_3d

    lib = _lib.engine()
    import ctypes as C
    free, total, cus = C.c_int64(), C.c_int64(), C.c_int32()
    name = C.create_string_buffer(128)
    lib.check(lib.device_info(C.byref(free), C.byref(total), C.byref(cus), name, 128), "device_info")
    log(f"device {name.value.decode()} CUs={cus.value} HBM free={free.value/2**30:.1f} GiB")

    t0 = time.perf_counter()
    nsteps_model = args.warmup + args.steps + args.timing_steps
    model = plane_wave_spheres_3d(n=args.n, steps=nsteps_model)
    integrals = 0 if args.no_integrals else None
    desc = model.describe(integrals=integrals)
    N = int(np.prod(desc.dims))
    log(f"model built in {time.perf_counter()-t0:.1f}s: {desc.dims} cells, Nt={desc.Nt}, "
        f"materials={len(desc.materials)}, integrals={desc.integrals}")

    h = Handle(desc, lib)
    lib.check(lib.model_set_options(h.h, 0, args.graph_steps), "opt")
    lib.check(lib.model_set_options(h.h, 1, args.zchunk), "opt")
    lib.check(lib.model_set_options(h.h, 5, args.ty), "opt")
    lib.check(lib.model_set_options(h.h, 6, args.kform), "opt")
    h.step(args.warmup)
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    h.step(args.steps)
    wall = time.perf_counter() - t0
    gcells = N * args.steps / wall / 1e9
    log(f"timed {args.steps} steps: {wall*1e3:.1f} ms, {gcells:.2f} Gcell-updates/s")

    # dominant-kernel timing: hipEvent pairs around the H and E launches on the engine's
    # stream, eager launches (no graph) over timing_steps steps
    lib.check(lib.model_set_options(h.h, 3, 1), "opt")
    h.step(args.timing_steps)
    tot, cnt = C.c_double(), C.c_int32()
    lib.check(lib.model_timing(h.h, 0, C.byref(tot), C.byref(cnt)), "timing")
    h_ms = tot.value / max(cnt.value, 1)
    lib.check(lib.model_timing(h.h, 1, C.byref(tot), C.byref(cnt)), "timing")
    e_ms = tot.value / max(cnt.value, 1)
    bh, be = bytes_per_cell_update(model)
    log(f"H kernel {h_ms*1e3:.1f} us ({N*bh/h_ms/1e6:.0f} GB/s), E kernel {e_ms*1e3:.1f} us "
        f"({N*be/e_ms/1e6:.0f} GB/s) over {cnt.value} launches")
    h.close()

    e_bytes = N * be
    achieved = e_bytes / (e_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic_plane_wave_spheres_3d_256.json")
    if os.path.exists(tpath) and args.n == 256:
        with open(tpath) as f:
            traffic = json.load(f).get("e_kernel_bytes_per_launch")

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "tests"))
        from oracle_util import field_errors, oracle
        orc = oracle("f64")
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        threads = orc.set_threads(threads)
        ho = Handle(desc, orc)
        t0 = time.perf_counter()
        ns = 0
        while time.perf_counter() - t0 < args.cpu_seconds and ns < desc.Nt:
            ho.step(1)
            ns += 1
        cw = time.perf_counter() - t0
        cpu = {"value": round(N * ns / cw / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
               "kind": "port",
               "sample": f"first {ns} steps of the same {args.n}^3 model, FP64, "
                         f"{threads} OpenMP threads (oracle/fdtd_oracle.c)"}
        log(f"cpu baseline: {ns} steps in {cw:.1f}s -> {cpu['value']} Gcell/s")
        hg = Handle(desc, lib)
        hg.step(ns)
        errs = field_errors(hg, ho, 3)
        parity = {"steps": ns, "rel_l2_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()}}
        log(f"parity after {ns} steps: {errs}")
        hg.close()
        ho.close()

    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded geometry and sources)",
        "config": {"workload": f"plane_wave_spheres_3d: {args.n}^3 cells, 10-cell CPML, TFSF "
                               f"plane wave, 40 dielectric spheres, Sa/E2 integrals "
                               f"{'off' if args.no_integrals else 'on'}",
                   "cells": N, "model_steps": desc.Nt, "graph_steps": args.graph_steps,
                   "zchunk": args.zchunk, "rows_per_workgroup": args.ty, "kernel_form": args.kform},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_e3d_p (E half-step)", "bytes_per_cell": round(be, 3),
                     "avg_launch_us": round(e_ms * 1e3, 2),
                     "h_kernel": {"avg_launch_us": round(h_ms * 1e3, 2), "bytes_per_cell": round(bh, 3),
                                  "achieved": round(N * bh / h_ms / 1e6, 1)}},
        "step_bandwidth_GBs": round(N * (bh + be) * args.steps / wall / 1e9, 1),
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
