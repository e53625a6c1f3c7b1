#!/usr/bin/env python3
"""Benchmark of the MI355X per-time-step engine.

Default workload: the 3D plane-wave / 40-sphere model of
maxwellfdtd_amd.workloads.plane_wave_spheres_3d -- 256^3 cells, 10-cell CPML on every
face, TFSF plane wave, FP32 fields -- stepped W (warm-up) + K (timed) time steps.  A
"step" is one full time step of the model: H half, E half, TFSF, integral variables.

Prints one JSON line with whole-job throughput (Gcell-updates/s), the HBM roofline of
the dominant kernel (E half-step, hipEvent-timed on the engine's stream), the CPU
baseline (the oracle, a C restatement of the reference algorithm, on the host cores for
a bounded number of steps of the same model) and the relative L2 error of the GPU
fields against that FP64 run at the same step.

Other workloads (--workload): dispersive_3d, batched_2d (batch vs one solve at a time),
served_3d (continuous-batching server vs one request at a time), slab_1d.  Each prints
its own JSON line.
"""
import argparse
import ctypes as C
import itertools
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


def set_opts(lib, h, opts):
    for k, v in opts.items():
        lib.check(lib.model_set_options(h.h, k, v), "set_options")


def kernel_timing(lib, h, steps):
    """Average hipEvent time (ms) of the H and E launches over `steps` eager steps."""
    lib.check(lib.model_set_options(h.h, 3, 1), "opt")
    lib.check(lib.model_timing(h.h, 2, C.byref(C.c_double()), C.byref(C.c_int32())), "timing")
    h.step(steps)
    tot, cnt = C.c_double(), C.c_int32()
    lib.check(lib.model_timing(h.h, 0, C.byref(tot), C.byref(cnt)), "timing")
    h_ms = tot.value / max(cnt.value, 1)
    lib.check(lib.model_timing(h.h, 1, C.byref(tot), C.byref(cnt)), "timing")
    e_ms = tot.value / max(cnt.value, 1)
    lib.check(lib.model_set_options(h.h, 3, 0), "opt")
    return h_ms, e_ms, cnt.value


def host_cpu_info():
    """What the CPU baseline ran on: logical CPUs of the host, CPUs this process may run
    on, the cgroup CPU quota (if any), OMP_NUM_THREADS and the CPU model."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            info["cgroup_cpus"] = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    info["model"] = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def baseline_threads():
    """All CPUs this process may use (sched_getaffinity), bounded by OMP_NUM_THREADS when
    it is set (the CPU share a shared host grants) and by the cgroup quota."""
    n = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if omp > 0:
        n = min(n, omp)
    q = host_cpu_info().get("cgroup_cpus")
    if q:
        n = min(n, max(1, int(q)))
    return max(1, n)


def oracle_rate(orc, desc, seconds, N, label=""):
    """Steps the FP64 oracle for about `seconds`; returns (steps, wall, handle)."""
    from maxwellfdtd_amd.models import Handle
    ho = Handle(desc, orc)
    t0 = time.perf_counter()
    t_note = t0 + 60.0
    ns = 0
    while time.perf_counter() - t0 < seconds and ns < desc.Nt:
        ho.step(1)
        ns += 1
        if time.perf_counter() > t_note:    # long samples report as they go
            log(f"cpu baseline{label}: {ns} of {desc.Nt} steps after {time.perf_counter() - t0:.0f} s")
            t_note += 60.0
    return ns, time.perf_counter() - t0, ho


def cpu_leg(desc, seconds, N, label, gpu_lib, seconds_1t=5.0):
    """The oracle (FP64) on all usable host cores for a bounded number of steps, then on
    one thread for a shorter sample, then the GPU engine for the first sample's steps:
    CPU rates and GPU-vs-FP64 error (fields, integrals, total field energy, DFT spectra)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle_util import field_energy, field_errors, oracle, rel_l2, spectra_errors
    from maxwellfdtd_amd import _lib
    from maxwellfdtd_amd.models import Handle
    orc = oracle("f64")
    info = host_cpu_info()
    threads = orc.set_threads(baseline_threads())
    ns, cw, ho = oracle_rate(orc, desc, seconds, N)
    one = None
    if seconds_1t > 0:
        orc.set_threads(1)
        n1, w1, h1 = oracle_rate(orc, desc, seconds_1t, N, " (1 thread)")
        h1.close()
        one = {"value": round(N * n1 / w1 / 1e9, 5), "steps": n1}
        orc.set_threads(threads)
    cpu = {"value": round(N * ns / cw / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
           "kind": "port",
           "sample": f"first {ns} steps of the same {label} model, FP64, {threads} OpenMP "
                     f"threads (oracle/fdtd_oracle.c, the reference CPU algorithm restated in C)",
           "one_thread": one, "host": info}
    log(f"cpu baseline: {ns} steps in {cw:.1f}s on {threads} threads -> {cpu['value']} Gcell/s; "
        f"1 thread: {one}; host {info}")
    hg = Handle(desc, gpu_lib)
    hg.step(ns)
    errs = field_errors(hg, ho, desc.ndims)
    if desc.integrals & 1:
        errs["Sa"] = rel_l2(hg.output(_lib.MFD_OUT_SA), ho.output(_lib.MFD_OUT_SA))
    if desc.integrals & 2:
        errs["E2"] = rel_l2(hg.output(_lib.MFD_OUT_E2), ho.output(_lib.MFD_OUT_E2))
    ge, gh = hg.energy()
    oe, oh = field_energy(ho)
    parity = {"steps": ns, "rel_l2_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
              "energy_rel_err": float(f"{abs((ge + gh) / (oe + oh) - 1):.3e}") if oe + oh > 0 else None,
              **spectra_errors(hg, ho, desc)}
    log(f"parity after {ns} steps: {parity}")
    hg.close()
    ho.close()
    return cpu, parity


def graph_steps_for(k, default):
    """Steps per graph replay for K timed steps: K itself up to 64, else the largest
    divisor of K from 32 down to 8, else `default` (eager steps cover the remainder)."""
    if default <= 0:
        return default
    if k <= 64:
        return k
    for g in range(32, 7, -1):
        if k % g == 0:
            return g
    return default


def base_record(args, value, wall, steps, dtype="f32"):
    return {"metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": round(wall / max(steps, 1) * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded geometry and sources)"}


# ---------------------------------------------------------------------------------------
def bench_3d(args, lib):
    from maxwellfdtd_amd.models import Handle
    from maxwellfdtd_amd.workloads import bytes_per_cell_update, plane_wave_spheres_3d

    t0 = time.perf_counter()
    nsteps_model = args.warmup + args.steps + args.timing_steps
    model = plane_wave_spheres_3d(n=args.n, steps=nsteps_model)
    integrals = 0 if args.no_integrals else None
    desc = model.describe(integrals=integrals)
    N = int(np.prod(desc.dims))
    log(f"model built in {time.perf_counter()-t0:.1f}s: {desc.dims} cells, Nt={desc.Nt}, "
        f"materials={len(desc.materials)}, integrals={desc.integrals}")

    h = Handle(desc, lib)
    gsteps = graph_steps_for(args.steps, args.graph_steps)
    set_opts(lib, h, {0: gsteps, 1: args.zchunk, 5: args.ty, 6: args.kform,
                      10: args.snake, 11: args.dless, 12: args.pml_split, 14: args.zchunk_h,
                      15: args.xcd_h, 16: args.xcd_block, 18: args.tfsf_fold})
    lib.check(lib.model_upload(h.h), "upload")
    # the first step call builds, instantiates and uploads the step graph (whatever the
    # warm-up length), so the timed steps are whole graph replays
    h.step(args.warmup)
    log(f"warmup {args.warmup} steps done (graph of {gsteps} steps built)")
    t0 = time.perf_counter()
    h.step(args.steps)
    wall = time.perf_counter() - t0
    gcells = N * args.steps / wall / 1e9
    log(f"timed {args.steps} steps: {wall*1e3:.1f} ms, {gcells:.2f} Gcell-updates/s")

    # dominant-kernel timing: hipEvent pairs around the H and E launches on the engine's
    # stream, eager launches (no graph)
    h_ms, e_ms, cnt = kernel_timing(lib, h, args.timing_steps)
    bh, be = bytes_per_cell_update(model)
    fused = args.kform == 3
    if fused:
        # one launch does both half-steps (the second timing window holds only the state
        # pass of ADE/plasma cells, empty here)
        log(f"fused step kernel {h_ms*1e3:.1f} us ({N*(bh+be)/h_ms/1e6:.0f} GB/s at {bh+be:.3f} B/cell) "
            f"over {cnt} launches")
    else:
        log(f"H kernel {h_ms*1e3:.1f} us ({N*bh/h_ms/1e6:.0f} GB/s), E kernel {e_ms*1e3:.1f} us "
            f"({N*be/e_ms/1e6:.0f} GB/s) over {cnt} launches")
    h.close()

    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic_plane_wave_spheres_3d_256.json")
    if os.path.exists(tpath) and args.n == 256:
        with open(tpath) as f:
            traffic = json.load(f).get("fused_kernel_bytes_per_launch" if fused else "e_kernel_bytes_per_launch")

    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = cpu_leg(desc, args.cpu_seconds, N, f"{args.n}^3", lib)

    out = base_record(args, gcells, wall, args.steps)
    out["config"] = {"workload": f"plane_wave_spheres_3d: {args.n}^3 cells, 10-cell CPML, TFSF "
                                 f"plane wave, 40 dielectric spheres, Sa/E2 integrals "
                                 f"{'off' if args.no_integrals else 'on'}",
                     "cells": N, "model_steps": desc.Nt, "graph_steps": gsteps,
                     "zchunk": args.zchunk, "rows_per_workgroup": args.ty, "kernel_form": args.kform,
                     "snake_order": args.snake, "plain_cells_without_d": args.dless,
                     "cpml_split": args.pml_split, "zchunk_h": args.zchunk_h, "xcd_chunked_h": args.xcd_h, "xcd_block": args.xcd_block,
                     "tfsf_fold": args.tfsf_fold}
    if fused:
        achieved = N * (bh + be) / (h_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": traffic,
                           "kernel": "k_step3d_f (fused H+E step)", "bytes_per_cell": round(bh + be, 3),
                           "avg_launch_us": round(h_ms * 1e3, 2)}
    else:
        achieved = N * be / (e_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": traffic,
                           "kernel": "k_e3d_p (E half-step)", "bytes_per_cell": round(be, 3),
                           "avg_launch_us": round(e_ms * 1e3, 2),
                           "h_kernel": {"avg_launch_us": round(h_ms * 1e3, 2), "bytes_per_cell": round(bh, 3),
                                        "achieved": round(N * bh / h_ms / 1e6, 1)}}
    out["step_bandwidth_GBs"] = round(N * (bh + be) * args.steps / wall / 1e9, 1)
    out["cpu_baseline"] = cpu
    out["parity"] = parity
    return out


# ---------------------------------------------------------------------------------------
def bench_dispersive_3d(args, lib):
    from maxwellfdtd_amd.models import Handle
    from maxwellfdtd_amd.workloads import cpml_slab_cells, dispersive_3d

    steps = args.steps if args.steps_given else 5000
    t0 = time.perf_counter()
    model, mons = dispersive_3d(steps=steps + args.timing_steps)
    desc = model.describe(monitors=mons)
    N = int(np.prod(desc.dims))
    geo = model.geometry
    f_sub, f_metal = float(np.mean(geo == 1)), float(np.mean(geo == 2))
    s = cpml_slab_cells(model) / N
    # algorithmic bytes per cell-update: base 86 + CPML, + per dispersive cell 12 (E^n)
    # + 16 per Drude/Lorentz pole per component (3 poles x 3) or 8 per Debye component
    b_cell = 86 + 32 * s + f_metal * (12 + 3 * 3 * 16) + f_sub * (12 + 3 * 8)
    log(f"model built in {time.perf_counter()-t0:.1f}s: {desc.dims}, f_metal={f_metal:.4f}, "
        f"f_sub={f_sub:.4f}, B_cell={b_cell:.2f}")
    h = Handle(desc, lib)
    set_opts(lib, h, {0: args.graph_steps, 1: args.zchunk, 5: args.ty, 6: args.kform, 14: args.zchunk_h,
                      15: args.xcd_h, 16: args.xcd_block})
    t0 = time.perf_counter()
    h.step(steps)
    wall = time.perf_counter() - t0
    gcells = N * steps / wall / 1e9
    log(f"{steps} steps: {wall:.2f} s, {gcells:.2f} Gcell/s")
    h_ms, e_ms, cnt = kernel_timing(lib, h, args.timing_steps)
    bh = 37 + 16 * s
    be = b_cell - bh
    h.close()
    fused = args.kform == 3
    if fused:
        # the fused launch does H and E of every cell; the state pass finishes E of the
        # ADE cells (its poles are part of b_cell): the two windows together
        achieved = N * b_cell / ((h_ms + e_ms) * 1e-3) / 1e9
        log(f"fused step kernel {h_ms*1e3:.1f} us + state pass {e_ms*1e3:.1f} us "
            f"({achieved:.0f} GB/s at {b_cell:.3f} B/cell)")
    else:
        achieved = N * be / (e_ms * 1e-3) / 1e9
    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = cpu_leg(desc, args.cpu_seconds, N, "384x384x192 dispersive", lib)
    out = base_record(args, gcells, wall, steps)
    out["config"] = {"workload": "dispersive_3d: 384x384x192 cells, dx=5 nm, 10-cell CPML, "
                                 "broadband TFSF, 6x6 Drude+2 Lorentz spheres on a Debye "
                                 "substrate, 384x384 DFT plane at 16 frequencies",
                     "cells": N, "model_steps": desc.Nt, "f_metal": round(f_metal, 5),
                     "f_sub": round(f_sub, 5), "bytes_per_cell_update": round(b_cell, 3)}
    if fused:
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": "k_step3d_f (fused H+E step, MODE_FULL) + k_e_state",
                           "bytes_per_cell": round(b_cell, 3), "avg_launch_us": round((h_ms + e_ms) * 1e3, 2),
                           "fused_us": round(h_ms * 1e3, 2), "state_pass_us": round(e_ms * 1e3, 2)}
    else:
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": "k_e3d_p (E half-step, MODE_FULL)", "bytes_per_cell": round(be, 3),
                           "avg_launch_us": round(e_ms * 1e3, 2),
                           "h_kernel": {"avg_launch_us": round(h_ms * 1e3, 2),
                                        "bytes_per_cell": round(bh, 3)}}
    out["step_bandwidth_GBs"] = round(N * b_cell * steps / wall / 1e9, 1)
    out["cpu_baseline"] = cpu
    out["parity"] = parity
    return out


# ---------------------------------------------------------------------------------------
def percentiles(lat):
    lat = np.sort(np.asarray(lat, dtype=float))
    if len(lat) == 0:
        return 0.0, 0.0
    k = lambda p: lat[min(len(lat), max(1, int(np.ceil(p / 100 * len(lat))))) - 1]
    return float(k(50)), float(k(99))


def bench_batched_2d(args, lib):
    from maxwellfdtd_amd import _lib
    from maxwellfdtd_amd.batch import Batch
    from maxwellfdtd_amd.models import Handle
    from maxwellfdtd_amd.workloads import algorithmic_bytes, batched_2d

    t0 = time.perf_counter()
    models = list(batched_2d(count=args.count, smin=max(2, int(2000 * args.step_scale)),
                             smax=max(2, int(10000 * args.step_scale))))
    descs = [m.describe() for m in models]
    step_bytes = np.array([algorithmic_bytes(m) for m in models])
    del models
    cells = np.array([int(np.prod(d.dims)) for d in descs], dtype=float)
    work = float(np.sum(cells * np.array([d.Nt for d in descs])))
    log(f"{len(descs)} instances built in {time.perf_counter()-t0:.1f}s: "
        f"{cells.sum()/1e6:.1f} M cells, {work/1e12:.3f} T cell-updates")

    # batched: one batch, autotuned launch shape, solves retire as they finish.  The
    # autotuner's trial steps are real steps; throughput counts the steps after tuning
    # over the time after tuning (as for the plain runs below), latency runs from the
    # batch's start
    t0 = time.perf_counter()
    b = Batch(descs)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    tune = b.autotune(trial_steps=8)
    t_tune = time.perf_counter() - t0
    it_tuned = np.array([h.it for h in b.handles], dtype=float)
    log(f"autotune: {tune} in {t_tune:.2f} s")
    done_t = np.zeros(len(descs))
    while b.active > 0:
        try:
            b.run(args.chunk)
        except _lib.EngineError as e:      # a solve went non-finite; the others go on
            log(f"batch: {e}")
        now = time.perf_counter() - t0
        for i, h in enumerate(b.handles):
            if done_t[i] == 0 and (h.it >= descs[i].Nt or h.status != 0):
                done_t[i] = now
    failed = [i for i, h in enumerate(b.handles) if h.status != 0]
    if failed:
        log(f"non-finite solves: {failed} dims {[descs[i].dims for i in failed]} "
            f"Nt {[descs[i].Nt for i in failed]} it {[b.handles[i].it for i in failed]}")
    wall_b = time.perf_counter() - t0
    done_steps = np.array([h.it for h in b.handles], dtype=float)
    work = float(np.sum(cells * done_steps))
    run_b = wall_b - t_tune
    g_batch = float(np.sum(cells * (done_steps - it_tuned))) / run_b / 1e9
    bytes_b = float(np.sum(step_bytes * (done_steps - it_tuned)))
    p50_b, p99_b = percentiles(done_t * 1e3)
    log(f"batched: {wall_b:.2f} s ({t_build:.1f} s host build+copy outside), {g_batch:.2f} Gcell/s, "
        f"p50 {p50_b:.0f} ms p99 {p99_b:.0f} ms")
    b.close()

    # plain: the same solves one at a time, each a batch of one tuned by the same
    # autotuner over the same candidate launch shapes; its throughput counts the steps
    # after tuning over the time after tuning (the tuning is left out, in the plain
    # runs' favour), its latency is the queue's wall clock
    t_start = time.perf_counter()
    busy = 0.0
    lat_p = []
    nplain = len(descs) if args.plain_count <= 0 else min(args.plain_count, len(descs))
    work_p = 0.0
    shapes = {}
    for j, d in enumerate(descs[:nplain]):
        if j % 32 == 0:
            log(f"plain: solve {j} of {nplain}, {busy:.1f} s so far")
        one = Batch([d])
        t_one = one.autotune(trial_steps=8)
        key = (t_one["zchunk"], t_one["rows"], t_one["persistent"])
        shapes[key] = shapes.get(key, 0) + 1
        h = one.handles[0]
        it0 = h.it
        t1 = time.perf_counter()
        try:
            while one.active > 0:
                one.run(0)
        except _lib.EngineError as e:
            log(f"plain: {e}")
        busy += time.perf_counter() - t1
        lat_p.append((time.perf_counter() - t_start) * 1e3)
        work_p += float(np.prod(d.dims)) * (h.it - it0)
        one.close()
    g_plain = work_p / busy / 1e9
    p50_p, p99_p = percentiles(lat_p)
    log(f"plain ({nplain} solves): {busy:.2f} s GPU-busy, {g_plain:.2f} Gcell/s, p50 {p50_p:.0f} ms "
        f"p99 {p99_p:.0f} ms")
    out = base_record(args, g_batch, run_b, int(max(d.Nt for d in descs)))
    out["config"] = {"workload": f"batched_2d: {len(descs)} uneven 2D instances, Nx,Nz in "
                                 f"256..2048, Lorentz rod + dielectrics, steps 2000..10000"
                                 + (f" x {args.step_scale}" if args.step_scale != 1.0 else ""),
                     "instances": len(descs), "cells": int(cells.sum()), "cell_updates": work,
                     "autotune": tune, "autotune_s": round(t_tune, 3)}
    achieved = bytes_b / run_b / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "whole batch: algorithmic bytes of every solve's steps (base, "
                                 "CPML psi, Lorentz state) over the batch's wall time after tuning"}
    out["batched"] = {"gcells_per_s": round(g_batch, 3), "wall_s": round(wall_b, 3),
                      "p50_ms": round(p50_b, 1), "p99_ms": round(p99_b, 1),
                      "latency": "batch start -> solve finished, all submitted at t=0, "
                                 f"observed every {args.chunk} steps"}
    out["plain"] = {"gcells_per_s": round(g_plain, 3), "solves": nplain, "busy_s": round(busy, 3),
                    "p50_ms": round(p50_p, 1), "p99_ms": round(p99_p, 1),
                    "latency": "queue of all solves at t=0 run one after another",
                    "launch_shapes": {f"zchunk={k[0]},rows={k[1]},persistent={k[2]}": v
                                      for k, v in sorted(shapes.items())},
                    "tuning": "each solve autotuned over the batch's candidate shapes; "
                              "throughput over the steps after tuning"}
    out["speedup_vs_plain"] = round(g_batch / g_plain, 3)
    out["nonfinite_solves"] = failed
    return out


def bench_served_3d(args, lib):
    from maxwellfdtd_amd.batch import Server, replay
    from maxwellfdtd_amd.models import Handle
    from maxwellfdtd_amd.workloads import algorithmic_bytes, served_3d

    t0 = time.perf_counter()
    reqs, job_bytes = [], 0.0
    for k, (t_arr, m) in enumerate(served_3d(count=args.count, smin=max(2, int(1000 * args.step_scale)),
                                             smax=max(2, int(4000 * args.step_scale)))):
        job_bytes += algorithmic_bytes(m) * m.Nt
        reqs.append((t_arr, m.describe()))
        if k % 32 == 31:
            log(f"built {k + 1} of {args.count} requests")
    cells = np.array([int(np.prod(d.dims)) for _, d in reqs], dtype=float)
    work = float(np.sum(cells * np.array([d.Nt for _, d in reqs])))
    log(f"{len(reqs)} requests built in {time.perf_counter()-t0:.1f}s: {work/1e12:.3f} T cell-updates, "
        f"arrivals over {reqs[-1][0]:.2f} s")

    srv = Server(options={8: args.quantum, 9: args.fill, 13: args.sched})
    ndone = [0]

    def on_done(ticket, handle, lat_ms):
        ndone[0] += 1
        if ndone[0] % 16 == 0:
            log(f"served: {ndone[0]} of {len(reqs)} requests done after {time.perf_counter() - t0:.1f} s")

    t0 = time.perf_counter()
    replay(srv, reqs, describe=lambda d: d, on_done=on_done)
    wall_s = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    log(f"served: {wall_s:.2f} s, {st}")

    # plain: one request at a time in arrival order (a FIFO queue in front of a single
    # solver); service times measured, waiting derived from the arrival times
    clock, busy, lat_p = 0.0, 0.0, []
    for j, (t_arr, d) in enumerate([] if args.no_plain else reqs):
        if j % 16 == 0:
            log(f"plain: request {j} of {len(reqs)}, {busy:.1f} s busy so far")
        h = Handle(d, lib)
        h.lib.check(h.lib.model_upload(h.h), "upload")
        t1 = time.perf_counter()
        h.step(d.Nt)
        svc = time.perf_counter() - t1
        h.close()
        start = max(clock, t_arr)
        clock = start + svc
        busy += svc
        lat_p.append((clock - t_arr) * 1e3)
    out = base_record(args, st["gcells_per_s"], wall_s, 0)
    out["config"] = {"workload": f"served_3d: first {len(reqs)} requests of the Poisson trace "
                                 "(20 req/s), N^3 with N in 64..320, Kerr + 6-photon plasma "
                                 "sphere, TFSF 800 nm, steps 1000..4000"
                                 + (f" x {args.step_scale}" if args.step_scale != 1.0 else ""),
                     "requests": len(reqs), "cell_updates": work, "quantum_steps": args.quantum,
                     "fill_mcells": args.fill,
                     "server_admission": "least work first" if args.sched else "arrival order"}
    out["served"] = {"gcells_per_s": round(st["gcells_per_s"], 3), "p50_ms": round(st["p50_ms"], 1),
                     "p99_ms": round(st["p99_ms"], 1), "mean_ms": round(st["mean_ms"], 1),
                     "done": st["done"], "failed": st["failed"], "busy_s": round(st["busy_s"], 3)}
    achieved = job_bytes / st["busy_s"] / 1e9 if st["busy_s"] > 0 else 0.0
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "whole trace: algorithmic bytes of every request's steps (base, "
                                 "CPML psi, plasma state) over the server's busy time (first "
                                 "admission to last completion)"}
    if args.no_plain:
        return out
    span = clock - reqs[0][0]
    g_plain = work / span / 1e9
    p50_p, p99_p = percentiles(lat_p)
    log(f"plain: {busy:.2f} s busy over a {span:.2f} s span, {g_plain:.2f} Gcell/s, "
        f"p50 {p50_p:.0f} ms p99 {p99_p:.0f} ms")
    out["plain"] = {"gcells_per_s": round(g_plain, 3), "p50_ms": round(p50_p, 1),
                    "p99_ms": round(p99_p, 1), "busy_s": round(busy, 3), "span_s": round(span, 3),
                    "throughput": "cell-updates / (last completion - first arrival), "
                                  "as for the server"}
    out["speedup_vs_plain"] = round(st["gcells_per_s"] / g_plain, 3) if g_plain else None
    out["latency_gain_vs_plain"] = {"p50": round(p50_p / st["p50_ms"], 3), "p99": round(p99_p / st["p99_ms"], 3)}
    return out


def bench_slab_1d(args, lib):
    """The 1D configuration, run as the reference defines it: FP64 state (MFD_OPT_F64,
    --f32 for the FP32 state), the whole 8000 steps on the GPU and on the oracle."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle_util import field_energy, field_errors, oracle, rel_l2, spectra_errors
    from maxwellfdtd_amd import _lib
    from maxwellfdtd_amd.models import Handle, MONITOR_COMPONENTS
    from maxwellfdtd_amd.workloads import slab_1d
    model, mons = slab_1d()
    desc = model.describe(monitors=mons)
    N = int(np.prod(desc.dims))
    f64 = not args.f32
    h = Handle(desc, lib)
    if f64:
        lib.check(lib.model_set_options(h.h, _lib.MFD_OPT_F64, 1), "set_options(F64)")
    lib.check(lib.model_upload(h.h), "upload")
    t0 = time.perf_counter()
    h.step(desc.Nt)
    wall = time.perf_counter() - t0
    cpu = parity = None
    if not args.no_cpu:
        orc = oracle("f64")
        threads = orc.set_threads(1)
        ho = Handle(desc, orc)
        t1 = time.perf_counter()
        ho.step(desc.Nt)
        cw = time.perf_counter() - t1
        cpu = {"value": round(N * desc.Nt / cw / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
               "kind": "port", "sample": f"all {desc.Nt} steps of the same model, FP64, 1 thread "
                                         "(oracle/fdtd_oracle.c; the 1D update runs on one core)",
               "host": host_cpu_info()}
        errs = field_errors(h, ho, 1)
        for k in range(2):
            errs[("Sa", "E2")[k]] = rel_l2(h.output(k), ho.output(k))
        pm = np.concatenate([np.ravel(h.monitor(0, c)) for c in MONITOR_COMPONENTS[1]])
        po = np.concatenate([np.ravel(ho.monitor(0, c)) for c in MONITOR_COMPONENTS[1]])
        errs["point_monitor"] = rel_l2(pm, po)
        ge, gh = h.energy()
        oe, oh = field_energy(ho)
        parity = {"steps": desc.Nt, "reference": "FP64 oracle, all steps",
                  "rel_l2_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
                  "energy_rel_err": float(f"{abs((ge + gh) / (oe + oh) - 1):.3e}"),
                  **spectra_errors(h, ho, desc)}
        log(f"parity over all {desc.Nt} steps: {parity}")
        ho.close()
    h.close()
    out = base_record(args, N * desc.Nt / wall / 1e9, wall, desc.Nt, dtype="f64" if f64 else "f32")
    out["config"] = {"workload": "slab_1d: Nz=4000, dz=40 nm, CN=0.99, 8000 steps, 10-cell CPML, eps=2.25 "
                                 "slab on cells 2000-2600, soft 800 nm 10 fs Gaussian at 500, point + "
                                 "32-wavelength DFT monitor at 3500", "cells": N,
                     "state": "FP64" if f64 else "FP32"}
    out["roofline"] = None
    out["cpu_baseline"] = cpu
    out["parity"] = parity
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="plane_wave_spheres_3d",
                    choices=["plane_wave_spheres_3d", "dispersive_3d", "batched_2d", "served_3d",
                             "slab_1d"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-integrals", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=16)
    ap.add_argument("--zchunk", type=int, default=None,
                    help="z planes per tile (default: 10 for the fused step, 16 for the other forms)")
    ap.add_argument("--zchunk-h", type=int, default=8, help="H half z planes per tile (0 = --zchunk)")
    ap.add_argument("--xcd-block", type=int, default=64,
                    help="tiles per block dealt to the XCDs (0: one chunk per XCD; default 64)")
    ap.add_argument("--xcd-h", type=int, default=1,
                    help="H tiles in per-XCD contiguous chunks (1) or in grid order (0)")
    ap.add_argument("--ty", type=int, default=None,
                    help="rows per lane / workgroup (default: 2 for the fused step, 4 for the other forms)")
    ap.add_argument("--kform", type=int, default=3,
                    help="3D kernel form: 3 fused step (default), 2 two-pass, 0 one cell per thread")
    ap.add_argument("--snake", type=int, default=1)
    ap.add_argument("--dless", type=int, default=1, help="plain cells keep no D (default 1)")
    ap.add_argument("--tfsf-fold", type=int, default=1,
                    help="fused step stores the TFSF x-face H corrections itself (0: the face launch does)")
    ap.add_argument("--pml-split", type=int, default=0,
                    help="tiles outside every CPML layer run the slab-free kernels (default 0)")
    ap.add_argument("--count", type=int, default=None, help="instances / requests")
    ap.add_argument("--step-scale", type=float, default=1.0)
    ap.add_argument("--chunk", type=int, default=200)
    ap.add_argument("--plain-count", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true",
                    help="served_3d: skip the one-request-at-a-time comparison")
    ap.add_argument("--quantum", type=int, default=64)
    ap.add_argument("--fill", type=int, default=16, help="server: active cells bound, Mi-cells")
    ap.add_argument("--f32", action="store_true",
                    help="slab_1d: FP32 state instead of the configuration's FP64")
    ap.add_argument("--sched", type=int, default=1,
                    help="server admission: 0 arrival order, 1 least work first (default 1)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU z-slab sharding is "
                                   "a design only, so --gpus must be 1"}))
        sys.exit(2)
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 2000
    if args.count is None:
        args.count = 512 if args.workload == "batched_2d" else 256
    # the 3D tile shape: 10-plane tiles of 2 rows for the fused step (the engine's own
    # choice at these sizes), 16 x 4 for the two-pass forms
    if args.zchunk is None:
        args.zchunk = 10 if args.kform == 3 else 16
    if args.ty is None:
        args.ty = 2 if args.kform == 3 else 4

    from maxwellfdtd_amd import _lib
    lib = _lib.engine()
    free, total, cus = C.c_int64(), C.c_int64(), C.c_int32()
    name = C.create_string_buffer(128)
    lib.check(lib.device_info(C.byref(free), C.byref(total), C.byref(cus), name, 128), "device_info")
    log(f"device {name.value.decode()} CUs={cus.value} HBM free={free.value/2**30:.1f} GiB")
    fn = {"plane_wave_spheres_3d": bench_3d, "dispersive_3d": bench_dispersive_3d,
          "batched_2d": bench_batched_2d, "served_3d": bench_served_3d,
          "slab_1d": bench_slab_1d}[args.workload]
    out = fn(args, lib)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
