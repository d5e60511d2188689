"""Benchmark of the MI355X FDTD engine (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--precision fp32]

A step is one full FDTD timestep (Engine::IterateTS(1): voltage and current
updates with the fused UPML, excitation, probe and field-box DFT hooks) over
the whole domain.  The default workload is the large single-GPU dipole case:
1537 x 1537 x 769 mesh lines (1.82 G cells) at 1 mm, 8-cell UPML on all
faces, 200 random dielectric boxes, a Gaussian (f0 = fc = 2 GHz) z dipole and
FP64 DFT at 16 frequencies on the six faces of an nf2ff box, stated for
20000 steps; the operator comes from the streamed build.  MCells/s counts
mesh-line products per second (openems.cpp:1337-1338).

Beside the timed GPU run the line carries
* roofline: algorithmic bytes of the slower update phase over its average
  duration (HIP events on the engine's stream) against 8 TB/s, with the HBM
  traffic of the committed rocprofv3 counter summary;
* cpu_baseline: the restated Engine_SSE_Compressed + Engine_Multithread on
  every host core this process may use, on the same workload for a bounded
  number of steps;
* error_vs_fp64: the same workload run for its full stated length against
  the FP64 engine (for the large dipole on a stated smaller mesh of the same
  geometry: the FP64 run at full size would take most of an hour).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCells/s + % HBM peak at 1 GPU; FP16-store & batched vs plain FP32 at stated err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    # name: (builder, kwargs, description)
    "graded_port_256": ("graded_port_scenario", {"n_cells": 256, "steps": 10000},
                        "256^3 cells (257^3 lines), graded 0.5-2 mm, UPML-8 on 6 faces, 32 lossy eps boxes, 50 ohm lumped port, 10000 steps"),
    "uniform_512": ("uniform_bench", {"n_lines": 512, "steps": 1000},
                    "512^3 lines uniform 1 mm vacuum, PEC, soft Ez Gaussian at centre, 1000 steps"),
    "cavity_64": ("cavity_point_source", {"n_cells": 64, "steps": 2000},
                  "64^3 cells (65^3 lines) 1 mm vacuum, PEC, soft Ez Gaussian, 2000 steps"),
    "dispersive_c3": ("dispersive_scenario", {"steps": 20000, "material_scale": 1e-3, "ts_factor": 0.9},
                      "385x385x193 lines at 0.25 mm, UPML-8, Drude slab + 16 Lorentz spheres (material frequencies x1e-3, "
                      "relaxation times x1e3, timestep factor 0.9: the stated THz/PHz values diverge in the reference ADE "
                      "scheme), TFSF plane wave +z / Ex (30/30 GHz Gaussian), 20000 steps"),
    "large_dipole": ("large_dipole_scenario", {"num_lines": (1537, 1537, 769), "n_boxes": 200, "steps": 20000, "verbose": True},
                     "1537x1537x769 lines (1.82 G cells) at 1 mm, UPML-8, 200 random dielectric boxes (eps_r ~ U[1,12]), "
                     "z dipole of 20 edges (Gaussian f0=fc=2 GHz), FP64 DFT of E and H at 16 frequencies on the six faces "
                     "of an nf2ff box 10 cells inside the UPML, streamed compressed operator build, 20000 steps"),
    "large_dipole_385": ("large_dipole_scenario", {"num_lines": (385, 385, 193), "n_boxes": 200, "steps": 20000},
                         "the large dipole workload on 385x385x193 lines (28.6 M cells)"),
    "batched_64": ("batched_port_scenarios", {"count": 64, "steps": 5000},
                   "64 independent domains, per-axis cells U{48..160} at 1 mm, UPML-8, 8 eps boxes, 50 ohm port (3/3 GHz), "
                   "5000 steps, one launch per update phase for all 64 (graph-captured), vs 64 sequential single-domain runs"),
}


def kernel_bytes(eng, data):
    """Algorithmic HBM bytes per cell of one launch of the voltage (or current)
    update kernel: read 3 own + 3 other field components, write 3 own, plus
    the coefficients (12 floats in full mode, a 2- or 4-byte index with the
    compressed table, a 2-byte index with the factored table; tables and mesh
    factors stay in cache)."""
    fb = 8 if eng.precision == 1 else (4 if eng.precision == 0 else 2)
    # coefficients are FP32 for the reduced-precision storage modes
    cb = 8 if eng.precision == 1 else 4
    field = 9 * fb
    if eng.factored:
        coef = 2
    elif eng.compressed:
        coef = 2 if (data.table is not None and len(data.table) <= 65536) else 4
    else:
        coef = 6 * cb
    return field + coef


def read_traffic(kernel, workload, precision="fp32"):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel, workload and storage precision, if there is one
    (profiles/<kernel>_<workload>_pmc.md for FP32 storage,
    profiles/<kernel>_<workload>_<precision>_pmc.md for the others)."""
    name = "%s_%s_pmc.md" % (kernel, workload) if precision == "fp32" else "%s_%s_%s_pmc.md" % (kernel, workload, precision)
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if ":" in line and not line.startswith(("|", "#", " ")):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get("precision", "fp32") != precision:
        return None
    try:
        return float(fields["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline_leg(data, interval, cells, threads, steps):
    from oracle.cpu_baseline import CpuBaseline
    t0 = time.time()
    b = CpuBaseline(data, threads)
    log("cpu baseline: operator compressed to %d unique 4-lane tuples in %.1f s, %d threads" % (b.unique, time.time() - t0, b.nthreads))
    dt = b.run(steps, interval)
    val = cells * steps / dt / 1e6
    hooks = " with its Lorentz/Drude ADE and TFSF hooks on thread 0" if (data.ade or data.tfsf) else ""
    return {"value": round(val, 2), "unit": "MCells/s", "cores": int(b.nthreads), "kind": "port",
            "sample": "%d steps of the same workload (restated Engine_SSE_Compressed + Engine_Multithread%s, %d threads, %.1f s)" % (
                steps, hooks, b.nthreads, dt)}


def run_batched(args, desc):
    from wipems_amd import scenarios
    from wipems_amd.autotune import tune_batch
    from wipems_amd.engine import Batch, copy_bandwidth, device_name, engine_from_operator

    steps = args.steps if args.steps is not None else 5000
    t0 = time.time()
    scs = scenarios.batched_port_scenarios(count=64, steps=steps)
    datas = [sc.data for sc in scs]
    cells = sum(d.cells for d in datas)
    log("workload batched_64: %d domains, %d cells in total, operator builds %.1f s" % (len(datas), cells, time.time() - t0))
    if args.batch_shape is not None:
        best = tuple(args.batch_shape)
        log("batch shape (threads per block, split) set: %s" % (best,))
    else:
        best, trials = tune_batch(datas, precision=args.precision)
        log("autotuned batch shape (threads per block, split): %s (update ms/step %s)" % (best, {"%d/%d" % k: round(v, 4) for k, v in trials.items()}))
    engines = [engine_from_operator(d, precision=args.precision) for d in datas]
    b = Batch(engines)
    b.set_block(best[0])
    b.set_split(best[1])
    if args.profile_steps:
        b.iterate(args.profile_steps)
        b.synchronize()
        gb = copy_bandwidth(1 << 30, 3)
        log("profile mode: %d steps + 4 launches of the 1 GiB copy kernel (%.0f GB/s)" % (args.profile_steps, gb))
        return
    b.iterate(args.warmup)
    b.synchronize()
    t1 = time.perf_counter()
    b.iterate(steps)
    b.synchronize()
    wall = time.perf_counter() - t1
    value = cells * steps / wall / 1e6
    log("batched: %d steps in %.3f s -> %.1f MCells/s (%.3f ms/step), device %s" % (steps, wall, value, 1e3 * wall / steps, device_name()))
    roofline = None
    if not args.no_timing_pass:
        b.set_kernel_timing(True)
        nt = min(steps, 1000)
        b.iterate(nt)
        tv, tc, n = b.kernel_times()
        b.set_kernel_timing(False)
        avg_v, avg_c = tv / n, tc / n
        # algorithmic bytes of one batched volt launch: per domain, its cells
        # times its per-cell bytes (compressed index or full coefficients)
        alg = 0.0
        for e, d in zip(engines, datas):
            alg += kernel_bytes(e, d) * d.cells
        achieved = alg / (avg_v * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic("volt_batch", "batched_64", args.precision),
                    "kernel": "volt_batch", "avg_ms": round(avg_v, 5), "curr_avg_ms": round(avg_c, 5)}
        log("batched kernel avg: volt %.4f ms, curr %.4f ms -> %.1f GB/s" % (avg_v, avg_c, achieved))
    b.close()
    for e in engines:
        e.close()
    # the same 64 domains run one after another as plain single-domain engines
    seq_steps = min(steps, 500)
    seq_time = 0.0
    for d in datas:
        e = engine_from_operator(d, precision=args.precision)
        e.iterate(5)
        e.synchronize()
        ts = time.perf_counter()
        e.iterate(seq_steps)
        e.synchronize()
        seq_time += time.perf_counter() - ts
        e.close()
    seq_value = cells * seq_steps / seq_time / 1e6
    log("sequential single-domain runs: %.1f MCells/s (%d steps each) -> batched speedup %.2fx" % (seq_value, seq_steps, value / seq_value))
    cpu = None
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else 20
    if cpu_steps > 0 and args.precision == "fp32":
        threads = args.cpu_threads or host_cores()
        from oracle.cpu_baseline import CpuBaseline
        tot, tcells = 0.0, 0
        for d, sc in list(zip(datas, scs))[:8]:
            cb = CpuBaseline(d, threads)
            tot += cb.run(cpu_steps, sc.meta["interval"])
            tcells += d.cells
        cpu = {"value": round(tcells * cpu_steps / tot / 1e6, 2), "unit": "MCells/s", "cores": threads, "kind": "port",
               "sample": "domains 0-7 of the batch, %d steps each, sequentially (restated Engine_SSE_Compressed + Engine_Multithread)" % cpu_steps}
        log("cpu baseline: %.1f MCells/s" % cpu["value"])
    copy_gbs = copy_bandwidth(1 << 30, 10)
    if roofline is not None:
        roofline["copy_kernel_GBs"] = round(copy_gbs, 1)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "MCells/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {"fp32": "f32", "fp64": "f64", "fp16": "f16-store/f32", "bf16": "bf16-store/f32"}[args.precision],
        "data": "synthetic (seeded geometry)",
        "config": {"workload": "batched_64", "description": desc, "cells": cells, "domains": len(datas), "block": best[0], "split": best[1],
                   "sequential_value": round(seq_value, 2), "batched_speedup": round(value / seq_value, 3), "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def host_cores():
    """Threads this process may run on (its CPU affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def mem_available():
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="large_dipole", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16", "bf16"])
    ap.add_argument("--coeffs", default="auto", choices=["auto", "full", "compressed", "factored"])
    ap.add_argument("--split", type=int, default=None, help="UPML launch form (fdtd_set_pml_split): 0 or 1; default: the engine's")
    ap.add_argument("--pml-rows", type=int, default=None, help="UPML coefficient rows (fdtd_set_pml_compression): 0 or 1; default: the engine's")
    ap.add_argument("--fuse-hooks", type=int, default=None, help="fused sparse hooks (fdtd_set_fuse_hooks): 0 or 1; default: the engine's")
    ap.add_argument("--shell-align", type=int, default=None, help="z alignment of the interior launch in vectors (fdtd_set_shell_align); default: the engine's")
    ap.add_argument("--tile-order", type=int, nargs=2, default=None, metavar=("BAND_TILES", "XCD"),
                    help="tile order of the update launches (fdtd_set_tile_order); default: the engine's")
    ap.add_argument("--cpu-steps", type=int, default=None, help="steps of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-timing-pass", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the FP64-engine error check")
    ap.add_argument("--check-lines", type=int, nargs=3, default=(193, 193, 97),
                    help="mesh lines of the smaller mesh the large dipole's FP64 check runs on")
    ap.add_argument("--batch-shape", type=int, nargs=2, default=None, metavar=("THREADS", "SPLIT"),
                    help="batched workload: threads per block and split instead of the autotuner's choice")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="counter-collection mode: run this many steps of the workload, then a 1 GiB device copy "
                         "of known size (the calibration kernel for FETCH_SIZE/WRITE_SIZE), and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one MI355X and the engine has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from wipems_amd import scenarios
    from wipems_amd.engine import copy_bandwidth, device_name, engine_from_operator, pilot_scales

    builder, kw, desc = WORKLOADS[args.workload]
    if args.workload == "batched_64":
        return run_batched(args, desc)
    kw = dict(kw)
    stated = kw["steps"]
    steps = args.steps if args.steps is not None else (40 if args.workload == "large_dipole" else min(stated, 200))
    t0 = time.time()
    sc = getattr(scenarios, builder)(**kw)
    data = sc.data
    cells = data.cells
    log("workload %s: %d cells, dT=%.4e s, operator build %.1f s" % (args.workload, cells, data.dT, time.time() - t0))

    compressed = {"auto": None, "full": False, "compressed": True, "factored": None}[args.coeffs]
    factored = {"factored": True, "full": False, "compressed": False}.get(args.coeffs)
    scales = None
    if args.precision in ("fp16", "bf16"):
        t0 = time.time()
        pilot = max(64, data.excitation.length) if data.vv is None else None
        scales = pilot_scales(data, steps=pilot, compressed=compressed, samples=8)
        log("storage scales (pilot FP32 run, %.1f s): volt %s curr %s" % (time.time() - t0, scales[0], scales[1]))
    eng = engine_from_operator(data, precision=args.precision, compressed=compressed, scales=scales, factored=factored)
    if args.split is not None:
        eng.set_pml_split(args.split)
    if args.pml_rows is not None:
        eng.set_pml_compression(args.pml_rows)
    if args.fuse_hooks is not None:
        eng.set_fuse_hooks(args.fuse_hooks)
    if args.tile_order is not None:
        eng.set_tile_order(*args.tile_order)
    if args.shell_align is not None:
        eng.set_shell_align(args.shell_align)
    log("device %s, coefficients %s, UPML split %s, tile order (band tiles, XCD) %s, device bytes %.2f GB" % (
        device_name(), "factored" if eng.factored else ("compressed" if eng.compressed else "full"),
        eng.pml_split(), eng.tile_order(), eng.device_bytes() / 1e9))

    if args.profile_steps:
        eng.iterate(args.profile_steps)
        eng.synchronize()
        gb = copy_bandwidth(1 << 30, 3)
        log("profile mode: %d steps + 4 launches of the 1 GiB copy kernel (%.0f GB/s)" % (args.profile_steps, gb))
        return

    eng.iterate(args.warmup)
    eng.synchronize()
    t1 = time.perf_counter()
    eng.iterate(steps)
    eng.synchronize()
    t2 = time.perf_counter()
    wall = t2 - t1
    value = cells * steps / wall / 1e6
    log("timed %d steps in %.3f s -> %.1f MCells/s (%.3f ms/step)" % (steps, wall, value, 1e3 * wall / steps))

    roofline = None
    if not args.no_timing_pass:
        # second pass of the same length with HIP events around every launch of
        # the two update phases, on the engine's stream
        eng.set_kernel_timing(True)
        eng.iterate(steps)
        eng.synchronize()
        tv, tc, nv, nc = eng.kernel_times()
        eng.set_kernel_timing(False)
        avg_v, avg_c = tv / nv, tc / nc
        kern, avg = ("volt_update", avg_v) if tv >= tc else ("curr_update", avg_c)
        bpc = kernel_bytes(eng, data)
        alg = bpc * cells
        achieved = alg / (avg * 1e-3) / 1e9
        traffic = read_traffic(kern, args.workload, args.precision)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": kern, "avg_ms": round(avg, 5), "bytes_per_cell": bpc,
                    "volt_avg_ms": round(avg_v, 5), "curr_avg_ms": round(avg_c, 5)}
        log("kernel avg: volt %.4f ms, curr %.4f ms -> %s %.1f GB/s (%.1f%% of %.0f)" % (avg_v, avg_c, kern, achieved, 100 * achieved / HBM_PEAK_GBS, HBM_PEAK_GBS))
    eng.close()

    copy_gbs = copy_bandwidth(1 << 30, 10)
    if roofline is not None:
        roofline["copy_kernel_GBs"] = round(copy_gbs, 1)
    log("plain copy kernel: %.1f GB/s" % copy_gbs)

    cpu = None
    if args.precision == "fp32":
        cpu = cpu_leg(args, data, sc, cells)

    err = None
    if not args.no_fp64_check and args.precision != "fp64":
        if args.workload.startswith("large_dipole"):
            err = subbox_check(args, kw, stated)
        else:
            err = fp64_check(builder, kw, sc, stated, args, compressed, scales, factored)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "MCells/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {"fp32": "f32", "fp64": "f64", "fp16": "f16-store/f32", "bf16": "bf16-store/f32"}[args.precision],
        "data": "synthetic (seeded geometry)",
        "config": {"workload": args.workload, "description": desc, "cells": cells, "stated_steps": stated,
                   "coefficients": "factored" if eng.factored else ("compressed" if eng.compressed else "full"), "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu, "error_vs_fp64": err,
    }
    print(json.dumps(out), flush=True)


def cpu_leg(args, data, sc, cells):
    """The CPU baseline on the same workload: the multithreaded restatement
    (with its thread-0 ADE and TFSF hooks where the workload has them) on
    every core of this process's affinity set."""
    cpu_steps = args.cpu_steps
    if cpu_steps is None:
        cpu_steps = max(4, int(2e10 / cells))   # about 10-30 s of host work at ~1 GCells/s
    if cpu_steps <= 0:
        return None
    threads = args.cpu_threads or host_cores()
    # fields of the 4-lane engine (V and I, 24 B per cell), its 4-lane index and the UPML slabs
    need = int(cells * 24 * 1.1 + cells + sum(48 * 1.1 * s.arrays[0].size for s in data.upml))
    avail = mem_available()
    if avail and need > 0.8 * avail:
        log("cpu baseline skipped: needs %.1f GB of host memory, %.1f GB available" % (need / 1e9, avail / 1e9))
        return None
    out = cpu_baseline_leg(data, sc.meta.get("interval", 0), cells, threads, cpu_steps)
    log("cpu baseline: %.1f MCells/s on %d threads" % (out["value"], out["cores"]))
    return out


def _rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def subbox_check(args, kw, stated):
    """Error of FP32 and FP16 storage against the FP64 engine for the full
    stated number of steps, on a smaller mesh of the same workload (same
    seed, geometry scaled with the domain): feed-probe rel-L2 and the DFT
    accumulators of the twelve nf2ff faces (max rel-L2 over boxes).  The FP32
    and FP16 runs use the streamed float build, the FP64 run the full double
    build."""
    import numpy as np
    from wipems_amd import scenarios
    from wipems_amd.engine import engine_from_operator, pilot_scales
    t0 = time.time()
    lines = tuple(int(v) for v in args.check_lines)
    kws = dict(kw)
    kws.update(num_lines=lines, verbose=False)
    sc32 = scenarios.large_dipole_scenario(**kws)
    sc64 = scenarios.large_dipole_scenario(prec=1, streamed=False, **kws)
    d32, d64 = sc32.data, sc64.data
    runs = {}
    e64 = engine_from_operator(d64, precision="fp64", compressed=False)
    for name, prec in (("fp32", "fp32"), ("fp16", "fp16")):
        scales = pilot_scales(d32, steps=max(64, d32.excitation.length), samples=8) if prec == "fp16" else None
        runs[name] = engine_from_operator(d32, precision=prec, scales=scales)
    done = 0
    while done < stated:
        n = min(2000, stated - done)
        for e in [e64] + list(runs.values()):
            e.iterate(n)
        e64.synchronize()
        done += n
        log("fp64 check on %dx%dx%d lines: %d of %d steps" % (lines + (done, stated)))
    out = {"mesh_lines": list(lines), "cells": d32.cells, "steps": stated}
    _, ref = e64.read_probe(0)
    refd = [e64.read_dft(k) for k in range(len(d64.dft_boxes))]
    for name, e in runs.items():
        _, v = e.read_probe(0)
        worst = max(_rel_l2(e.read_dft(k), refd[k]) for k in range(len(refd)))
        out[name] = {"feed_probe_rel_l2": _rel_l2(v, ref), "dft_faces_max_rel_l2": worst}
        e.close()
    e64.close()
    log("fp64 check (%.1f s): %s" % (time.time() - t0, out))
    return out


def fp64_check(builder, kw, sc, stated, args, compressed, scales, factored=None):
    """Run the workload for its full stated number of steps on the timed
    storage precision and on an FP64 handle (double operator build, double
    fields and coefficients), and compare the probe series; for a lumped port
    also max |S11 - S11_fp64| over the excited band."""
    import numpy as np
    from wipems_amd import scenarios
    from wipems_amd.engine import engine_from_operator
    t0 = time.time()
    kw64 = dict(kw)
    kw64["prec"] = 1
    sc64 = getattr(scenarios, builder)(**kw64)
    e64 = engine_from_operator(sc64.data, precision="fp64", compressed=False)
    e = engine_from_operator(sc.data, precision=args.precision, compressed=compressed, scales=scales, factored=factored)
    done = 0
    while done < stated:
        n = min(2000, stated - done)
        e64.iterate(n)
        e.iterate(n)
        e64.synchronize()
        done += n
        log("fp64 check: %d of %d steps" % (done, stated))
    out = {"steps": stated}
    for pid, p in enumerate(sc.data.probes):
        _, v = e.read_probe(pid)
        _, v64 = e64.read_probe(pid)
        out["probe_%s_rel_l2" % (p.name or pid)] = _rel_l2(v, v64)
    if sc.port is not None:
        f0, fc = sc.meta["f0"], sc.meta["fc"]
        freqs = np.linspace(max(f0 - fc, 0.0) + 0.05 * fc, f0 + fc, 96)
        tu, u = e.read_probe(0); ti, i = e.read_probe(1)
        tu6, u6 = e64.read_probe(0); ti6, i6 = e64.read_probe(1)
        s = scenarios.s11_from_probes(tu, u, ti, i, freqs)
        s6 = scenarios.s11_from_probes(tu6, u6, ti6, i6, freqs)
        out["s11_max_abs_err"] = float(np.max(np.abs(s - s6)))
        out["band_GHz"] = [round(float(freqs[0]) / 1e9, 3), round(float(freqs[-1]) / 1e9, 3)]
    e64.close()
    e.close()
    log("fp64 check (%.1f s): %s" % (time.time() - t0, out))
    return out


if __name__ == "__main__":
    main()
