"""Benchmark of the MI355X FDTD engine (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--precision fp32]

A step is one full FDTD timestep (Engine::IterateTS(1): voltage and current
updates with the fused UPML, excitation and probe hooks) over the whole
domain.  The default workload is the graded-mesh lumped-port case: 256^3
cells (257^3 mesh lines) with cell sizes U[0.5,2] mm per axis, 8-cell UPML on
all faces, 32 random lossy dielectric boxes, a 50 Ohm lumped port with a
Gaussian (f0 = fc = 5 GHz) excitation, 10000 steps.  MCells/s counts mesh-line
products per second (openems.cpp:1337-1338).

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
    "large_dipole": ("large_dipole_scenario", {"num_lines": (1537, 1537, 769), "n_boxes": 200, "steps": 3000, "verbose": True},
                     "1537x1537x769 lines (1.82 G cells) at 1 mm, UPML-8, 200 random dielectric boxes (eps_r ~ U[1,12]), "
                     "z dipole of 20 edges (Gaussian f0=fc=2 GHz), FP64 DFT of E and H at 16 frequencies on the six faces "
                     "of an nf2ff box 10 cells inside the UPML, streamed compressed operator build"),
    "large_dipole_385": ("large_dipole_scenario", {"num_lines": (385, 385, 193), "n_boxes": 200, "steps": 3000},
                         "the large dipole workload on 385x385x193 lines (28.6 M cells)"),
    "batched_64": ("batched_port_scenarios", {"count": 64, "steps": 5000},
                   "64 independent domains, per-axis cells U{48..160} at 1 mm, UPML-8, 8 eps boxes, 50 ohm port (3/3 GHz), "
                   "5000 steps, one launch per update phase for all 64 (graph-captured), vs 64 sequential single-domain runs"),
}


def kernel_bytes(eng, data):
    """Algorithmic HBM bytes per cell of one launch of the voltage (or current)
    update kernel: read 3 own + 3 other field components, write 3 own, plus
    the coefficients (12 floats in full mode, a 2- or 4-byte index with the
    compressed table)."""
    fb = 8 if eng.precision == 1 else (4 if eng.precision == 0 else 2)
    # coefficients are FP32 for the reduced-precision storage modes
    cb = 8 if eng.precision == 1 else 4
    field = 9 * fb
    if eng.compressed:
        coef = 2 if (data.table is not None and len(data.table) <= 65536) else 4
    else:
        coef = 6 * cb
    return field + coef


def read_traffic(kernel, workload, precision="fp32"):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel, workload and storage precision, if there is one."""
    path = os.path.join(ROOT, "profiles", "%s_%s_pmc.md" % (kernel, workload))
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
    return {"value": round(val, 2), "unit": "MCells/s", "cores": int(b.nthreads), "kind": "port",
            "sample": "%d steps of the same workload (restated Engine_SSE_Compressed + Engine_Multithread, %d threads, %.1f s)" % (steps, b.nthreads, dt)}


def scalar_baseline_leg(data, cells, steps):
    from oracle.oracle import oracle_from_operator
    o = oracle_from_operator(data, ftz=True)
    o.iterate(1)
    t0 = time.perf_counter()
    o.iterate(steps)
    dt = time.perf_counter() - t0
    return {"value": round(cells * steps / dt / 1e6, 2), "unit": "MCells/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same workload on the scalar restatement of the basic engine with every hook "
                      "(one thread, %.1f s)" % (steps, dt)}


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
    best, trials = tune_batch(datas, precision=args.precision)
    log("autotuned batch block size: %d threads (ms/step %s)" % (best, {k: round(v, 4) for k, v in trials.items()}))
    engines = [engine_from_operator(d, precision=args.precision) for d in datas]
    b = Batch(engines)
    b.set_block(best)
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
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
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
        "config": {"workload": "batched_64", "description": desc, "cells": cells, "domains": len(datas), "block": best,
                   "sequential_value": round(seq_value, 2), "batched_speedup": round(value / seq_value, 3), "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="graded_port_256", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16", "bf16"])
    ap.add_argument("--coeffs", default="auto", choices=["auto", "full", "compressed"])
    ap.add_argument("--cpu-steps", type=int, default=None, help="steps of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-timing-pass", action="store_true")
    ap.add_argument("--variant", type=int, default=None, help="update-kernel shape (fdtd_set_tuning), default: the engine's")
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the FP64-engine error check")
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
    steps = args.steps if args.steps is not None else kw["steps"]
    kw["steps"] = max(steps, kw["steps"])
    t0 = time.time()
    sc = getattr(scenarios, builder)(**kw)
    data = sc.data
    cells = data.cells
    log("workload %s: %d cells, dT=%.4e s, operator build %.1f s" % (args.workload, cells, data.dT, time.time() - t0))

    compressed = {"auto": None, "full": False, "compressed": True}[args.coeffs]
    scales = None
    if args.precision in ("fp16", "bf16"):
        t0 = time.time()
        # streamed (large) workloads: the pilot covers the excitation only,
        # where the fields near the source peak
        pilot = max(64, data.excitation.length) if data.vv is None else None
        scales = pilot_scales(data, steps=pilot, compressed=compressed, samples=2 if data.vv is None else 8)
        log("storage scales (pilot FP32 run, %.1f s): volt %s curr %s" % (time.time() - t0, scales[0], scales[1]))
    eng = engine_from_operator(data, precision=args.precision, compressed=compressed, scales=scales)
    if args.variant is not None:
        eng.set_tuning(args.variant, 4, 1)
    log("device %s, coefficients %s, device bytes %.2f GB" % (device_name(), "compressed" if eng.compressed else "full", eng.device_bytes() / 1e9))

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
        # the two update kernels, on the engine's stream
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

    cpu = None
    cpu_steps = args.cpu_steps
    if cpu_steps is None:
        cpu_steps = max(10, int(1.5e9 / cells))   # about 10-30 s of host work at ~100 MCells/s
    if data.vv is None:
        log("cpu baseline skipped: the streamed build keeps no full coefficient arrays for the CPU engines")
    elif cpu_steps > 0 and args.precision == "fp32" and not data.ade and not data.tfsf:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline_leg(data, sc.meta.get("interval", 0), cells, threads, cpu_steps)
        log("cpu baseline: %.1f MCells/s on %d threads" % (cpu["value"], cpu["cores"]))
    elif cpu_steps > 0 and args.precision == "fp32":
        # the multithreaded restatement has no ADE/TFSF hooks: time the scalar
        # oracle (basic engine loop order, one thread) on the same workload
        cpu = scalar_baseline_leg(data, cells, args.cpu_steps if args.cpu_steps else max(3, int(5e8 / cells)))
        log("cpu baseline (scalar): %.1f MCells/s" % cpu["value"])

    copy_gbs = copy_bandwidth(1 << 30, 10)
    if roofline is not None:
        roofline["copy_kernel_GBs"] = round(copy_gbs, 1)
    log("plain copy kernel: %.1f GB/s" % copy_gbs)

    err = None
    if data.vv is None:
        if args.precision in ("fp16", "bf16") and not args.no_fp64_check:
            err = fp32_check(data, eng, sc)
    elif not args.no_fp64_check and args.precision != "fp64":
        err = fp64_check(builder, kw, eng, sc, steps, args)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "MCells/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {"fp32": "f32", "fp64": "f64", "fp16": "f16-store/f32", "bf16": "bf16-store/f32"}[args.precision],
        "data": "synthetic (seeded geometry)",
        "config": {"workload": args.workload, "description": desc, "cells": cells,
                   "coefficients": "compressed" if eng.compressed else "full", "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu, "error_vs_fp64": err,
    }
    print(json.dumps(out), flush=True)


def fp32_check(data, eng, sc):
    """Streamed (float-only) workloads have no FP64 operator: compare a reduced-
    precision run with the FP32 engine on the same operator, probe series and
    the DFT accumulators of every field box."""
    import numpy as np
    from wipems_amd.engine import engine_from_operator
    t0 = time.time()
    e32 = engine_from_operator(data, precision="fp32")
    total = eng.GetNumberOfTimesteps()
    t1 = time.perf_counter()
    done = 0
    while done < total:                 # in chunks, with a progress line each
        n = min(1000, total - done)
        e32.iterate(n)
        e32.synchronize()
        done += n
        log("fp32 reference: %d of %d steps" % (done, total))
    wall = time.perf_counter() - t1
    out = {"reference": "fp32 engine", "steps": total,
           "fp32_ms_per_step": round(1e3 * wall / total, 4), "fp32_MCells_s": round(data.cells * total / wall / 1e6, 1)}
    for pid, p in enumerate(data.probes):
        _, v = eng.read_probe(pid)
        _, v32 = e32.read_probe(pid)
        out["probe_%s_rel_l2" % (p.name or pid)] = float(np.linalg.norm(v - v32) / max(np.linalg.norm(v32), 1e-300))
    worst = 0.0
    for k, b in enumerate(data.dft_boxes):
        a, r = eng.read_dft(k), e32.read_dft(k)
        worst = max(worst, float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300)))
    if data.dft_boxes:
        out["dft_max_rel_l2"] = worst
    e32.close()
    log("fp32 check (%.1f s): %s" % (time.time() - t0, out))
    return out


def fp64_check(builder, kw, eng, sc, steps, args):
    """Run the same workload on an FP64 handle (double operator build, double
    fields and coefficients) and compare the probe series of the timed engine;
    for a lumped port also max |S11 - S11_fp64| over the excited band."""
    import numpy as np
    from wipems_amd import scenarios
    from wipems_amd.engine import engine_from_operator
    t0 = time.time()
    kw64 = dict(kw)
    kw64["prec"] = 1
    sc64 = getattr(scenarios, builder)(**kw64)
    e64 = engine_from_operator(sc64.data, precision="fp64", compressed=False)
    total = eng.GetNumberOfTimesteps()
    e64.iterate(total)
    e64.synchronize()
    out = {"steps": total}
    for pid, p in enumerate(sc.data.probes):
        t, v = eng.read_probe(pid)
        t64, v64 = e64.read_probe(pid)
        out["probe_%s_rel_l2" % (p.name or pid)] = float(np.linalg.norm(v - v64) / max(np.linalg.norm(v64), 1e-300))
    if sc.port is not None:
        f0, fc = sc.meta["f0"], sc.meta["fc"]
        freqs = np.linspace(max(f0 - fc, 0.0) + 0.05 * fc, f0 + fc, 96)
        tu, u = eng.read_probe(0); ti, i = eng.read_probe(1)
        tu6, u6 = e64.read_probe(0); ti6, i6 = e64.read_probe(1)
        s = scenarios.s11_from_probes(tu, u, ti, i, freqs)
        s6 = scenarios.s11_from_probes(tu6, u6, ti6, i6, freqs)
        out["s11_max_abs_err"] = float(np.max(np.abs(s - s6)))
        out["band_GHz"] = [round(freqs[0] / 1e9, 3), round(freqs[-1] / 1e9, 3)]
    e64.close()
    log("fp64 check (%.1f s): %s" % (time.time() - t0, out))
    return out


if __name__ == "__main__":
    main()
