"""Scenario builders for the benchmark configurations and the parity tests.

Each builder describes a simulation through ``Operator`` (mesh, boundary
conditions, materials, sources, probes) and returns the built
``OperatorData``.  Geometry is drawn from ``numpy.random.default_rng`` with
the stated seeds, so every run sees the same inputs.

Mesh sizes are counted in mesh lines, as openEMS counts cells
(FDTD/operator.cpp:502-507): "64x64x64 cells" is 65 lines per axis.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .operator import BC_PEC, BC_PML, Operator, OperatorData, ProbeDef

SEED_BASE = 20261203


@dataclass
class Scenario:
    name: str
    data: OperatorData
    steps: int
    op: Operator
    port: dict = None          # lumped port description for S11 post-processing
    meta: dict = None


def _finish(op, data, probes):
    data.probes = list(probes)
    return data


def cavity_point_source(n_cells=64, steps=2000, prec=0, delta_mm=1.0, f0=0.0, fc=10e9, bc=BC_PEC, factor=0.99):
    """Plain vacuum cube, PEC walls, one soft Ez edge at the centre driven by a
    Gaussian pulse; voltage probe on that edge, current probe around it.
    With the defaults this is the 64x64x64-cell / 1 mm / 2000-step case."""
    op = Operator()
    lines = np.arange(n_cells + 1, dtype=np.float64) * delta_mm
    op.set_mesh([lines, lines, lines], 1e-3)
    op.set_boundary([bc] * 6, [8] * 6)
    op.set_timestep_factor(factor)
    op.set_gauss_excite(f0, fc)
    c = lines[n_cells // 2]
    d = lines[n_cells // 2 + 1]
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (c, c, c), (c, c, d), priority=10)
    data = op.build(prec, nr_ts=steps)
    interval = max(1, data.nyquist // 4)
    probes = [
        op.voltage_probe((c, c, c), (c, c, d), 1.0, td_interval=interval, name="ut"),
        op.current_probe((c, c, 0.5 * (c + d)), (c, c, 0.5 * (c + d)), 2, 1.0, td_interval=interval, name="it"),
    ]
    _finish(op, data, probes)
    return Scenario("cavity_%d" % n_cells, data, steps, op, meta={"interval": interval})


def plane_wave_box(n_cells=40, steps=400, prec=0, pml=8, margin=6, delta_mm=1.0, prop_dir=(1.0, 0.0, 0.0),
                   e_vec=(0.0, 0.0, 1.0), f0=5e9, fc=5e9, phase_freq=0.0, graded=False, seed=SEED_BASE + 3):
    """Vacuum cube with UPML on all faces and a total-field/scattered-field
    plane wave on a box `margin` cells inside the PML (openEMS
    AddPlaneWaveExcite, Operator_Ext_TFSF); Gaussian pulse f0/fc.  Probes:
    the E edge of the polarization at the box centre and at a point in the
    scattered-field region between box and PML."""
    if graded:
        rng = np.random.default_rng(seed)
        lines = [graded_lines(rng, n_cells, 0.7, 1.3) for _ in range(3)]
    else:
        lines = [np.arange(n_cells + 1, dtype=np.float64) * delta_mm for _ in range(3)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PML if pml > 0 else BC_PEC] * 6, [pml] * 6)
    op.set_gauss_excite(f0, fc)
    lo = [lines[n][pml + margin] for n in range(3)]
    hi = [lines[n][-1 - pml - margin] for n in range(3)]
    op.set_plane_wave(lo, hi, prop_dir, e_vec, phase_freq)
    data = op.build(prec, nr_ts=steps)
    ez = int(np.argmax(np.abs(e_vec)))
    c = n_cells // 2
    pc = [lines[n][c] for n in range(3)]
    pc2 = list(pc)
    pc2[ez] = lines[ez][c + 1]
    out = [lines[n][pml + margin // 2] for n in range(3)]
    out2 = list(out)
    out2[ez] = lines[ez][pml + margin // 2 + 1]
    probes = [
        op.voltage_probe(tuple(pc), tuple(pc2), 1.0, td_interval=1, name="u_total"),
        op.voltage_probe(tuple(out), tuple(out2), 1.0, td_interval=1, name="u_scat"),
    ]
    _finish(op, data, probes)
    return Scenario("plane_wave_%d" % n_cells, data, steps, op,
                    meta={"box": (lo, hi), "lines": lines, "centre": pc, "pol_axis": ez})


def dispersive_scenario(num_lines=(385, 385, 193), delta_mm=0.25, steps=20000, prec=0, seed=SEED_BASE + 3, pml=8,
                        n_spheres=16, f0=30e9, fc=30e9, drude=True, lorentz=True, material_scale=1.0, ts_factor=1.0):
    """Drude slab and Lorentz spheres lit by a TFSF plane wave: uniform
    0.25 mm mesh of num_lines lines (defaults: 385 x 385 x 193), UPML-8.
    Slab: eps_inf = 1, f_p = 2e15/(2 pi) Hz, tau = 1e-13 s, across the
    interior in x and y, z in [0.55, 0.65] of the domain.  Spheres:
    eps_inf ~ U[1,3], d_eps ~ U[1,5], f_0 ~ U[50,300] THz, tau = 1e-13 s,
    as EpsPlasmaFrequency = f_0 sqrt(d_eps/eps_inf), EpsLorPoleFrequency =
    f_0.  Plane wave along +z, E along x, Gaussian f0/fc.  Probes: the x edge
    at the centre of the plane below the slab, and one above it.
    material_scale multiplies every material frequency and divides every
    relaxation time (1.0: the values above); ts_factor is the timestep factor
    (Operator::SetTimestepFactor)."""
    rng = np.random.default_rng(seed)
    lines = [np.arange(n, dtype=np.float64) * delta_mm for n in num_lines]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PML] * 6, [pml] * 6)
    op.set_timestep_factor(ts_factor)
    op.set_gauss_excite(f0, fc)
    ext = [lines[n][-1] for n in range(3)]
    if drude:
        m = op.add_lorentz_material(eps_r=1.0, plasma_freq=material_scale * 2e15 / (2 * math.pi),
                                    relax_time=1e-13 / material_scale)
        lo = (lines[0][pml + 2], lines[1][pml + 2], 0.55 * ext[2])
        hi = (lines[0][-pml - 3], lines[1][-pml - 3], 0.65 * ext[2])
        op.add_box(m, lo, hi, priority=1)
    spheres = []
    if lorentz:
        for k in range(n_spheres):
            eps_inf = rng.uniform(1, 3)
            d_eps = rng.uniform(1, 5)
            f_0 = rng.uniform(50e12, 300e12)
            r = rng.uniform(0.03, 0.08) * min(ext)
            c = [rng.uniform(lines[n][pml] + r, lines[n][-1 - pml] - r) for n in range(3)]
            fs = material_scale * f_0
            m = op.add_lorentz_material(eps_r=eps_inf, plasma_freq=fs * math.sqrt(d_eps / eps_inf),
                                        relax_time=1e-13 / material_scale, lor_pole_freq=fs)
            op.add_sphere(m, tuple(c), r, priority=2)
            spheres.append({"centre": c, "radius": r, "eps_inf": eps_inf, "d_eps": d_eps, "f0": f_0})
    margin = 4
    blo = [lines[n][pml + margin] for n in range(3)]
    bhi = [lines[n][-1 - pml - margin] for n in range(3)]
    op.set_plane_wave(blo, bhi, (0.0, 0.0, 1.0), (1.0, 0.0, 0.0))
    data = op.build(prec, nr_ts=steps)
    interval = max(1, data.nyquist // 4)
    cx, cy = num_lines[0] // 2, num_lines[1] // 2
    zlo = int(0.4 * (num_lines[2] - 1))
    zhi = int(0.8 * (num_lines[2] - 1))
    probes = []
    for z, nm in ((zlo, "u_below"), (zhi, "u_above")):
        a = (lines[0][cx], lines[1][cy], lines[2][z])
        b = (lines[0][cx + 1], lines[1][cy], lines[2][z])
        probes.append(op.voltage_probe(a, b, 1.0, td_interval=interval, name=nm))
    _finish(op, data, probes)
    return Scenario("dispersive_%dx%dx%d" % tuple(num_lines), data, steps, op,
                    meta={"interval": interval, "spheres": spheres, "seed": seed})


def large_dipole_scenario(num_lines=(1537, 1537, 769), steps=20000, seed=SEED_BASE + 5, n_boxes=200, pml=8,
                          f0=2e9, fc=2e9, n_freq=16, box_margin=10, dipole_cells=20, threads=16, max_entries=65536,
                          verbose=False, prec=0, streamed=True):
    """Large uniform 1 mm domain (defaults: 1537 x 1537 x 769 lines, 1.82 G
    cells) with UPML-8, `n_boxes` random lossless dielectric boxes (eps_r ~
    U[1,12]), a z-directed soft dipole of `dipole_cells` edges at the centre
    (Gaussian f0/fc) and an nf2ff recording box `box_margin` cells inside the
    UPML with FP64 DFT at n_freq frequencies over (0, f0+fc].  Built with the
    streamed operator build (compressed coefficients only, float); with
    streamed=False the full build in precision `prec` (0 float, 1 double),
    whose float result equals the streamed one.  Box sizes and positions are
    fractions of the domain, so smaller num_lines give the same geometry
    scaled down."""
    rng = np.random.default_rng(seed)
    lines = [np.arange(n, dtype=np.float64) for n in num_lines]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PML] * 6, [pml] * 6)
    op.set_threads(threads)
    op.set_verbose(1 if verbose else 0)
    op.set_gauss_excite(f0, fc)
    ext = [lines[n][-1] for n in range(3)]
    boxes = []
    for k in range(n_boxes):
        size = [rng.uniform(0.02, 0.12) * ext[n] for n in range(3)]
        a = [rng.uniform(lines[n][pml + 2], lines[n][-pml - 3] - size[n]) for n in range(3)]
        eps = rng.uniform(1.0, 12.0)
        m = op.add_material(eps_r=eps)
        op.add_box(m, tuple(a), tuple(a[n] + size[n] for n in range(3)), priority=1)
        boxes.append({"start": a, "size": size, "eps_r": eps})
    c = [lines[n][len(lines[n]) // 2] for n in range(3)]
    z0 = lines[2][len(lines[2]) // 2 - dipole_cells // 2]
    z1 = lines[2][len(lines[2]) // 2 - dipole_cells // 2 + dipole_cells]
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (c[0], c[1], z0), (c[0], c[1], z1), priority=10)
    data = op.build_streamed(nr_ts=steps, max_entries=max_entries) if streamed else op.build(prec, nr_ts=steps)
    interval = max(1, data.nyquist // 4)
    feed = (c[0], c[1], lines[2][len(lines[2]) // 2])
    feed2 = (c[0], c[1], lines[2][len(lines[2]) // 2 + 1])
    probes = [op.voltage_probe(feed, feed2, 1.0, td_interval=interval, name="u_feed")]
    _finish(op, data, probes)
    lo = [lines[n][pml + box_margin] for n in range(3)]
    hi = [lines[n][-1 - pml - box_margin] for n in range(3)]
    freqs = np.linspace((f0 + fc) / n_freq, f0 + fc, n_freq)
    op.nf2ff_box(data, lo, hi, freqs)
    return Scenario("large_dipole_%dx%dx%d" % tuple(num_lines), data, steps, op,
                    meta={"interval": interval, "boxes": boxes, "seed": seed, "f0": f0, "fc": fc, "freqs": freqs})


def graded_lines(rng, n_cells, lo=0.5, hi=2.0):
    gaps = rng.uniform(lo, hi, n_cells)
    return np.concatenate([[0.0], np.cumsum(gaps)])


def add_lumped_port(op, lines, port_idx, z_span, R=50.0, excite=1.0, priority=100):
    """openEMS LumpedPort (python/openEMS/ports.py:143-194): a z-directed
    parallel-R lumped element with metal caps, a soft E excitation with value
    -direction*excite, a voltage probe along the port (weight -1) and a
    current probe at the mid plane (weight = direction)."""
    xi, yi, zi = port_idx
    x, y = lines[0][xi], lines[1][yi]
    z0, z1 = lines[2][zi], lines[2][zi + z_span]
    start, stop = (x, y, z0), (x, y, z1)
    ny = 2
    direction = 1.0 if stop[ny] > start[ny] else -1.0
    le = op.add_lumped_element(ny, R=R, caps=True)
    op.add_box(le, start, stop, priority=priority)
    exc = op.add_excitation(0, (0.0, 0.0, -direction * excite))
    op.add_box(exc, start, stop, priority=priority)
    zmid = 0.5 * (z0 + z1)
    return {"start": start, "stop": stop, "R": R, "direction": direction, "zmid": zmid}


def port_probes(op, port, interval, fd_interval=0, freqs=()):
    s, e = port["start"], port["stop"]
    u_start = (0.5 * (s[0] + e[0]), 0.5 * (s[1] + e[1]), s[2])
    u_stop = (0.5 * (s[0] + e[0]), 0.5 * (s[1] + e[1]), e[2])
    i_pt = (s[0], s[1], port["zmid"])
    return [
        op.voltage_probe(u_start, u_stop, -1.0, td_interval=interval, fd_interval=fd_interval, freqs=tuple(freqs), name="port_ut"),
        op.current_probe(i_pt, i_pt, 2, port["direction"], td_interval=interval, fd_interval=fd_interval, freqs=tuple(freqs), name="port_it"),
    ]


def random_boxes(rng, op, lines, count, pml, eps_range=(1.0, 12.0), kappa_log=(-3.0, 0.0), min_frac=0.04, max_frac=0.25,
                 avoid=None, priority=1):
    """Random axis-aligned dielectric boxes inside the non-PML interior."""
    boxes = []
    for k in range(count):
        start, stop = [], []
        for n in range(3):
            lo, hi = lines[n][pml], lines[n][-1 - pml]
            span = hi - lo
            size = rng.uniform(min_frac, max_frac) * span
            a = rng.uniform(lo, hi - size)
            start.append(a)
            stop.append(a + size)
        eps = rng.uniform(*eps_range)
        kappa = 10 ** rng.uniform(*kappa_log) if kappa_log is not None else 0.0
        m = op.add_material(eps_r=eps, kappa=kappa)
        op.add_box(m, tuple(start), tuple(stop), priority=priority)
        boxes.append({"start": start, "stop": stop, "eps_r": eps, "kappa": kappa})
    return boxes


def graded_port_scenario(n_cells=256, steps=10000, prec=0, seed=SEED_BASE + 2, n_boxes=32, pml=8, f0=5e9, fc=5e9,
                         kappa_log=(-3.0, 0.0), eps_range=(1.0, 12.0)):
    """Graded mesh (cell sizes U[0.5,2] mm per axis), 8-cell UPML on all faces,
    random lossy dielectric boxes (eps_r ~ U[1,12], kappa ~ LogU[1e-3,1] S/m)
    and a 50 Ohm lumped port with a Gaussian excitation.  Defaults give the
    256^3-cell, 32-box, 10000-step case."""
    rng = np.random.default_rng(seed)
    lines = [graded_lines(rng, n_cells) for _ in range(3)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PML if pml > 0 else BC_PEC] * 6, [pml] * 6)
    op.set_gauss_excite(f0, fc)
    boxes = random_boxes(rng, op, lines, n_boxes, pml, eps_range=eps_range, kappa_log=kappa_log)
    c = n_cells // 2
    port = add_lumped_port(op, lines, (c, c, c - 2), 4)
    data = op.build(prec, nr_ts=steps)
    interval = max(1, data.nyquist // 4)
    probes = port_probes(op, port, interval)
    _finish(op, data, probes)
    return Scenario("graded_port_%d" % n_cells, data, steps, op, port=port,
                    meta={"interval": interval, "boxes": boxes, "seed": seed, "f0": f0, "fc": fc})


def uniform_bench(n_lines=512, steps=1000, prec=0):
    """Uniform vacuum cube of n_lines^3 mesh lines, PEC walls, one soft Ez edge
    at the centre (Gaussian f0=0, fc=10 GHz), 0.99 Courant."""
    return cavity_point_source(n_cells=n_lines - 1, steps=steps, prec=prec)


def dft_time2freq(t, val, freqs):
    """2 dt sum_n val(t_n) exp(-2 pi i f t_n) at each frequency, dt = t[1] - t[0]
    (DFT_time2freq, python/openEMS/utilities.py:21-35)."""
    t = np.asarray(t)
    val = np.asarray(val)
    out = np.empty(len(freqs), dtype=np.complex128)
    for k, fk in enumerate(freqs):
        out[k] = np.sum(val * np.exp(-1j * 2 * np.pi * fk * t))
    return 2 * out * (t[1] - t[0])


def s11_from_probes(t_u, u, t_i, i, freqs, Z_ref=50.0):
    """S11 of a lumped port from its U/I time series, as python/openEMS/ports.py
    Port.CalcPort (:103-141): uf_inc = (U + I Z_ref)/2, uf_ref = U - uf_inc,
    S11 = uf_ref / uf_inc."""
    uf = dft_time2freq(t_u, u, freqs)
    itf = dft_time2freq(t_i, i, freqs)
    uf_inc = 0.5 * (uf + itf * Z_ref)
    uf_ref = uf - uf_inc
    return uf_ref / uf_inc


def batched_port_scenarios(count=64, lo=48, hi=160, steps=5000, prec=0, seed=SEED_BASE + 4, pml=8, n_boxes=8, f0=3e9, fc=3e9):
    """`count` independent domains with per-axis cell counts drawn from
    U{lo..hi} (1 mm uniform mesh), 8-cell UPML on all faces, `n_boxes` random
    lossless dielectric boxes (eps_r ~ U[1,10]) and one 50 Ohm lumped port
    with a Gaussian excitation (f0 = fc = 3 GHz).  Defaults give the 64-domain,
    48..160-cell, 5000-step case."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = [int(v) for v in rng.integers(lo, hi + 1, size=3)]
        lines = [np.arange(m + 1, dtype=np.float64) for m in n]
        op = Operator()
        op.set_mesh(lines, 1e-3)
        op.set_boundary([BC_PML] * 6, [pml] * 6)
        op.set_gauss_excite(f0, fc)
        boxes = random_boxes(rng, op, lines, n_boxes, pml, eps_range=(1.0, 10.0), kappa_log=None)
        c = [m // 2 for m in n]
        port = add_lumped_port(op, lines, (c[0], c[1], c[2] - 2), 4)
        data = op.build(prec, nr_ts=steps)
        interval = max(1, data.nyquist // 4)
        _finish(op, data, port_probes(op, port, interval))
        out.append(Scenario("batch_%d_%dx%dx%d" % (k, n[0], n[1], n[2]), data, steps, op, port=port,
                            meta={"interval": interval, "boxes": boxes, "cells": n}))
    return out
