"""Mur ABC (Operator_Ext_Mur_ABC + Engine_Ext_Mur_ABC), pinned on the CPU oracle.

* Coefficients: k = (c dT - delta) / (c dT + delta) with c = c0 / sqrt(eps_r mu_r)
  of the material half a cell inside the boundary (operator_ext_mur_abc.cpp:117-180),
  or the set phase velocity; stored in the build's float type.
* Start rule: a voltage excitation on the boundary plane delays the face to
  max delay + signal length + 10 (engine_ext_mur_abc.cpp:42-58).
* Absorption: a TEM pulse on a parallel-plate line (PEC top/bottom, PMC
  sides) hits the far end.  Against a 3x longer line, time-gated, the Mur end
  reflects below -50 dB (-65 dB measured); a PEC end reflects at 0 dB, and
  coefficients with the sign flipped reflect far above the Mur level.
"""
import math

import numpy as np
import pytest

from oracle.oracle import oracle_from_operator
from wipems_amd.operator import BC_MUR, BC_PEC, BC_PMC, Operator

C0 = 299792458.0


def line_op(nx, far=BC_MUR, n_t=4):
    op = Operator()
    lines = [np.arange(nx + 1, dtype=float), np.arange(n_t + 1, dtype=float), np.arange(n_t + 1, dtype=float)]
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_MUR, far, BC_PMC, BC_PMC, BC_PEC, BC_PEC], [0] * 6)
    op.set_gauss_excite(5e9, 5e9)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (50.0, 0.0, 0.0), (50.0, float(n_t), float(n_t)), priority=5)
    return op, lines


def probe_series(nx, steps, far=BC_MUR, flip=False):
    op, lines = line_op(nx, far)
    d = op.build(0, nr_ts=steps)
    if flip:
        for m in d.mur:
            m.k_nyP = -m.k_nyP
            m.k_nyPP = -m.k_nyPP
    d.probes = [op.voltage_probe((150.0, 2.0, 1.0), (150.0, 2.0, 3.0), td_interval=1, name="u")]
    o = oracle_from_operator(d)
    o.iterate(steps)
    return o.read_probe(0)[1], d


def reflection_db(u, ref, n):
    return 20 * math.log10(np.abs(u[:n] - ref[:n]).max() / np.abs(ref[:n]).max())


def test_mur_end_absorbs_tem_pulse():
    steps = 1700
    ref, d = probe_series(600, steps)
    # the long line's own far-end reflection reaches the probe after ~1000 cells
    n = int(900e-3 / C0 / d.dT)
    assert n < steps
    u, _ = probe_series(200, steps)
    mur = reflection_db(u, ref, n)
    pec, _ = probe_series(200, steps, far=BC_PEC)
    bad, _ = probe_series(200, steps, flip=True)
    assert mur < -50, mur
    assert reflection_db(pec, ref, n) > -1.0
    assert reflection_db(bad, ref, n) > mur + 20


def test_coefficients_closed_form_and_material():
    op = Operator()
    lines = [np.arange(13, dtype=float), np.arange(9, dtype=float), np.arange(11, dtype=float)]
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_MUR] * 6, [0] * 6)
    op.set_gauss_excite(0, 10e9)
    m = op.add_material(eps_r=4.0)
    op.add_box(m, (0.0, 2.0, 3.0), (3.0, 5.0, 6.0), priority=1)     # touches the x-lo face
    d = op.build(0, nr_ts=10)
    assert [f.face for f in d.mur] == [0, 1, 2, 3, 4, 5]
    x_lo = d.mur[0]
    assert x_lo.k_nyP.shape == (9, 11)                               # [Ny][Nz]
    delta = 1e-3

    def k(c):
        return np.float32((c * d.dT - delta) / (c * d.dT + delta))
    assert x_lo.k_nyP[0, 0] == k(C0) and x_lo.k_nyPP[8, 10] == k(C0)
    assert abs(float(x_lo.k_nyP[3, 4]) - float(k(C0 / 2))) < 1e-7
    assert d.mur[1].k_nyP[3, 4] == k(C0)                             # x-hi face: vacuum
    assert all(f.start_ts == 0 for f in d.mur)
    # phase velocity overrides the material
    op.set_mur_phase_velocity(0, 0.5 * C0)
    d2 = op.build(0, nr_ts=10)
    assert abs(float(d2.mur[0].k_nyP[0, 0]) - float(k(0.5 * C0))) < 1e-7


def test_excitation_on_the_boundary_plane_delays_the_face():
    op = Operator()
    lines = [np.arange(9, dtype=float)] * 3
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC, BC_PEC, BC_MUR, BC_PEC, BC_PEC, BC_PEC], [0] * 6)
    op.set_gauss_excite(0, 10e9)
    e = op.add_excitation(0, (0.0, 0.0, 1.0), delay=0.0)
    op.add_box(e, (4.0, 0.0, 3.0), (4.0, 0.0, 4.0), priority=5)      # E_z on the y-lo plane
    d = op.build(0, nr_ts=100)
    assert d.mur[0].face == 2
    assert d.mur[0].start_ts == d.excitation.length + 10
