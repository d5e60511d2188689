"""Lumped RLC (Operator_Ext_LumpedRLC + Engine_Ext_LumpedRLC), pinned on the
CPU oracle by the circuits the extension builds.

With the current coefficients zeroed, an RLC edge is decoupled from the
grid and its voltage follows the element's own recursion:

* parallel LC (R = 0): leapfrog x' = x - y', y' = y + (dT^2 / LC) x, so the
  ringing frequency is exactly acos(1 - dT^2 / (2 LC)) / (2 pi dT);
* series RLC with the cell capacitance Cd across it: the continuous loop
  L J' + R J + q / C = V, Cd V' = -J rings at sqrt(1/(L Ceff) - (R/2L)^2),
  Ceff = C Cd / (C + Cd), and its oscillation decays as exp(-R t / 2L);
  at a tenth of the Courant step (~130 steps per period) the discrete
  recursion is within 0.5% (frequency) and 5% (decay rate) of that.

A planted 10% error in the inductance term moves the parallel frequency
out, and doubling R doubles the series decay rate.  The parallel element's
coefficient lists are checked against their closed forms.
"""
import math

import numpy as np

from oracle.oracle import oracle_from_operator
from wipems_amd.operator import BC_PEC, Operator


def rlc_edge(rlc_type, R, L, C, steps, ts_factor=1.0):
    op = Operator()
    lines = [np.arange(7, dtype=float)] * 3
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_timestep_factor(ts_factor)
    op.set_gauss_excite(2e9, 2e9)
    e = op.add_lumped_rlc(2, rlc_type, R=R, L=L, C=C, caps=False)
    op.add_box(e, (3.0, 3.0, 3.0), (3.0, 3.0, 4.0), priority=5)
    d = op.build(0, nr_ts=steps)
    for a in (d.ii, d.iv):
        a[...] = 0                      # decouple the edge from the grid
    d.excitation = None
    d.probes = [op.voltage_probe((3.0, 3.0, 3.0), (3.0, 3.0, 4.0), td_interval=1, name="v")]
    return op, d


def ring(d, steps, v0=1.0):
    o = oracle_from_operator(d)
    V = o.get_field(0)
    V[2, 3, 3, 3] = v0
    o.set_field(0, V)
    o.iterate(steps)
    return o.read_probe(0)[1]


def frequency(v, dT):
    """Mean frequency from linearly interpolated upward zero crossings."""
    s = np.sign(v)
    idx = np.where((s[:-1] < 0) & (s[1:] >= 0))[0]
    t = idx + v[idx] / (v[idx] - v[idx + 1])
    return (len(t) - 1) / ((t[-1] - t[0]) * dT)


def test_parallel_lc_rings_at_the_discrete_frequency():
    L, C, steps = 1e-9, 1e-12, 3000
    op, d = rlc_edge(0, 0.0, L, C, steps)
    r = d.rlc
    assert r.coef.shape == (9, 1) and int(r.dir[0]) == 2
    dT = d.dT
    # i2v = dT/C (G = 0), ilv = dT/L; series lists zero, vvd = 1
    assert abs(float(r.coef[1, 0]) - dT / C) <= 1e-6 * dT / C
    assert abs(float(r.coef[0, 0]) - dT / L) <= 1e-6 * dT / L
    assert float(r.coef[5, 0]) == 1.0 and not np.any(r.coef[2:5]) and not np.any(r.coef[6:])
    v = ring(d, steps)
    f = frequency(v, dT)
    a = float(np.float32(np.float32(dT / C) * np.float32(dT / L)))
    f_exact = math.acos(1 - a / 2) / (2 * math.pi * dT)
    assert abs(f / f_exact - 1) < 2e-4, (f, f_exact)
    assert abs(f_exact * 2 * math.pi * math.sqrt(L * C) - 1) < 1e-3
    # teeth: a 10% stronger inductor term
    d.rlc.coef[0] *= np.float32(1.1)
    assert abs(frequency(ring(d, steps), dT) / f_exact - 1) > 0.03


def test_series_rlc_matches_the_continuous_loop():
    # L C resonant at fMax = 4 GHz keeps the element's impedance small there,
    # so the cell keeps its own capacitance Cd (no raise); a tenth of the
    # Courant step gives ~130 steps per period of the L-Cd loop
    R, L, C, steps = 5.0, 1.583e-9, 1e-12, 4000
    op, d = rlc_edge(1, R, L, C, steps, ts_factor=0.1)
    Cd = float(op.get_ec()[0][2, 3, 3, 3])          # the cell capacitance the extension settled on
    assert abs(Cd / (8.854e-12 * 1e-3) - 1) < 1e-3   # eps0 * area / length of a 1 mm cell
    ceff = C * Cd / (C + Cd)
    alpha = R / (2 * L)
    wd = math.sqrt(1 / (L * ceff) - alpha ** 2)
    v = ring(d, steps)
    f = frequency(v, d.dT)
    assert abs(f / (wd / (2 * math.pi)) - 1) < 5e-3, (f, wd / (2 * math.pi))
    # decay of the oscillation about its (charge-sharing) offset: half the
    # peak-to-peak over three periods, early against late
    per = int(round(1 / (f * d.dT)))
    i1, i2 = 2 * per, len(v) - 4 * per

    def amp(i):
        w = v[i:i + 3 * per]
        return 0.5 * (w.max() - w.min())
    rate = math.log(amp(i1) / amp(i2)) / ((i2 - i1) * d.dT)
    assert abs(rate / alpha - 1) < 0.05, (rate, alpha)
    # teeth: twice the resistance decays twice as fast
    _, d2 = rlc_edge(1, 2 * R, L, C, steps, ts_factor=0.1)
    v2 = ring(d2, steps)
    w1, w2 = v2[i1:i1 + 3 * per], v2[i2:i2 + 3 * per]
    rate2 = math.log((w1.max() - w1.min()) / (w2.max() - w2.min())) / ((i2 - i1) * d.dT)
    assert abs(rate2 / (2 * alpha) - 1) < 0.05
