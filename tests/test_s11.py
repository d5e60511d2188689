"""S11 post-processing (python/openEMS/ports.py:103-141 with the DFT of
python/openEMS/utilities.py:21-35), as scenarios.s11_from_probes restates it.

Pinned on the CPU, without the GPU:
* the DFT 2 dt sum v(t) exp(-2 pi i f t) against the closed-form geometric
  series of a sampled damped complex exponential;
* S11 = (U - U_inc)/U_inc with U_inc = (U + I Z)/2 against the reflection
  coefficient of ideal terminations: matched (0), open (+1), short (-1) and
  a resistor R ((R - Z)/(R + Z)), on arbitrary test spectra;
* an FDTD lumped port (the oracle) at the middle of a parallel-plate line of
  two 100 Ohm halves (PMC sides, PEC plates, Mur ends) sees 50 Ohm: |S11|
  stays below 0.06 across 1-8 GHz (0.045 measured), while the same waves
  against Z_ref = 100 Ohm, or the port on a line of half the impedance,
  reflect as the closed form (Z - Z_ref)/(Z + Z_ref) says."""
import math

import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.operator import BC_MUR, BC_PEC, BC_PMC, Operator

Z0 = 376.730313668


def test_dft_closed_form():
    """dft_time2freq of a sampled damped cosine equals 2 dt times the two
    geometric series sum_n r^n = (1 - r^N)/(1 - r) it splits into."""
    dt = 1.3e-12
    n = np.arange(900)
    t = n * dt
    a, f1 = 2.1e9, 3.7e9
    v = np.exp(-a * t) * np.cos(2 * np.pi * f1 * t)
    freqs = np.array([0.0, 1e9, 3.7e9, 8e9])
    ref = []
    for f in freqs:
        tot = 0j
        for sgn in (1, -1):
            r = np.exp((-a + sgn * 2j * np.pi * f1 - 2j * np.pi * f) * dt)
            tot += 0.5 * (1 - r ** len(n)) / (1 - r)
        ref.append(2 * dt * tot)
    np.testing.assert_allclose(scenarios.dft_time2freq(t, v, freqs), np.array(ref), rtol=1e-9)
    # the opposite phase sign or a missing factor 2 would not match
    wrong = 2 * dt * np.array([np.sum(v * np.exp(2j * np.pi * f * t)) for f in freqs])
    assert np.max(np.abs(wrong - np.array(ref))) > 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("R", [10.0, 50.0, 75.0, 1e3])
def test_s11_of_ideal_terminations(R):
    rng = np.random.default_rng(3)
    dt = 1e-12
    t = np.arange(400) * dt
    u = rng.standard_normal(400)
    freqs = np.linspace(1e9, 20e9, 11)
    z = 50.0
    s = scenarios.s11_from_probes(t, u, t, u / R, freqs, Z_ref=z)
    np.testing.assert_allclose(s, (R - z) / (R + z), rtol=1e-10, atol=1e-12)
    s_short = scenarios.s11_from_probes(t, 0 * u, t, u, freqs, Z_ref=z)
    np.testing.assert_allclose(s_short, -1.0, rtol=1e-12)


def _port_on_line(W=30, h=8, L=240, steps=2600):
    """Parallel-plate line along y (PMC at x = 0 and W, PEC at z = 0 and h,
    Mur at both y ends), TEM impedance Z0 h / W on each side of a lumped
    port that fills the cross-section at the middle."""
    from oracle.oracle import oracle_from_operator
    op = Operator()
    lines = [np.arange(W + 1, dtype=float), np.arange(L + 1, dtype=float), np.arange(h + 1, dtype=float)]
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PMC, BC_PMC, BC_MUR, BC_MUR, BC_PEC, BC_PEC], [0] * 6)
    op.set_gauss_excite(4e9, 4e9)
    y0 = float(L // 2)
    start, stop = (0.0, y0, 0.0), (float(W), y0, float(h))
    le = op.add_lumped_element(2, R=50.0, caps=True)
    op.add_box(le, start, stop, priority=100)
    exc = op.add_excitation(0, (0.0, 0.0, -1.0))
    op.add_box(exc, start, stop, priority=100)
    d = op.build(0, nr_ts=steps)
    d.probes = [op.voltage_probe((W / 2, y0, 0.0), (W / 2, y0, float(h)), -1.0, td_interval=1, name="u"),
                op.current_probe((0.0, y0, h / 2), (float(W), y0, h / 2), 2, 1.0, td_interval=1, name="i")]
    o = oracle_from_operator(d, ftz=True)
    o.iterate(steps)
    tu, u = o.read_probe(0)
    ti, i = o.read_probe(1)
    return tu, u, ti, i


def test_matched_port_on_parallel_plate_line():
    tu, u, ti, i = _port_on_line()
    freqs = np.linspace(1e9, 8e9, 15)
    z_line = Z0 * 8 / 30 / 2          # two 100.5 Ohm halves in parallel
    s = scenarios.s11_from_probes(tu, u, ti, i, freqs, Z_ref=50.0)
    print("matched: max|S11| = %.4f (line %.2f Ohm)" % (np.max(np.abs(s)), z_line))
    assert np.max(np.abs(s)) < 0.06
    # the same waves against a 100 Ohm reference: (50 - 100)/(50 + 100) = -1/3
    s100 = scenarios.s11_from_probes(tu, u, ti, i, freqs, Z_ref=100.0)
    np.testing.assert_allclose(s100.real, (z_line - 100.0) / (z_line + 100.0), atol=0.05)
    assert np.max(np.abs(s100.imag)) < 0.05


def test_port_on_wider_line_sees_its_impedance():
    """The same port across a line twice as wide (two 50 Ohm halves, 25 Ohm
    seen) reflects (25 - 50)/(25 + 50) = -1/3 against 50 Ohm."""
    tu, u, ti, i = _port_on_line(W=60)
    freqs = np.linspace(1e9, 8e9, 15)
    z_line = Z0 * 8 / 60 / 2
    s = scenarios.s11_from_probes(tu, u, ti, i, freqs, Z_ref=50.0)
    print("wide line: S11 %s" % np.round(s.real, 3))
    np.testing.assert_allclose(s.real, (z_line - 50.0) / (z_line + 50.0), atol=0.05)
