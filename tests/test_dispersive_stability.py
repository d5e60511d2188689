"""Stability of the dispersive workload, checked on the CPU oracle.

The explicit ADE of Engine_Ext_LorentzMaterial couples J and V once per
step.  For an undamped Drude medium the local update is stable only for
v_ext = (w_p dT)^2 <= 4, and on a uniform mesh at timestep factor 1 the
main update already sits at the Courant limit, so any plasma term pushes
the scheme past it.  The Drude slab / Lorentz spheres workload with its
stated parameters (w_p = 2e15 rad/s on a 0.25 mm mesh: w_p dT ~ 960) diverges
within a few steps in this restatement of the reference scheme; the
benchmarked variant scales the material frequencies by 1e-3 (relaxation
times by 1e3) and uses timestep factor 0.9, and stays bounded.
"""
import math

import numpy as np

from oracle.oracle import oracle_from_operator
from wipems_amd import scenarios

SMALL = dict(num_lines=(65, 65, 49), n_spheres=6)


def energies(data, chunks, chunk):
    o = oracle_from_operator(data)
    out = []
    for _ in range(chunks):
        o.iterate(chunk)
        out.append(o.energy())
    return np.array(out)


def test_drude_coupling_closed_form():
    sc = scenarios.dispersive_scenario(steps=10, lorentz=False, **SMALL)
    d = sc.data
    a = d.ade[0]
    wp, tau = 2e15, 1e-13
    # slab edges away from material interfaces: v_ext = (w_p dT)^2 / (1 + dT/(2 tau)) * (vi * C / dT) = ... / 1
    expect = (wp * d.dT) ** 2 / (1 + d.dT / (2 * tau))
    v = np.asarray(a.lists[1][0], dtype=np.float64)
    assert abs(np.median(v[v > 0]) / expect - 1) < 1e-5
    # far beyond the local stability bound 2 (1 + v_int)
    vint = np.asarray(a.lists[0][0], dtype=np.float64)
    assert np.median(v[v > 0]) > 1e3 * 2 * (1 + np.median(vint))


def test_stated_parameters_diverge():
    sc = scenarios.dispersive_scenario(steps=200, **SMALL)
    e = energies(sc.data, 4, 5)
    assert not np.all(np.isfinite(e)) or e[-1] > 1e20 * max(e[0], 1e-30)


def test_benchmarked_variant_stays_bounded():
    sc = scenarios.dispersive_scenario(steps=4000, material_scale=1e-3, ts_factor=0.9, **SMALL)
    e = energies(sc.data, 20, 200)
    assert np.all(np.isfinite(e))
    peak = e.max()
    assert e[-1] < peak and e[-1] < 0.5 * peak      # decays after the pulse has passed
    # the same materials at timestep factor 1 are not stable
    sc1 = scenarios.dispersive_scenario(steps=4000, material_scale=1e-3, ts_factor=1.0, **SMALL)
    e1 = energies(sc1.data, 10, 200)
    assert not np.all(np.isfinite(e1)) or e1[-1] > 1e6 * e1[0]
