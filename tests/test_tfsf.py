"""Plane-wave (TFSF) builder and hook, pinned on the CPU oracle.

* In vacuum the scattered-field region outside the box stays empty: the
  reference's own check for this extension (total-field/scattered-field
  separation).  A planted sign error in the current injection list breaks it.
* Inside the box the field is the incident plane wave: the centre edge
  voltage equals E0 * edge * g(t - d/c - dT) for the Gaussian g of
  FDTD/excitation.cpp:151-177, d the distance from the reference plane one
  line before the box (operator_ext_tfsf.cpp:185-195).
"""
import math

import numpy as np
import pytest

from oracle.oracle import oracle_from_operator
from wipems_amd import scenarios

C0 = 299792458.0


def gauss(t, f0, fc):
    return np.cos(2 * np.pi * f0 * (t - 9 / (2 * np.pi * fc))) * np.exp(-(2 * np.pi * fc * t / 3 - 3) ** 2)


def leakage_db(o, steps):
    o.iterate(steps)
    _, u = o.read_probe(0)
    _, us = o.read_probe(1)
    return 20 * math.log10(np.abs(us).max() / np.abs(u).max()), u


def test_axis_plane_wave_separates_and_matches_closed_form():
    sc = scenarios.plane_wave_box(steps=500)
    d = sc.data
    assert d.tfsf is not None and d.tfsf[0]["comp"].size == d.tfsf[1]["comp"].size > 0
    o = oracle_from_operator(d)
    db, u = leakage_db(o, 500)
    assert db < -80, db
    t, _ = o.read_probe(0)
    lines = sc.meta["lines"]
    x0 = lines[0][8 + 6 - 1]
    dist = (sc.meta["centre"][0] - x0) * 1e-3
    exp = 1e-3 * gauss(t - dist / C0 - d.dT, 5e9, 5e9)
    err = np.linalg.norm(u - exp) / np.linalg.norm(exp)
    assert err < 2e-3, err
    # teeth: a quarter-step timing error is already visible
    exp_late = 1e-3 * gauss(t - dist / C0 - 1.25 * d.dT, 5e9, 5e9)
    assert np.linalg.norm(u - exp_late) / np.linalg.norm(exp_late) > 4 * err


def test_planted_sign_error_leaks():
    sc = scenarios.plane_wave_box(steps=500)
    d = sc.data
    d.tfsf[1]["amp"] = -d.tfsf[1]["amp"]
    db, _ = leakage_db(oracle_from_operator(d), 500)
    assert db > -20, db


@pytest.mark.parametrize("prop_dir,e_vec", [((1.0, 1.0, 0.0), (0.0, 0.0, 1.0)),
                                            ((0.0, -1.0, 0.0), (1.0, 0.0, 0.0)),
                                            ((1.0, 1.0, 1.0), (1.0, -1.0, 0.0))])
def test_oblique_and_reversed_directions_separate(prop_dir, e_vec):
    sc = scenarios.plane_wave_box(steps=450, prop_dir=prop_dir, e_vec=e_vec)
    db, _ = leakage_db(oracle_from_operator(sc.data), 450)
    # staircased oblique incidence leaks more than the axis case but stays small
    assert db < -30, (prop_dir, db)


def test_graded_mesh_and_numerical_phase_velocity():
    sc0 = scenarios.plane_wave_box(steps=400, graded=True)
    sc1 = scenarios.plane_wave_box(steps=400, graded=True, phase_freq=5e9)
    d0, d1 = sc0.data, sc1.data
    # a numerical phase velocity below c0 lengthens the delays
    assert np.all(d1.tfsf[1]["delay"] + d1.tfsf[1]["delta"] >= d0.tfsf[1]["delay"] + d0.tfsf[1]["delta"])
    for d in (d0, d1):
        db, _ = leakage_db(oracle_from_operator(d), 400)
        assert db < -30, db


def test_polarization_is_made_orthogonal():
    """E not perpendicular to k: the builder removes the parallel part, so a
    slanted E vector gives the same tables as its perpendicular projection."""
    a = scenarios.plane_wave_box(steps=10, e_vec=(0.5, 0.0, 1.0)).data
    b = scenarios.plane_wave_box(steps=10, e_vec=(0.0, 0.0, 1.0)).data
    for w in (0, 1):
        np.testing.assert_array_equal(a.tfsf[w]["amp"], b.tfsf[w]["amp"])
        np.testing.assert_array_equal(a.tfsf[w]["delay"], b.tfsf[w]["delay"])
