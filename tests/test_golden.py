"""The committed fixture tests/golden/cavity64.npz (written by
tests/golden/make_golden.py) against the current operator and oracle."""
import os

import numpy as np

from wipems_amd import scenarios
from oracle.oracle import oracle_from_operator

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return np.load(os.path.join(HERE, "golden", "cavity64.npz"), allow_pickle=False)


def test_operator_constants_match_fixture():
    g = load()
    sc = scenarios.cavity_point_source(64, 2000, prec=0)
    d = sc.data
    assert d.dT == float(g["dT"])
    assert d.excitation.length == int(g["length"])
    assert d.nyquist == int(g["nyquist"])
    np.testing.assert_array_equal(d.excitation.signal_volt, g["signal_volt"])
    from tests.golden.make_golden import coeff_hash
    assert coeff_hash(d) == str(g["coeff_sha256"])


def test_oracle_series_match_fixture():
    g = load()
    sc = scenarios.cavity_point_source(64, 2000, prec=0)
    o = oracle_from_operator(sc.data, ftz=True)
    o.iterate(2000)
    tu, u = o.read_probe(0)
    ti, i = o.read_probe(1)
    np.testing.assert_array_equal(tu, g["t_u"])
    np.testing.assert_array_equal(u, g["u"])
    np.testing.assert_array_equal(i, g["i"])
    assert o.energy() == float(g["energy"])
