"""Writes tests/golden/cavity64.npz: the 64^3-cell, 1 mm, PEC, soft-Ez
Gaussian (f0=0, fc=10 GHz), 2000-step case run by the CPU oracle (float,
FTZ), with the operator constants and a hash of the coefficient arrays.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from wipems_amd import scenarios  # noqa: E402
from oracle.oracle import oracle_from_operator  # noqa: E402


def coeff_hash(d):
    h = hashlib.sha256()
    for a in (d.vv, d.vi, d.ii, d.iv):
        h.update(np.ascontiguousarray(a, dtype=np.float32).tobytes())
    return h.hexdigest()


def main():
    sc = scenarios.cavity_point_source(64, 2000, prec=0)
    d = sc.data
    o = oracle_from_operator(d, ftz=True)
    o.iterate(2000)
    tu, u = o.read_probe(0)
    ti, i = o.read_probe(1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cavity64.npz")
    np.savez(out, t_u=tu, u=u, t_i=ti, i=i, dT=np.float64(d.dT), length=np.int64(d.excitation.length),
             nyquist=np.int64(d.nyquist), interval=np.int64(sc.meta["interval"]), energy=np.float64(o.energy()),
             coeff_sha256=np.array(coeff_hash(d)), signal_volt=d.excitation.signal_volt)
    print("wrote", out, len(u), "samples")


if __name__ == "__main__":
    main()
