"""The end-criterion time loop (engine.run_fdtd, openEMS::RunFDTD), run over
the oracle on the CPU and once over the device engine."""
import math

import numpy as np
import pytest

from oracle.oracle import oracle_from_operator
from wipems_amd import scenarios
from wipems_amd.engine import run_fdtd
from wipems_amd.operator import BC_MUR, BC_PEC, BC_PML


class OracleRun:
    """The loop's view of an Oracle."""

    def __init__(self, o, period=0):
        self.o, self.period = o, period
        self.energies = []

    def IterateTS(self, n):
        self.o.iterate(n)

    def GetNumberOfTimesteps(self):
        return self.o.num_timesteps()

    def energy(self):
        e = self.o.energy()
        self.energies.append((self.o.num_timesteps(), e))
        return e

    def GetLastDiff(self):
        return self.o.GetLastDiff()

    def steady_state_period(self):
        return self.period


def _pulse(bc):
    sc = scenarios.cavity_point_source(n_cells=24, steps=4000, bc=bc, f0=8e9, fc=4e9)
    return sc.data


def test_energy_criterion_stops_an_open_domain():
    d = _pulse(BC_PML)
    r = OracleRun(oracle_from_operator(d))
    n, change = run_fdtd(r, 4000, d.max_excite_step, end_crit=1e-4, check_steps=100)
    assert n < 4000 and change <= 1e-4
    assert n % 100 == 0
    # the step of maximum excitation was sampled, and change is the last
    # sample over the largest one
    assert d.max_excite_step in [t for t, _ in r.energies]
    top = max(e for _, e in r.energies)
    assert change == r.energies[-1][1] / top
    # one check earlier the criterion was not met
    assert r.energies[-2][1] / max(e for _, e in r.energies[:-1]) > 1e-4


def test_energy_criterion_never_met_in_a_closed_cavity():
    d = _pulse(BC_PEC)
    r = OracleRun(oracle_from_operator(d))
    n, change = run_fdtd(r, 1500, d.max_excite_step, end_crit=1e-4, check_steps=250)
    assert n == 1500 and change > 0.1


def test_steady_state_criterion():
    from tests.test_steady_state import line_op
    d = line_op(5e9).build(0)
    o = oracle_from_operator(d)
    p = o.set_steady_state(d.steady_state)
    r = OracleRun(o, p)
    n, change = run_fdtd(r, 60 * p, d.max_excite_step, end_crit=1e-4, check_steps=p)
    assert change <= 1e-4 and n < 60 * p and n % p == 0
    assert not r.energies                       # the energy is not consulted


@pytest.mark.gpu
def test_gpu_run_loop_matches_oracle(gpu):
    from wipems_amd.engine import engine_from_operator
    print("case: end-criterion loop on the device engine", flush=True)
    d = _pulse(BC_PML)
    eng = engine_from_operator(d, precision="fp32")
    n, change = run_fdtd(eng, 4000, d.max_excite_step, end_crit=1e-4, check_steps=100)
    r = OracleRun(oracle_from_operator(d))
    n2, change2 = run_fdtd(r, 4000, d.max_excite_step, end_crit=1e-4, check_steps=100)
    assert n == n2
    assert change == pytest.approx(change2, rel=1e-9)
