"""Steady-state detection (Operator_Ext_SteadyState + Engine_Ext_SteadyState).

Pinned on the CPU oracle, then the device engine against the oracle:

* Linear ramp, closed form.  With vv = 1, vi = 0 and a constant soft source of
  1 on one edge, that edge records V = n at step n.  The reference's ring
  (slot n mod 2p) and window choice (new = slots [0, p), old = [p, 2p) at every
  evaluation) then give, at n = k p:
    k even: new = {kp, (k-2)p+1 .. (k-2)p+p-1},  old = {(k-1)p .. (k-1)p+p-1}
    k odd:  new = {(k-1)p .. (k-1)p+p-1},        old = {kp, (k-2)p+1 .. (k-2)p+p-1}
  so diff_pow = p^3 for every k and curr_pow is the sum of squares of the new
  window.  An edge with vv = -1 alternates sign (ratio 4 for odd p) and counts
  only while its power is above 1% of the largest; an all-zero edge never
  counts.  A constant current edge adds fixed energy and lets the probe term
  win.  All values are small integers, so every sum is exact and the expected
  value is compared bit for bit.
* An exactly periodic field gives 0.
* A sine-driven parallel-plate line with Mur ends settles below the reference's
  default end criterion (1e-5); with the excitation's periodic wrap removed
  the sine stops after two periods and the detection never settles.
* The operator gives a sine excitation the reference's default probe set
  (openems.cpp:1060-1091, 21 probes) and the excitation wrap int(T/dT).
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, oracle_from_operator
from wipems_amd.operator import (BC_MUR, BC_PEC, BC_PMC, Excitation, Operator, SteadyState,
                                 default_steady_state_probes)

EPS0 = 8.85418781762e-12
MUE0 = 1.256637062e-6

N = (6, 5, 7)
P = 7                                  # odd: the alternating edge flips sign between windows
DT = 1e-12
EDGE_A = (2, (2, 2, 3))                # (component, node): the ramp
EDGE_B = (0, (3, 1, 4))                # alternating +-C_B
EDGE_Z = (1, (4, 3, 5))                # always zero
EDGE_I = (1, (1, 1, 1))                # constant current C_I
C_B = 3.0


def ramp_setup(target, dtype, c_i):
    """Loads the ramp case into an Oracle or an Engine (same methods)."""
    sh = (3,) + N
    vv = np.ones(sh, dtype)
    vv[EDGE_B[0]][EDGE_B[1]] = -1
    vi = np.zeros(sh, dtype)
    ii = np.ones(sh, dtype)
    iv = np.zeros(sh, dtype)
    target.set_coeffs(vv, vi, ii, iv)
    a = np.array([[EDGE_A[1][0]], [EDGE_A[1][1]], [EDGE_A[1][2]]], np.uint32)
    e = np.zeros((3, 0), np.uint32)
    target.set_excitation(Excitation(2, 0, np.ones(2), np.ones(2), a, np.array([EDGE_A[0]], np.uint32), np.ones(1),
                                     np.zeros(1, np.uint32), e, np.zeros(0, np.uint32), np.zeros(0), np.zeros(0, np.uint32)))
    V = np.zeros(sh)
    V[EDGE_B[0]][EDGE_B[1]] = C_B
    I = np.zeros(sh)
    I[EDGE_I[0]][EDGE_I[1]] = c_i
    target.set_field(0, V)
    target.set_field(1, I)
    pos = np.array([EDGE_A[1], EDGE_B[1], EDGE_Z[1]], np.uint32).T.copy()
    d = np.array([EDGE_A[0], EDGE_B[0], EDGE_Z[0]], np.uint32)
    target.set_steady_state(SteadyState(P * DT, pos, d))


def ramp_expected(kmax, c_i):
    """GetLastDiff after each evaluation k = 2..kmax, from the window contents
    derived in the module docstring."""
    p = P
    out, last_e = {}, 0.0
    for k in range(2, kmax + 1):
        if k % 2 == 0:
            new = [k * p] + [(k - 2) * p + j for j in range(1, p)]
        else:
            new = [(k - 1) * p + j for j in range(p)]
        curr_a = float(sum(v * v for v in new))
        diff_a = float(p ** 3)
        curr_b, diff_b = p * C_B * C_B, p * (2 * C_B) ** 2
        e = EPS0 * (float((k * p) ** 2) + C_B * C_B) + MUE0 * (c_i * c_i)
        diff, valid = 0.0, False
        if last_e > 0:
            diff, valid = abs(e - last_e) / last_e, True
        last_e = e
        mx = max(curr_a, curr_b, 0.0)
        for c, d in ((curr_a, diff_a), (curr_b, diff_b), (0.0, 0.0)):
            if c > mx * 1e-2:
                diff, valid = max(diff, d / c), True
        if not valid or diff > 1:
            diff = 1.0
        out[k] = diff
    return out


def run_ramp(target, kmax, step_fn):
    """GetLastDiff after every step up to k = kmax evaluations."""
    got = {}
    for n in range(kmax * P + 1):
        step_fn(1)
        if n >= 2 * P and n % P == 0:
            got[n // P] = target.GetLastDiff()
        elif n < 2 * P:
            assert target.GetLastDiff() == 1.0
    return got


@pytest.mark.parametrize("c_i", [0.0, 1000.0])
@pytest.mark.parametrize("dbl", [False, True])
def test_ramp_closed_form_oracle(dbl, c_i):
    o = Oracle(N, DT, use_double=dbl)
    ramp_setup(o, np.float64 if dbl else np.float32, c_i)
    kmax = 14
    exp = ramp_expected(kmax, c_i)
    got = run_ramp(o, kmax, o.iterate)
    assert got == exp
    # the case exercises each branch: capped at 1, the alternating edge
    # counted then dropped by the 1% rule, energy or probe term winning
    assert exp[2] == exp[3] == exp[4] == 1.0
    # k = 4: the alternating edge (power 63) is above 1% of the ramp's (2639);
    # k = 5: below 1% of 6727, so it no longer counts
    assert exp[5] < 0.6 and exp[kmax] < 0.2
    if c_i == 0:
        assert exp[kmax] == pytest.approx((kmax ** 2 - (kmax - 1) ** 2) * P * P / ((kmax - 1) ** 2 * P * P + C_B ** 2), rel=1e-12)
    else:
        assert exp[kmax] == P ** 3 / sum(v * v for v in [kmax * P] + [(kmax - 2) * P + j for j in range(1, P)])


def test_ramp_window_rule_is_discriminating():
    """The reference's windows differ from the 'last two full periods' reading;
    the closed form distinguishes them at every even k."""
    p = P
    for k in range(2, 10, 2):
        ref = sum(v * v for v in [k * p] + [(k - 2) * p + j for j in range(1, p)])
        naive = sum(((k - 1) * p + j) ** 2 for j in range(p))
        assert ref != naive


def test_exactly_periodic_field_gives_zero():
    o = Oracle(N, DT, use_double=False)
    ramp_setup(o, np.float32, 0.0)
    # integer signal with zero sum over its period P: the source edge repeats
    # exactly every P steps; the alternating edge repeats every 2 steps, so use
    # an even detection period
    sig = np.array([1.0, 2.0, -1.0, -2.0, 3.0, -3.0, 0.0])
    a = np.array([[EDGE_A[1][0]], [EDGE_A[1][1]], [EDGE_A[1][2]]], np.uint32)
    e = np.zeros((3, 0), np.uint32)
    o.set_excitation(Excitation(P, P, sig, sig, a, np.array([EDGE_A[0]], np.uint32), np.ones(1),
                                np.zeros(1, np.uint32), e, np.zeros(0, np.uint32), np.zeros(0), np.zeros(0, np.uint32)))
    pos = np.array([EDGE_A[1], EDGE_B[1]], np.uint32).T.copy()
    o.set_steady_state(SteadyState(2 * P * DT, pos, np.array([EDGE_A[0], EDGE_B[0]], np.uint32)))
    seen = []
    for n in range(10 * 2 * P + 1):
        o.iterate(1)
        if n >= 4 * P and n % (2 * P) == 0:
            seen.append(o.GetLastDiff())
    assert seen and all(v == 0.0 for v in seen)


def test_default_probe_set():
    pos, d = default_steady_state_probes((10, 12, 14))
    assert pos.shape == (3, 21) and d.shape == (21,)
    cols = [tuple(int(v) for v in pos[:, i]) for i in range(21)]
    assert cols[:3] == [(5, 6, 7)] * 3
    assert cols[3:9] == [(0, 6, 7)] * 6
    assert cols[9:15] == [(5, 0, 7)] * 6
    assert cols[15:21] == [(5, 6, 0)] * 6
    assert list(d) == [0, 1, 2] * 7


def line_op(f0, nx=200, n_t=4):
    op = Operator()
    lines = [np.arange(nx + 1, dtype=float), np.arange(n_t + 1, dtype=float), np.arange(n_t + 1, dtype=float)]
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_MUR, BC_MUR, BC_PMC, BC_PMC, BC_PEC, BC_PEC], [0] * 6)
    op.set_excite_signal(1, f0)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (50.0, 0.0, 0.0), (50.0, float(n_t), float(n_t)), priority=5)
    return op


def test_operator_sine_period_and_default_detection():
    f0 = 5e9
    d = line_op(f0).build(0)
    T = 1 / f0
    assert d.signal_period == T
    ts = round(T / d.dT)
    assert abs(ts * d.dT - T) < 1e-9 * T                      # dT matched to the period
    assert d.excitation.period == int(T / d.dT)
    assert d.excitation.length == round(2 / f0 / d.dT)
    assert d.steady_state is not None and d.steady_state.period == T
    assert d.steady_state.pos.shape == (3, 21)
    g = line_op(f0)
    g.set_gauss_excite(f0, f0)
    dg = g.build(0, nr_ts=100)
    assert dg.signal_period == 0 and dg.steady_state is None and dg.excitation.period == 0


def _settle(o, periods, p):
    diffs = []
    for _ in range(periods):
        o.iterate(p)
        diffs.append(o.GetLastDiff())
    return diffs


def test_sine_line_settles():
    d = line_op(5e9).build(0)
    o = oracle_from_operator(d)
    p = o.set_steady_state(d.steady_state)
    assert p == round(d.signal_period / d.dT)
    diffs = _settle(o, 40, p)
    assert diffs[0] == 1.0 and diffs[1] == 1.0
    assert min(diffs[-5:]) < 1e-5, diffs[-5:]
    assert max(diffs[-5:]) < 1e-4, diffs[-5:]


def test_sine_without_wrap_does_not_settle():
    d = line_op(5e9).build(0)
    d.excitation.period = 0                  # the sine then stops after its two stored periods
    o = oracle_from_operator(d)
    p = o.set_steady_state(d.steady_state)
    diffs = _settle(o, 40, p)
    assert min(diffs[5:]) > 0.1, diffs


# ---------------------------------------------------------------------------
# device engine
# ---------------------------------------------------------------------------
def _engine(precision):
    from wipems_amd.engine import Engine
    return Engine(N, DT, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("c_i", [0.0, 1000.0])
@pytest.mark.parametrize("precision,dtype", [("fp32", np.float32), ("fp64", np.float64), ("fp16", np.float32)])
def test_gpu_ramp_closed_form(gpu, precision, dtype, c_i):
    """Integer-valued records and energies: the device result equals the
    closed form exactly (FP16 stores these small integers exactly)."""
    print("case: steady-state ramp, %s, c_i %g" % (precision, c_i), flush=True)
    eng = _engine(precision)
    ramp_setup(eng, dtype, c_i)
    assert eng.steady_state_period() == P
    kmax = 14
    got = run_ramp(eng, kmax, eng.iterate)
    assert got == ramp_expected(kmax, c_i)


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [True, False])
def test_gpu_ramp_chunked_matches_oracle(gpu, use_graph):
    """Chunks that straddle evaluation steps (graph replays of 16 steps and
    single steps around them) give the oracle's value at every chunk end."""
    print("case: steady-state ramp chunked, graph %s" % use_graph, flush=True)
    eng = _engine("fp32")
    eng.set_use_graph(use_graph)
    ramp_setup(eng, np.float32, 1000.0)
    o = Oracle(N, DT, use_double=False)
    ramp_setup(o, np.float32, 1000.0)
    for c in (3, 11, 40, 1, 17, 33, 5):
        eng.iterate(c)
        o.iterate(c)
        assert eng.GetLastDiff() == o.GetLastDiff()
    assert eng.GetNumberOfTimesteps() == 110


@pytest.mark.gpu
@pytest.mark.parametrize("prec,precision", [(0, "fp32"), (1, "fp64")])
def test_gpu_sine_line_matches_oracle(gpu, prec, precision):
    """Operator case with the default detection: fields bitwise, the
    detection value at every chunk end within the energy sums' rounding.
    The device sums the energy in FP64 in another order; a relative error e
    of each energy moves |E1 - E0| / E0 by up to about 2e in absolute terms
    (e <= n * 2^-53 ~ 1e-12 for these 10^4 terms), whatever the ratio's size."""
    from wipems_amd.engine import engine_from_operator
    print("case: steady-state sine line, %s" % precision, flush=True)
    d = line_op(5e9).build(prec)
    eng = engine_from_operator(d, precision=precision)
    o = oracle_from_operator(d)
    assert eng.steady_state_period() == 104
    vals = []
    for _ in range(80):
        eng.iterate(37)
        o.iterate(37)
        g, r = eng.GetLastDiff(), o.GetLastDiff()
        assert g == pytest.approx(r, rel=1e-9, abs=2e-12), (eng.GetNumberOfTimesteps(), g, r)
        vals.append(g)
    for which in (0, 1):
        np.testing.assert_array_equal(eng.get_field(which), o.get_field(which))
    assert vals[-1] < 1e-4


@pytest.mark.gpu
def test_gpu_steady_state_rejections(gpu):
    from wipems_amd._native import FdtdError
    from wipems_amd.engine import Batch
    eng = _engine("fp32")
    pos = np.zeros((3, 1), np.uint32)
    with pytest.raises(FdtdError):
        eng.set_steady_state(SteadyState(0.2 * DT, pos, np.zeros(1, np.uint32)))     # below one timestep
    with pytest.raises(FdtdError):
        eng.set_steady_state(SteadyState(P * DT, np.full((3, 1), 99, np.uint32), np.zeros(1, np.uint32)))
    with pytest.raises(FdtdError):
        eng.set_steady_state(SteadyState(P * DT, pos, np.array([3], np.uint32)))
    with pytest.raises(FdtdError):
        eng.GetLastDiff()
    ramp_setup(eng, np.float32, 0.0)
    with pytest.raises(FdtdError):
        Batch([eng])
    eng.iterate(1)
    with pytest.raises(FdtdError):
        eng.set_steady_state(None)            # only before the first timestep
