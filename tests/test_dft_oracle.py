"""Pins of the field-box DFT restatement (oracle/dft.py) on the CPU:
node interpolation against closed forms of linear fields, the boundary
rules of NODE_INTERPOLATE, and the accumulation against a geometric-series
closed form.  Runs without a GPU."""
import math

import numpy as np
import pytest

from oracle.dft import DftBox, interp_e, interp_h
from wipems_amd.operator import BC_PEC, Operator, fd_interval_for


def graded(rng, n, lo=0.5, hi=2.0):
    return np.concatenate([[0.0], np.cumsum(rng.uniform(lo, hi, n - 1))])


def build_small(lines, unit=1e-3):
    op = Operator()
    op.set_mesh(lines, unit)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_gauss_excite(0.0, 10e9)
    return op, op.build(0, nr_ts=10)


def split(e, N):
    return [e[:N[0]], e[N[0]:N[0] + N[1]], e[N[0] + N[1]:]]


def test_edge_lengths_closed_form():
    rng = np.random.default_rng(5)
    lines = [graded(rng, n) for n in (7, 9, 6)]
    op, data = build_small(lines)
    N = data.num_lines
    P, D = split(data.edge_primary, N), split(data.edge_dual, N)
    for n in range(3):
        l = np.asarray(lines[n]) * 1e-3
        prim = np.append(np.diff(l), l[-1] - l[-2])
        dl = np.append(0.5 * (l[:-1] + l[1:]), l[-1] + 0.5 * (l[-1] - l[-2]))
        dual = np.concatenate([[l[1] - l[0]], np.diff(dl)])
        np.testing.assert_allclose(P[n], prim, rtol=1e-12)
        np.testing.assert_allclose(D[n], dual, rtol=1e-12)


def test_e_node_interpolation_of_linear_field_is_exact():
    """raw E at edge midpoints of a linear field, blended with delta/(delta+deltaDown),
    gives the field at the node exactly on a graded mesh."""
    rng = np.random.default_rng(7)
    N = (9, 8, 10)
    lines = [graded(rng, n) * 1e-3 for n in N]
    lens = [np.append(np.diff(l), l[-1] - l[-2]) for l in lines]
    a, b = 0.7, [3.0e2, -1.5e2, 2.2e2]
    V = np.zeros((3,) + N)
    for n in range(3):
        mid = lines[n] + 0.5 * lens[n]
        f = a + b[n] * mid
        sh = [1, 1, 1]; sh[n] = N[n]
        V[n] = np.broadcast_to((f * lens[n]).reshape(sh), N)
    start, stop = (0, 0, 0), tuple(k - 1 for k in N)
    E = interp_e(V, lens, start, stop)
    for n in range(3):
        sh = [1, 1, 1]; sh[n] = N[n]
        exact = np.broadcast_to((a + b[n] * lines[n]).reshape(sh), N)
        inner = [slice(None)] * 3
        inner[n] = slice(1, N[n] - 1)
        np.testing.assert_allclose(E[n][tuple(inner)], exact[tuple(inner)], rtol=1e-12, atol=1e-12)
        # boundary rules: node 0 takes the value of edge 0, the last node that of the last edge
        mid = lines[n] + 0.5 * lens[n]
        first = [slice(None)] * 3; first[n] = 0
        last = [slice(None)] * 3; last[n] = N[n] - 1
        np.testing.assert_allclose(E[n][tuple(first)], a + b[n] * mid[0], rtol=1e-12)
        np.testing.assert_allclose(E[n][tuple(last)], a + b[n] * mid[N[n] - 2], rtol=1e-12)
    # teeth: exchanging the two weights breaks exactness on this graded mesh
    n = 0
    i = 4
    r0 = (V[0][i] / lens[0][i])[0, 0]
    r1 = (V[0][i - 1] / lens[0][i - 1])[0, 0]
    rel = lens[0][i] / (lens[0][i] + lens[0][i - 1])
    swapped = r0 * rel + r1 * (1.0 - rel)
    assert abs(swapped - (a + b[0] * lines[0][i])) > 1e-6
    assert abs(E[0][i, 0, 0] - (a + b[0] * lines[0][i])) < 1e-12


def test_e_zero_length_edge_gives_zero():
    N = (4, 4, 4)
    lens = [np.array([1e-3, 0.0, 1e-3, 1e-3])] * 3
    V = np.ones((3,) + N)
    E = interp_e(V, lens, (0, 0, 0), (3, 3, 3))
    assert np.all(E[0][1] == 0.0)
    assert np.all(E[0][2] == 0.0)       # deltaDown = 0: rel = 1, so the lower raw value (0) is taken
    assert np.all(E[0][3] == 1e3)


def test_h_node_interpolation_uniform_linear_and_bounds():
    N = (7, 6, 8)
    h = 1e-3
    lens = [np.full(n, h) for n in N]
    x = [np.arange(n) * h for n in N]
    I = np.zeros((3,) + N)
    c = [0.3, -0.8, 1.1]
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        # H_n lives on dual edges at cell centres in the transverse axes
        g = np.zeros(N)
        P = np.meshgrid(*[x[k] + 0.5 * h for k in range(3)], indexing="ij")
        g = 2.0 + c[nP] * P[nP] * 1e3 + c[nPP] * P[nPP] * 1e3
        I[n] = g * h
    H = interp_h(I, lens, (0, 0, 0), tuple(k - 1 for k in N))
    X = np.meshgrid(*x, indexing="ij")
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        exact = 2.0 + c[nP] * X[nP] * 1e3 + c[nPP] * X[nPP] * 1e3
        ok = np.ones(N, bool)
        for k in range(3):
            ok &= X[k] < x[k][-1] - 0.5 * h
        ok &= (X[nP] > 0) & (X[nPP] > 0)
        np.testing.assert_allclose(H[n][ok], exact[ok], rtol=1e-12, atol=1e-12)
        assert np.all(H[n][~ok] == 0.0)


def _series(w0, w, tau, shift_t, K):
    """sum_{k=0..K} cos(w0 t_k) exp(-i w t_k), t_k = k tau + shift_t, in closed form."""
    total = 0
    for s in (1, -1):
        a = s * w0 - w
        q = np.exp(1j * a * tau)
        p0 = np.exp(1j * a * shift_t)
        geo = (K + 1) if abs(q - 1) < 1e-15 else (1 - q ** (K + 1)) / (1 - q)
        total += 0.5 * p0 * geo
    return total


def test_accumulation_matches_geometric_series():
    """A spatially constant field cos(w0 t) sampled every `interval` steps:
    sum_k cos(w0 t_k) exp(-i w t_k) 2 dT interval in closed form (E at
    t = ts dT, H at the dual time (ts + 1/2) dT)."""
    N = (3, 3, 3)
    dT = 1.3e-12
    interval = 7
    f0, freqs = 2.1e9, [1.0e9, 2.1e9, 3.3e9]
    K = 40
    for field in (0, 1):
        e = np.ones(9)
        box = DftBox(field, (1, 1, 1), (1, 1, 1), freqs, interval, e, e, N)
        shift = 0.5 if field else 0.0
        for k in range(0, K + 1):
            t = (k * interval + shift) * dT
            F = np.full((3,) + N, math.cos(2 * math.pi * f0 * t))
            box.accumulate(F, k * interval, dT)
        got = box.acc[:, 0, 0, 0, 0]
        for j, f in enumerate(freqs):
            w, w0, tau = 2 * math.pi * f, 2 * math.pi * f0, interval * dT
            ref = _series(w0, w, tau, shift * dT, K) * 2 * dT * interval
            assert abs(got[j] - ref) <= 1e-12 * abs(ref), (field, f, got[j], ref)
            # teeth: the opposite phase sign, or no dual-time shift for H, give other sums
            wrong = _series(w0, -w, tau, shift * dT, K) * 2 * dT * interval
            assert abs(got[j] - wrong) > 1e-3 * abs(ref)
            if field:
                unshifted = _series(w0, w, tau, 0.0, K) * 2 * dT * interval
                assert abs(got[j] - unshifted) > 1e-6 * abs(ref)
        assert box.samples == K + 1


def test_fd_interval_rule():
    # Processing::AddFrequency: min(Nyquist steps of the excitation, floor(1/f/2/dT))
    assert fd_interval_for([1e9], 1e-12, 26) == 26
    assert fd_interval_for([1e9, 30e9], 1e-12, 26) == 16
    assert fd_interval_for([1e9, 0.9e12], 1e-12, 26) == 26     # above the time-step limit: skipped
    with pytest.raises(ValueError):
        fd_interval_for([0.9e12], 1e-12, 26)
