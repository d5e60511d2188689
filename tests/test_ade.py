"""Lorentz/Drude ADE (Operator_Ext_LorentzMaterial + Engine_Ext_LorentzMaterial
/ Engine_Ext_Dispersive), pinned on the CPU oracle by the closed-form
dispersion of the circuit it builds (operator_ext_lorentzmaterial.cpp:233-250):

    eps(f) = eps_r * (1 + f_p^2 / (f_L^2 - f^2))        (no damping)

A PEC cavity filled with such a medium resonates where f^2 eps(f) = f_v^2,
f_v the vacuum resonance of the same mode.  For Drude (f_L = 0) that is
f^2 = f_v^2 + f_p^2; for Lorentz the two roots of
f^4 - (f_L^2 + f_p^2 + f_v^2) f^2 + f_v^2 f_L^2 = 0.  The measured TE
resonances (the cavity of TESTSUITE/combinedtests/cavity.m) must lie within
0.2 % of these roots (roots of two modes closer than 2.5 % are not checked);
a planted 10 % error in the plasma term moves them out.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, oracle_from_operator
from tests.test_oracle import fft_time2freq
from wipems_amd.operator import BC_PEC, Operator

C0 = 299792458.0
A, B, D = 5e-2, 2e-2, 6e-2
TE = [(1, 0, 1), (1, 0, 2), (2, 0, 1)]


def f_vac(m, n, l):
    return C0 / 2 * math.sqrt((m / A) ** 2 + (n / B) ** 2 + (l / D) ** 2)


def roots(fv, fp, fl):
    if fl == 0:
        return [math.sqrt(fv * fv + fp * fp)]
    s = fl * fl + fp * fp + fv * fv
    disc = math.sqrt(s * s - 4 * fv * fv * fl * fl)
    return [math.sqrt((s + disc) / 2), math.sqrt((s - disc) / 2)]


def filled_cavity(fp, fl, steps=20000):
    mx, my, mz = np.linspace(0, A, 26), np.linspace(0, B, 11), np.linspace(0, D, 32)
    op = Operator()
    op.set_mesh([mx, my, mz], 1.0)
    op.set_boundary([BC_PEC] * 6)
    op.set_gauss_excite(4.5e9, 4.5e9)
    m = op.add_lorentz_material(eps_r=1.0, plasma_freq=fp, lor_pole_freq=fl)
    op.add_box(m, (0, 0, 0), (A, B, D), priority=0)
    i, j, k = 26 * 2 // 3 - 1, 11 * 2 // 3 - 1, 32 * 2 // 3 - 1
    e = op.add_excitation(0, (1.0, 1.0, 1.0))
    op.add_box(e, (mx[i], my[j], mz[k]), (mx[i], my[j], mz[k + 1]), priority=10)
    op.add_box(e, (mx[i], my[j], mz[k + 1]), (mx[i + 1], my[j], mz[k + 1]), priority=10)
    op.add_box(e, (mx[i + 1], my[j], mz[k + 1]), (mx[i + 1], my[j + 1], mz[k + 1]), priority=10)
    data = op.build(0, nr_ts=steps)
    x4, y2, z5 = 26 // 4 - 1, 11 // 2 - 1, 32 // 5 - 1
    data.probes = [op.voltage_probe((mx[x4], my[y2], mz[z5]), (mx[x4], my[y2 + 1], mz[z5]), td_interval=1, name="uy")]
    return data


def peak_near(f, mag, f0, span=0.03):
    sel = np.where((f > f0 * (1 - span)) & (f < f0 * (1 + span)))[0]
    k = sel[np.argmax(mag[sel])]
    y0, y1, y2 = np.log(mag[k - 1]), np.log(mag[k]), np.log(mag[k + 1])
    shift = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    return f[k] + shift * (f[1] - f[0])


def spectrum(o, steps):
    o.iterate(steps)
    t, u = o.read_probe(0)
    i0 = int(np.argmin(np.abs(t - 7e-10)))
    f, v = fft_time2freq(t[i0:], u[i0:])
    return f, np.abs(v)


@pytest.mark.parametrize("fp,fl", [(3e9, 0.0), (2e9, 2e9)])
def test_filled_cavity_resonances_follow_dispersion(fp, fl):
    data = filled_cavity(fp, fl)
    assert len(data.ade) == 1 and data.ade[0].pos.shape[1] > 0
    f, mag = spectrum(oracle_from_operator(data), 20000)
    # every cavity mode's roots, to skip checks where two lines overlap
    allr = []
    for m in range(5):
        for n in range(2):
            for l in range(5):
                if (m > 0) + (n > 0) + (l > 0) >= 2:
                    allr += roots(f_vac(m, n, l), fp, fl)
    checked = 0
    for mode in TE:
        fv = f_vac(*mode)
        for fr in roots(fv, fp, fl):
            if fr > 9e9 or fr < 1.2e9:
                continue
            if sum(abs(r / fr - 1) < 0.025 for r in allr) > 1:
                continue
            got = peak_near(f, mag, fr, span=0.02)
            assert abs(got / fr - 1) < 2e-3, (mode, fr, got)
            assert abs(fr / fv - 1) > 2e-2       # the medium moved this mode measurably
            checked += 1
    assert checked >= 3


def test_planted_plasma_error_moves_resonance():
    data = filled_cavity(3e9, 0.0)
    a = data.ade[0]
    # 10 % too strong coupling v_ext (the L_D term) in the oracle
    lists = [np.array(x, copy=True) for x in a.lists]
    lists[1] = (lists[1] * np.float32(1.1)).astype(lists[1].dtype)
    a.lists = lists
    o = oracle_from_operator(data)
    f, mag = spectrum(o, 20000)
    fv = f_vac(1, 0, 1)
    fr = roots(fv, 3e9, 0.0)[0]
    got = peak_near(f, mag, fr, span=0.05)
    assert abs(got / fr - 1) > 6e-3, got


def test_drude_ade_coefficients_closed_form():
    """v_int = 1, v_ext = dT/L_D*VI with L_D = 1/(w_p^2 C) when undamped,
    and v_int = (2L - dT R)/(2L + dT R) with R = L/tau when damped."""
    for tau in (0.0, 1e-10):
        mx = np.linspace(0, 1e-2, 6)
        op = Operator()
        op.set_mesh([mx, mx, mx], 1.0)
        op.set_boundary([BC_PEC] * 6)
        op.set_gauss_excite(0, 5e9)
        m = op.add_lorentz_material(eps_r=2.0, plasma_freq=4e9, relax_time=tau)
        op.add_box(m, (0, 0, 0), (1e-2, 1e-2, 1e-2))
        d = op.build(1, nr_ts=10)
        ade = d.ade[0]
        eps0 = 8.85418781762e-12
        h = 2e-3
        wp = 2 * math.pi * 4e9
        k = int(np.where((ade.pos[0] == 2) & (ade.pos[1] == 2) & (ade.pos[2] == 2))[0][0])
        x, y, z = 2, 2, 2
        for n in range(3):
            C = eps0 * 2.0 * h * h / h        # interior edge, uniform filling
            LD = 1 / (wp * wp * C)
            R = LD / tau if tau > 0 else 0.0
            vi = d.vi[n, x, y, z]
            if vi == 0:
                continue
            assert ade.lists[0][n][k] == pytest.approx((2 * LD - d.dT * R) / (2 * LD + d.dT * R), rel=1e-12)
            assert ade.lists[1][n][k] == pytest.approx(d.dT / (LD + d.dT * R / 2) * vi, rel=1e-9)
