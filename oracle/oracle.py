"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never routes
through it.
"""
from __future__ import annotations

import ctypes
import math
import os
from ctypes import POINTER, c_double, c_int, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_create.restype = c_void_p
        L.orc_create.argtypes = [POINTER(c_uint32), c_double, c_int, c_int]
        L.orc_destroy.argtypes = [c_void_p]
        L.orc_set_coeffs.argtypes = [c_void_p, c_int] + [c_void_p] * 4
        L.orc_set_excitation.argtypes = [c_void_p, c_uint32, c_uint32, POINTER(c_double), POINTER(c_double),
                                         c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), POINTER(c_uint32),
                                         c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), POINTER(c_uint32)]
        L.orc_add_upml.argtypes = [c_void_p, POINTER(c_uint32), POINTER(c_uint32), c_int] + [c_void_p] * 6
        L.orc_add_ade.argtypes = [c_void_p, c_uint32, POINTER(c_uint32), c_int] + [c_void_p] * 6 + [c_uint32]
        L.orc_add_mur.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_uint32]
        L.orc_set_rlc.argtypes = [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32), c_int, c_void_p]
        L.orc_set_tfsf.argtypes = [c_void_p, c_int, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                   POINTER(c_double), POINTER(c_double), c_uint32]
        L.orc_add_probe.argtypes = [c_void_p, c_int, POINTER(c_uint32), POINTER(c_uint32), c_int, c_uint32, c_double,
                                    c_uint32, c_uint32, POINTER(c_double), c_uint32]
        L.orc_iterate.argtypes = [c_void_p, c_uint32]
        L.orc_num_timesteps.argtypes = [c_void_p]
        L.orc_num_timesteps.restype = c_uint32
        L.orc_get_field.argtypes = [c_void_p, c_int, POINTER(c_double)]
        L.orc_set_field.argtypes = [c_void_p, c_int, POINTER(c_double)]
        L.orc_read_probe.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_double), c_uint32]
        L.orc_read_probe_fd.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_double)]
        L.orc_energy.argtypes = [c_void_p]
        L.orc_energy.restype = c_double
        L.orc_set_steady_state.argtypes = [c_void_p, c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32)]
        L.orc_steady_state_last_diff.argtypes = [c_void_p]
        L.orc_steady_state_last_diff.restype = c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _up(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def _vp(a):
    return a.ctypes.data_as(c_void_p)


class Oracle:
    """CPU restatement of the reference engine (float with FTZ, or double)."""

    def __init__(self, num_lines, dT, use_double=False, ftz=True):
        self._L = lib()
        n = (c_uint32 * 3)(*num_lines)
        self._h = self._L.orc_create(n, float(dT), int(use_double), int(ftz))
        self.num_lines = tuple(num_lines)
        self.dT = float(dT)
        self._nprobe = 0

    def __del__(self):
        try:
            if self._h:
                self._L.orc_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_coeffs(self, vv, vi, ii, iv):
        arrs = [np.ascontiguousarray(a) for a in (vv, vi, ii, iv)]
        code = 1 if arrs[0].dtype == np.float64 else 0
        self._L.orc_set_coeffs(self._h, code, *[_vp(a) for a in arrs])

    def set_excitation(self, e):
        keep = []

        def arr(a, dt):
            b = np.ascontiguousarray(a, dtype=dt)
            if b.size == 0:
                b = np.zeros(1, dtype=dt)
            keep.append(b)
            return b
        sv, si = arr(e.signal_volt, np.float64), arr(e.signal_curr, np.float64)
        self._L.orc_set_excitation(self._h, e.length, e.period, _dp(sv), _dp(si),
                                   len(e.volt_dir), _up(arr(e.volt_pos, np.uint32)), _up(arr(e.volt_dir, np.uint32)),
                                   _dp(arr(e.volt_amp, np.float64)), _up(arr(e.volt_delay, np.uint32)),
                                   len(e.curr_dir), _up(arr(e.curr_pos, np.uint32)), _up(arr(e.curr_dir, np.uint32)),
                                   _dp(arr(e.curr_amp, np.float64)), _up(arr(e.curr_delay, np.uint32)))

    def add_upml(self, s):
        arrs = [np.ascontiguousarray(a) for a in s.arrays]
        code = 1 if arrs[0].dtype == np.float64 else 0
        st = (c_uint32 * 3)(*[int(x) for x in s.start])
        nl = (c_uint32 * 3)(*[int(x) for x in s.num_lines])
        self._L.orc_add_upml(self._h, st, nl, code, *[_vp(a) for a in arrs])

    def add_ade(self, a):
        pos = np.ascontiguousarray(a.pos, dtype=np.uint32)
        lists = [np.ascontiguousarray(x) for x in a.lists]
        code = 1 if lists[0].dtype == np.float64 else 0
        self._L.orc_add_ade(self._h, pos.shape[1], _up(pos), code, *[_vp(x) for x in lists], int(a.flags))

    def set_rlc(self, r):
        pos = np.ascontiguousarray(r.pos, dtype=np.uint32)
        d = np.ascontiguousarray(r.dir, dtype=np.uint32)
        coef = np.ascontiguousarray(r.coef)
        code = 1 if coef.dtype == np.float64 else 0
        self._L.orc_set_rlc(self._h, len(d), _up(pos), _up(d), code, _vp(coef))

    def set_steady_state(self, ss):
        """Operator_Ext_SteadyState::BuildExtension rounds the period to whole
        timesteps with C round() (half away from zero)."""
        x = ss.period / self.dT
        p = math.floor(x)
        if x - p >= 0.5:
            p += 1
        pos = np.ascontiguousarray(ss.pos, dtype=np.uint32).reshape(3, -1)
        d = np.ascontiguousarray(ss.dir, dtype=np.uint32)
        if self._L.orc_set_steady_state(self._h, int(p), len(d), _up(pos), _up(d)):
            raise ValueError("steady-state period below one timestep")
        return int(p)

    def GetLastDiff(self):
        return self._L.orc_steady_state_last_diff(self._h)

    def add_mur(self, m):
        kp = np.ascontiguousarray(m.k_nyP)
        kpp = np.ascontiguousarray(m.k_nyPP, dtype=kp.dtype)
        code = 1 if kp.dtype == np.float64 else 0
        self._L.orc_add_mur(self._h, int(m.face), code, _vp(kp), _vp(kpp), int(m.start_ts))

    def set_tfsf(self, which, pos, comp, delay, delta, amp, max_delay):
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        comp = np.ascontiguousarray(comp, dtype=np.uint32)
        delay = np.ascontiguousarray(delay, dtype=np.uint32)
        delta = np.ascontiguousarray(delta, dtype=np.float64)
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        self._L.orc_set_tfsf(self._h, int(which), len(comp), _up(pos), _up(comp), _up(delay), _dp(delta), _dp(amp), int(max_delay))

    def add_probe(self, p):
        fr = np.ascontiguousarray(p.freqs if len(p.freqs) else [0.0], dtype=np.float64)
        st = (c_uint32 * 3)(*p.start)
        sp = (c_uint32 * 3)(*p.stop)
        rc = self._L.orc_add_probe(self._h, int(p.kind), st, sp, int(p.norm_dir), int(p.inside_mask), float(p.weight),
                                   int(p.td_interval), int(p.fd_interval), _dp(fr), len(p.freqs))
        if rc < 0:
            raise RuntimeError("oracle probe table full")
        return rc

    def iterate(self, n):
        self._L.orc_iterate(self._h, int(n))

    def num_timesteps(self):
        return self._L.orc_num_timesteps(self._h)

    def get_field(self, which):
        out = np.empty((3,) + self.num_lines)
        self._L.orc_get_field(self._h, int(which), _dp(out))
        return out

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._L.orc_set_field(self._h, int(which), _dp(a))

    def read_probe(self, pid, max_n=10_000_000):
        t = np.empty(max_n if max_n < 100000 else 100000)
        v = np.empty_like(t)
        n = self._L.orc_read_probe(self._h, pid, _dp(t), _dp(v), len(t))
        if n > len(t):
            t = np.empty(n); v = np.empty(n)
            self._L.orc_read_probe(self._h, pid, _dp(t), _dp(v), n)
        return t[:n].copy(), v[:n].copy()

    def read_probe_fd(self, pid, n_freq):
        re = np.zeros(max(n_freq, 1)); im = np.zeros(max(n_freq, 1))
        self._L.orc_read_probe_fd(self._h, pid, _dp(re), _dp(im))
        return re[:n_freq] + 1j * im[:n_freq]

    def energy(self):
        return self._L.orc_energy(self._h)


def oracle_from_operator(data, use_double=None, ftz=True):
    """Oracle loaded with the same operator data an Engine gets."""
    dbl = data.prec == 1 if use_double is None else use_double
    o = Oracle(data.num_lines, data.dT, dbl, ftz)
    o.set_coeffs(data.vv, data.vi, data.ii, data.iv)
    if data.excitation is not None and data.excitation.length:
        o.set_excitation(data.excitation)
    for s in data.upml:
        o.add_upml(s)
    for a in data.ade:
        o.add_ade(a)
    for m in getattr(data, "mur", []):
        o.add_mur(m)
    if getattr(data, "rlc", None) is not None:
        o.set_rlc(data.rlc)
    if getattr(data, "steady_state", None) is not None:
        o.set_steady_state(data.steady_state)
    if data.tfsf:
        for which in (0, 1):
            t = data.tfsf[which]
            o.set_tfsf(which, t["pos"], t["comp"], t["delay"], t["delta"], t["amp"], data.tfsf["max_delay"])
    for p in data.probes:
        o.add_probe(p)
    return o
