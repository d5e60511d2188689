"""ctypes wrapper of the CPU baseline engine (oracle/libcpubaseline.so).

BENCH/TEST INFRASTRUCTURE ONLY: the restated Engine_SSE_Compressed +
Engine_Multithread timed on the host cores beside the GPU run.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libcpubaseline.so")
        if not os.path.exists(path):
            raise RuntimeError("CPU baseline not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.cpub_create.restype = c_void_p
        L.cpub_create.argtypes = [POINTER(c_uint32), c_int]
        L.cpub_destroy.argtypes = [c_void_p]
        L.cpub_set_coeffs.argtypes = [c_void_p] + [POINTER(ctypes.c_float)] * 4
        L.cpub_set_coeffs_compressed.argtypes = [c_void_p, POINTER(c_uint32), POINTER(c_double), c_uint32]
        L.cpub_set_excitation.argtypes = [c_void_p, c_uint32, c_uint32, POINTER(c_double), c_uint32, POINTER(c_uint32),
                                          POINTER(c_uint32), POINTER(c_double), POINTER(c_uint32)]
        L.cpub_add_upml.argtypes = [c_void_p, POINTER(c_uint32), POINTER(c_uint32)] + [POINTER(ctypes.c_float)] * 6
        L.cpub_set_signal_curr.argtypes = [c_void_p, c_uint32, POINTER(c_double)]
        L.cpub_add_ade.argtypes = [c_void_p, c_uint32, POINTER(c_uint32)] + [POINTER(ctypes.c_float)] * 6 + [c_uint32]
        L.cpub_set_tfsf.argtypes = [c_void_p, c_int, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                    POINTER(c_double), POINTER(c_double), c_uint32]
        L.cpub_add_probe.argtypes = [c_void_p, c_int, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int), c_double]
        L.cpub_run.restype = c_double
        L.cpub_run.argtypes = [c_void_p, c_uint32, c_uint32]
        L.cpub_get_field.argtypes = [c_void_p, c_int, POINTER(c_double)]
        L.cpub_num_unique.restype = c_uint32
        L.cpub_num_unique.argtypes = [c_void_p]
        L.cpub_num_threads.argtypes = [c_void_p]
        L.cpub_read_samples.argtypes = [c_void_p, c_int, POINTER(c_double), c_uint32]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _up(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def probe_edges(p, num_lines):
    """Edge list (component, x, y, z, sign) of a probe, built as the engine does."""
    comp, pos, sign = [], [], []
    a, b = p.start, p.stop
    if p.kind == 0:
        for n in range(3):
            if a[n] < b[n]:
                q = list(a)
                while q[n] < b[n]:
                    comp.append(n); pos.append(tuple(q)); sign.append(1); q[n] += 1
            elif a[n] > b[n]:
                q = list(b)
                while q[n] < a[n]:
                    comp.append(n); pos.append(tuple(q)); sign.append(-1); q[n] += 1
    else:
        m = p.inside_mask
        si = [(m >> k) & 1 for k in range(3)]
        so = [(m >> (3 + k)) & 1 for k in range(3)]

        def rng(lo, hi):
            return range(lo + 1, hi + 1)
        nd = p.norm_dir
        if nd == 0:
            if so[0] and si[2]: [comp.append(1) or pos.append((b[0], i, a[2])) or sign.append(1) for i in rng(a[1], b[1])]
            if so[0] and so[1]: [comp.append(2) or pos.append((b[0], b[1], i)) or sign.append(1) for i in rng(a[2], b[2])]
            if si[0] and so[2]: [comp.append(1) or pos.append((a[0], i, b[2])) or sign.append(-1) for i in rng(a[1], b[1])]
            if si[0] and si[1]: [comp.append(2) or pos.append((a[0], a[1], i)) or sign.append(-1) for i in rng(a[2], b[2])]
        elif nd == 1:
            if si[0] and si[1]: [comp.append(2) or pos.append((a[0], a[1], i)) or sign.append(1) for i in rng(a[2], b[2])]
            if so[1] and so[2]: [comp.append(0) or pos.append((i, b[1], b[2])) or sign.append(1) for i in rng(a[0], b[0])]
            if so[0] and so[1]: [comp.append(2) or pos.append((b[0], b[1], i)) or sign.append(-1) for i in rng(a[2], b[2])]
            if si[1] and si[2]: [comp.append(0) or pos.append((i, a[1], a[2])) or sign.append(-1) for i in rng(a[0], b[0])]
        else:
            if si[1] and si[2]: [comp.append(0) or pos.append((i, a[1], a[2])) or sign.append(1) for i in rng(a[0], b[0])]
            if so[0] and si[2]: [comp.append(1) or pos.append((b[0], i, a[2])) or sign.append(1) for i in rng(a[1], b[1])]
            if so[1] and so[2]: [comp.append(0) or pos.append((i, b[1], b[2])) or sign.append(-1) for i in rng(a[0], b[0])]
            if si[0] and so[2]: [comp.append(1) or pos.append((a[0], i, b[2])) or sign.append(-1) for i in rng(a[1], b[1])]
    return comp, pos, sign


class CpuBaseline:
    def __init__(self, data, nthreads):
        self._L = lib()
        if data.prec != 0:
            raise ValueError("the CPU baseline runs the float operator")
        n = (c_uint32 * 3)(*data.num_lines)
        self._h = self._L.cpub_create(n, int(nthreads))
        self.num_lines = tuple(data.num_lines)
        if data.vv is None:
            # streamed build: per-cell index + table, compressed to 4-lane tuples here
            idx = np.ascontiguousarray(data.index, dtype=np.uint32)
            tab = np.ascontiguousarray(data.table, dtype=np.float64)
            self.unique = self._L.cpub_set_coeffs_compressed(self._h, _up(idx), _dp(tab), int(tab.shape[0]))
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (data.vv, data.vi, data.ii, data.iv)]
            self.unique = self._L.cpub_set_coeffs(self._h, *[_fp(a) for a in arrs])
        e = data.excitation
        if e is not None and len(getattr(e, "curr_dir", ())):
            raise ValueError("the CPU baseline restates voltage sources only")
        if e is not None and e.length:
            sv = np.ascontiguousarray(e.signal_volt, dtype=np.float64)
            si = np.ascontiguousarray(e.signal_curr, dtype=np.float64)
            self._L.cpub_set_signal_curr(self._h, e.length, _dp(si))
            pos = np.ascontiguousarray(e.volt_pos, dtype=np.uint32)
            self._L.cpub_set_excitation(self._h, e.length, e.period, _dp(sv), len(e.volt_dir), _up(pos),
                                        _up(np.ascontiguousarray(e.volt_dir, np.uint32)), _dp(np.ascontiguousarray(e.volt_amp, np.float64)),
                                        _up(np.ascontiguousarray(e.volt_delay, np.uint32)))
        for s in data.upml:
            st = (c_uint32 * 3)(*[int(x) for x in s.start])
            nl = (c_uint32 * 3)(*[int(x) for x in s.num_lines])
            ua = [np.ascontiguousarray(a, dtype=np.float32) for a in s.arrays]
            self._L.cpub_add_upml(self._h, st, nl, *[_fp(a) for a in ua])
        if getattr(data, "mur", None) or getattr(data, "rlc", None) is not None or getattr(data, "steady_state", None) is not None:
            raise ValueError("the CPU baseline restates the main updates, UPML, ADE, TFSF and voltage excitation only")
        self._keep = []
        for a in data.ade:
            pos = np.ascontiguousarray(a.pos, dtype=np.uint32)
            lists = [None if x is None else np.ascontiguousarray(x, dtype=np.float32) for x in a.lists]
            self._keep += [pos] + lists
            ptrs = [None if x is None else _fp(x) for x in lists]
            if self._L.cpub_add_ade(self._h, pos.shape[1], _up(pos), *ptrs, int(a.flags)):
                raise ValueError("too many ADE orders")
        if data.tfsf:
            for which in (0, 1):
                t = data.tfsf[which]
                pos = np.ascontiguousarray(t["pos"], dtype=np.uint32)
                comp = np.ascontiguousarray(t["comp"], dtype=np.uint32)
                delay = np.ascontiguousarray(t["delay"], dtype=np.uint32)
                delta = np.ascontiguousarray(t["delta"], dtype=np.float64)
                amp = np.ascontiguousarray(t["amp"], dtype=np.float64)
                self._L.cpub_set_tfsf(self._h, which, len(comp), _up(pos), _up(comp), _up(delay), _dp(delta), _dp(amp),
                                      int(data.tfsf["max_delay"]))
        self.probes = data.probes
        for p in data.probes:
            comp, pos, sign = probe_edges(p, data.num_lines)
            c = np.array(comp, np.uint32)
            ps = np.ascontiguousarray(np.array(pos, np.uint32).reshape(-1, 3).T)
            sg = np.array(sign, np.int32)
            self._L.cpub_add_probe(self._h, int(p.kind), len(comp), _up(c), _up(ps), sg.ctypes.data_as(POINTER(c_int)), float(p.weight))
        self.nthreads = self._L.cpub_num_threads(self._h)

    def run(self, nsteps, interval):
        return self._L.cpub_run(self._h, int(nsteps), int(interval))

    def get_field(self, which):
        out = np.empty((3,) + self.num_lines)
        self._L.cpub_get_field(self._h, int(which), _dp(out))
        return out

    def samples(self, probe):
        out = np.empty(1 << 20)
        n = self._L.cpub_read_samples(self._h, probe, _dp(out), len(out))
        return out[:n].copy()

    def __del__(self):
        try:
            if self._h:
                self._L.cpub_destroy(self._h)
                self._h = None
        except Exception:
            pass
