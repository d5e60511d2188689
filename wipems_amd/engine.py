"""Python mirror of the openEMS Engine interface over the HIP engine C ABI.

``Engine`` wraps one ``fdtd_engine`` handle (include/wipems_fdtd.h).  Method
names follow FDTD/engine.h: ``IterateTS``, ``GetNumberOfTimesteps``,
``GetVolt``/``GetCurr``/``SetVolt``/``SetCurr``; probes follow
ProcessVoltage/ProcessCurrent.  All compute runs on the GPU; there is no CPU
fallback (constructing an Engine without the built library or without a GPU
raises FdtdError).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

from ._native import ExcitationDesc, FdtdError, GridDesc, check, engine_lib
from .operator import OperatorData

PREC_FP32, PREC_FP64, PREC_FP16, PREC_BF16 = 0, 1, 2, 3
_PREC_NAMES = {"fp32": 0, "fp64": 1, "fp16": 2, "bf16": 3}
VOLT, CURR = 0, 1


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _up(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def _vp(a):
    return a.ctypes.data_as(c_void_p)


def _dtype_code(a):
    if a.dtype == np.float64:
        return 1
    if a.dtype == np.float32:
        return 0
    raise FdtdError("arrays must be float32 or float64")


def device_name():
    lib = engine_lib()
    buf = ctypes.create_string_buffer(256)
    check(lib, lib.fdtd_device_name(buf, 256), "fdtd_last_error")
    return buf.value.decode()


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Achieved read+write GB/s of a plain device copy of nbytes (k_copy)."""
    lib = engine_lib()
    g = c_double()
    check(lib, lib.fdtd_copy_bench(int(nbytes), int(reps), byref(g)), "fdtd_last_error")
    return g.value


class Engine:
    """One FDTD domain on the GPU (Engine_HIP)."""

    def __init__(self, num_lines, dT, precision="fp32"):
        self._lib = engine_lib()
        prec = _PREC_NAMES[precision] if isinstance(precision, str) else int(precision)
        d = GridDesc((c_uint32 * 3)(*num_lines), float(dT), prec, 0)
        h = c_void_p()
        self._chk(self._lib.fdtd_create(byref(d), byref(h)))
        self._h = h
        self.num_lines = tuple(int(x) for x in num_lines)
        self.dT = float(dT)
        self.precision = prec
        self.compressed = False
        self.factored = False
        self.probe_ids = []
        self.dft_boxes = []
        self._keep = []
        self._td_callbacks = []
        self._td_errors = []
        self._td_dumps = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fdtd_destroy(self._h)
            self._h = None
        for d in getattr(self, "_td_dumps", []):
            d.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return check(self._lib, rc, "fdtd_last_error")

    # ---- operator ----
    def set_coeffs(self, vv, vi, ii, iv):
        arrs = [np.ascontiguousarray(a) for a in (vv, vi, ii, iv)]
        code = _dtype_code(arrs[0])
        arrs = [a.astype(arrs[0].dtype, copy=False) for a in arrs]
        self._chk(self._lib.fdtd_set_coeffs(self._h, code, *[_vp(a) for a in arrs]))

    def set_coeffs_compressed(self, index, table):
        idx = np.ascontiguousarray(index, dtype=np.uint32)
        tab = np.ascontiguousarray(table, dtype=np.float64)
        self._chk(self._lib.fdtd_set_coeffs_compressed(self._h, _up(idx), 1, _vp(tab), int(tab.shape[0])))

    def set_coeffs_factored(self, fac):
        """Factored coefficients (operator.FactoredCoeffs) for float handles."""
        idx = np.ascontiguousarray(fac.index, dtype=np.uint32)
        tab = np.ascontiguousarray(fac.table, dtype=np.float32)
        fv = np.ascontiguousarray(fac.fv, dtype=np.float32)
        fc = np.ascontiguousarray(fac.fc, dtype=np.float32)
        fp = lambda a: a.ctypes.data_as(POINTER(c_float))
        self._chk(self._lib.fdtd_set_coeffs_factored(self._h, _up(idx), fp(tab), int(tab.shape[0]), fp(fv), fp(fc)))

    def set_storage_scale(self, volt_scale, curr_scale):
        v = (c_double * 3)(*volt_scale)
        c = (c_double * 3)(*curr_scale)
        self._chk(self._lib.fdtd_set_storage_scale(self._h, v, c))

    def set_excitation(self, exc):
        keep = []

        def arr(a, dt):
            b = np.ascontiguousarray(a, dtype=dt)
            if b.size == 0:
                b = np.zeros(1, dtype=dt)
            keep.append(b)
            return b

        sv, si = arr(exc.signal_volt, np.float64), arr(exc.signal_curr, np.float64)
        vp, vd, va, vl = arr(exc.volt_pos, np.uint32), arr(exc.volt_dir, np.uint32), arr(exc.volt_amp, np.float64), arr(exc.volt_delay, np.uint32)
        cp, cd, ca, cl = arr(exc.curr_pos, np.uint32), arr(exc.curr_dir, np.uint32), arr(exc.curr_amp, np.float64), arr(exc.curr_delay, np.uint32)
        d = ExcitationDesc(exc.length, exc.period, _dp(sv), _dp(si),
                           len(exc.volt_dir), _up(vp), _up(vd), _dp(va), _up(vl),
                           len(exc.curr_dir), _up(cp), _up(cd), _dp(ca), _up(cl))
        self._chk(self._lib.fdtd_set_excitation(self._h, byref(d)))

    def add_upml(self, slab):
        arrs = [np.ascontiguousarray(a) for a in slab.arrays]
        code = _dtype_code(arrs[0])
        st = (c_uint32 * 3)(*[int(x) for x in slab.start])
        nl = (c_uint32 * 3)(*[int(x) for x in slab.num_lines])
        self._chk(self._lib.fdtd_add_upml(self._h, st, nl, code, *[_vp(a) for a in arrs]))

    def add_ade(self, order):
        pos = np.ascontiguousarray(order.pos, dtype=np.uint32)
        lists = [np.ascontiguousarray(a) for a in order.lists]
        code = _dtype_code(lists[0])
        self._chk(self._lib.fdtd_add_ade(self._h, pos.shape[1], _up(pos), code, *[_vp(a) for a in lists], int(order.flags)))

    def set_rlc(self, r):
        pos = np.ascontiguousarray(r.pos, dtype=np.uint32)
        d = np.ascontiguousarray(r.dir, dtype=np.uint32)
        coef = np.ascontiguousarray(r.coef)
        self._chk(self._lib.fdtd_set_rlc(self._h, len(d), _up(pos), _up(d), _dtype_code(coef), _vp(coef)))

    def set_steady_state(self, ss):
        """Steady-state detection (Operator_Ext_SteadyState + Engine_Ext_SteadyState);
        ss = SteadyState(period, pos, dir), or None to remove it."""
        if ss is None:
            self._chk(self._lib.fdtd_set_steady_state(self._h, 0.0, 0, None, None))
            return
        pos = np.ascontiguousarray(ss.pos, dtype=np.uint32).reshape(3, -1)
        d = np.ascontiguousarray(ss.dir, dtype=np.uint32)
        self._chk(self._lib.fdtd_set_steady_state(self._h, float(ss.period), len(d), _up(pos), _up(d)))

    def steady_state_period(self):
        v = c_uint32()
        self._chk(self._lib.fdtd_get_steady_state_period(self._h, byref(v)))
        return v.value

    def GetLastDiff(self):
        """Engine_Ext_SteadyState::GetLastDiff: relative change over the last period."""
        v = c_double()
        self._chk(self._lib.fdtd_steady_state_last_diff(self._h, byref(v)))
        return v.value

    def add_mur(self, m):
        kp = np.ascontiguousarray(m.k_nyP)
        kpp = np.ascontiguousarray(m.k_nyPP, dtype=kp.dtype)
        self._chk(self._lib.fdtd_add_mur(self._h, int(m.face), _dtype_code(kp), _vp(kp), _vp(kpp), int(m.start_ts)))

    def set_tfsf(self, which, pos, comp, delay, delta, amp, max_delay):
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        comp = np.ascontiguousarray(comp, dtype=np.uint32)
        delay = np.ascontiguousarray(delay, dtype=np.uint32)
        delta = np.ascontiguousarray(delta, dtype=np.float64)
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        self._chk(self._lib.fdtd_set_tfsf(self._h, int(which), len(comp), _up(pos), _up(comp), _up(delay), _dp(delta), _dp(amp), int(max_delay)))

    def add_probe(self, p):
        fr = np.ascontiguousarray(p.freqs if len(p.freqs) else [0.0], dtype=np.float64)
        st = (c_uint32 * 3)(*p.start)
        sp = (c_uint32 * 3)(*p.stop)
        i = c_int()
        self._chk(self._lib.fdtd_add_probe(self._h, int(p.kind), st, sp, int(p.norm_dir), int(p.inside_mask), float(p.weight),
                                           int(p.td_interval), int(p.fd_interval), _dp(fr), len(p.freqs), byref(i)))
        self.probe_ids.append(i.value)
        return i.value

    # ---- frequency-domain field boxes (ProcessFieldsFD) ----
    def set_edge_lengths(self, primary, dual):
        primary = np.ascontiguousarray(primary, dtype=np.float64)
        dual = np.ascontiguousarray(dual, dtype=np.float64)
        self._chk(self._lib.fdtd_set_edge_lengths(self._h, _dp(primary), _dp(dual)))

    def set_material(self, kind, values):
        """Averaged material per edge for J (kind 1, kappa), D (kind 0,
        epsilon) and B (kind 2, mu) dumps (fdtd_set_material): float
        [3][Nx][Ny][Nz], as Operator.materials holds them."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.shape != (3,) + tuple(self.num_lines):
            raise ValueError("material array must be [3][Nx][Ny][Nz], got %s" % (v.shape,))
        self._chk(self._lib.fdtd_set_material(self._h, int(kind), v.ctypes.data_as(POINTER(c_float))))

    def add_dft_box(self, b):
        """Frequency-domain box b (DftBoxDef): dump type b.field, interpolation
        b.interp (fdtd_add_dft_dump_box)."""
        fr = np.ascontiguousarray(b.freqs, dtype=np.float64)
        st = (c_uint32 * 3)(*b.start)
        sp = (c_uint32 * 3)(*b.stop)
        i = c_int()
        self._chk(self._lib.fdtd_add_dft_dump_box(self._h, int(b.field), int(getattr(b, "interp", 1)), st, sp, _dp(fr), len(fr),
                                                  int(b.fd_interval), byref(i)))
        self.dft_boxes.append(b)
        return i.value

    def add_td_box(self, field, start, stop, interval, fn, interp=1):
        """Time-domain dump box (fdtd_add_td_dump_box) of dump type `field`
        (0 E, 1 H, 2 J, 3 rot H, 4 D, 5 B) with interpolation `interp` (0 none,
        1 node, 2 cell): fn(ts, time, ptr, count) gets every sample, ptr a
        ctypes float pointer to [3][nx][ny][nz] valid during the call.  Errors
        raised by fn surface from the next IterateTS."""
        from .dump import TD_SINK

        def sink(user, bid, ts, t, ptr, n):
            try:
                fn(ts, t, ptr, n)
            except Exception as e:          # ctypes would swallow it
                self._td_errors.append(e)
        cb = TD_SINK(sink)
        self._td_callbacks.append(cb)
        st = (c_uint32 * 3)(*start)
        sp = (c_uint32 * 3)(*stop)
        i = c_int()
        self._chk(self._lib.fdtd_add_td_dump_box(self._h, int(field), int(interp), st, sp, int(interval), ctypes.cast(cb, c_void_p),
                                                 None, byref(i)))
        return i.value

    def add_td_dump(self, path, field, start, stop, interval, lines, unit, interp=1):
        """A TD field dump written to the HDF5 file `path` (wipems_amd.dump.TdDump)."""
        from .dump import TdDump
        d = TdDump(self, path, field, start, stop, interval, lines, unit, interp=interp)
        self._td_dumps.append(d)
        return d

    def td_num_samples(self, bid):
        n = c_uint64()
        self._chk(self._lib.fdtd_td_num_samples(self._h, int(bid), byref(n)))
        return n.value

    def dft_num_samples(self, bid):
        n = c_uint64()
        self._chk(self._lib.fdtd_dft_num_samples(self._h, int(bid), byref(n)))
        return n.value

    def read_dft(self, bid):
        """Complex accumulators [n_freq][3][nx][ny][nz] of box bid."""
        b = self.dft_boxes[bid]
        shape = [len(b.freqs), 3] + [b.stop[n] - b.start[n] + 1 for n in range(3)]
        out = np.empty(shape + [2], dtype=np.float64)
        self._chk(self._lib.fdtd_read_dft(self._h, int(bid), _dp(out)))
        return out[..., 0] + 1j * out[..., 1]

    # ---- stepping (Engine::IterateTS) ----
    def IterateTS(self, n):
        self._chk(self._lib.fdtd_iterate(self._h, int(n)))
        if self._td_errors:
            e, self._td_errors = self._td_errors[0], []
            raise e

    iterate = IterateTS

    def synchronize(self):
        self._chk(self._lib.fdtd_synchronize(self._h))

    def GetNumberOfTimesteps(self):
        t = c_uint32()
        self._chk(self._lib.fdtd_get_num_timesteps(self._h, byref(t)))
        return t.value

    def set_use_graph(self, on):
        self._chk(self._lib.fdtd_set_use_graph(self._h, int(bool(on))))

    def set_fuse_hooks(self, mode=-1):
        """1/-1: one launch for the sparse hooks of a half step (where allowed), 0: one launch per hook."""
        self._chk(self._lib.fdtd_set_fuse_hooks(self._h, int(mode)))

    def hooks_fused(self):
        a = c_int()
        self._chk(self._lib.fdtd_hooks_fused(self._h, byref(a)))
        return bool(a.value)

    def set_pml_compression(self, mode=-1):
        """1: UPML coefficient rows, 0: full slab arrays, -1: rows only with full main coefficients."""
        self._chk(self._lib.fdtd_set_pml_compression(self._h, int(mode)))

    def pml_compressed(self):
        a = c_int()
        self._chk(self._lib.fdtd_pml_compressed(self._h, byref(a)))
        return bool(a.value)

    def set_pml_split(self, mode=-1):
        """Interior/shell launches for UPML: 1 split, 0 one launch, -1 auto (2-byte storage, or FP32 with compressed coefficients, interior >= half)."""
        self._chk(self._lib.fdtd_set_pml_split(self._h, int(mode)))

    def pml_split(self):
        a = c_int()
        self._chk(self._lib.fdtd_pml_split(self._h, byref(a)))
        return bool(a.value)

    def set_shell_align(self, vectors=-1):
        """With the UPML split: the interior launch's z range aligned to `vectors` 16-byte vectors (1, 2, 4, 8; -1 auto)."""
        self._chk(self._lib.fdtd_set_shell_align(self._h, int(vectors)))

    def set_ade_compression(self, mode=-1):
        """ADE coefficient tables for the orders added afterwards: 0 lists, 1 / -1 tables where they fit."""
        self._chk(self._lib.fdtd_set_ade_compression(self._h, int(mode)))

    def ade_compressed(self, order):
        """(volt, curr): whether ADE order `order` reads its coefficients from tables."""
        v, c = c_int(), c_int()
        self._chk(self._lib.fdtd_ade_compressed(self._h, int(order), byref(v), byref(c)))
        return bool(v.value), bool(c.value)

    def set_tile_order(self, band_tiles=-1, xcd=-1):
        """Tile order of the update launches: band_tiles = tiles per y band (0: whole planes, -1: the
        engine's choice), xcd = n > 0: the XCDs take chunks of n tiles in turn, 0: off, -1: the engine's choice."""
        self._chk(self._lib.fdtd_set_tile_order(self._h, int(band_tiles), int(xcd)))

    def tile_order(self):
        """(tiles per y band, 0 = whole planes; XCD chunk width, 0 = off) of the voltage update."""
        a, b = c_int(), c_int()
        self._chk(self._lib.fdtd_tile_order(self._h, byref(a), byref(b)))
        return a.value, b.value

    def set_kernel_timing(self, on):
        self._chk(self._lib.fdtd_set_kernel_timing(self._h, int(bool(on))))

    def kernel_times(self):
        a, b = c_double(), c_double()
        na, nb = c_uint32(), c_uint32()
        self._chk(self._lib.fdtd_get_kernel_times(self._h, byref(a), byref(b), byref(na), byref(nb)))
        return a.value, b.value, na.value, nb.value

    # ---- field access ----
    def energy(self):
        e = c_double()
        self._chk(self._lib.fdtd_calc_energy(self._h, byref(e)))
        return e.value

    CalcFastEnergy = energy

    def energy_noalloc(self):
        """The same energy through fdtd_calc_energy_noalloc (no device
        allocation; partial sums in the library's static scratch array)."""
        e = c_double()
        self._chk(self._lib.fdtd_calc_energy_noalloc(self._h, byref(e)))
        return e.value

    def field_absmax(self, which):
        """Per-component max |value| of the volt (0) or curr (1) field, on the device."""
        out = (c_double * 3)()
        self._chk(self._lib.fdtd_field_absmax(self._h, int(which), out))
        return np.array(list(out))

    def get_field(self, which, dtype=np.float64):
        out = np.empty((3,) + self.num_lines, dtype=dtype)
        self._chk(self._lib.fdtd_get_field(self._h, int(which), 1 if dtype == np.float64 else 0, _vp(out)))
        return out

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr)
        self._chk(self._lib.fdtd_set_field(self._h, int(which), _dtype_code(a), _vp(a)))

    def GetVolt(self, n, x, y, z):
        v = c_double()
        self._chk(self._lib.fdtd_get_value(self._h, VOLT, n, x, y, z, byref(v)))
        return v.value

    def GetCurr(self, n, x, y, z):
        v = c_double()
        self._chk(self._lib.fdtd_get_value(self._h, CURR, n, x, y, z, byref(v)))
        return v.value

    def SetVolt(self, n, x, y, z, value):
        self._chk(self._lib.fdtd_set_value(self._h, VOLT, n, x, y, z, float(value)))

    def SetCurr(self, n, x, y, z, value):
        self._chk(self._lib.fdtd_set_value(self._h, CURR, n, x, y, z, float(value)))

    def read_probe(self, pid):
        n = c_uint32()
        self._chk(self._lib.fdtd_probe_num_samples(self._h, pid, byref(n)))
        t = np.empty(max(n.value, 1))
        v = np.empty(max(n.value, 1))
        self._chk(self._lib.fdtd_read_probe(self._h, pid, _dp(t), _dp(v), n.value))
        return t[:n.value], v[:n.value]

    def read_probe_fd(self, pid, n_freq):
        re = np.zeros(max(n_freq, 1))
        im = np.zeros(max(n_freq, 1))
        self._chk(self._lib.fdtd_read_probe_fd(self._h, pid, _dp(re), _dp(im)))
        return re[:n_freq] + 1j * im[:n_freq]

    def device_bytes(self):
        b = c_uint64()
        self._chk(self._lib.fdtd_device_bytes(self._h, byref(b)))
        return b.value


def pilot_scales(data: OperatorData, steps=None, headroom=1024.0, compressed=None, samples=8):
    """Per-component power-of-two storage scales for FP16/BF16 handles.

    Runs the FP32 engine for `steps` (default: twice the excitation length)
    and records max |V_n| and |I_n| at `samples` evenly spaced steps; the
    scale of a component is the power of two nearest to max/headroom, so
    stored values peak near `headroom` (FP16 max 65504) and scaling itself
    is exact."""
    if steps is None:
        steps = max(64, 2 * (data.excitation.length if data.excitation is not None else 64))
    eng = engine_from_operator(data, precision="fp32", compressed=compressed)
    peak_v = np.zeros(3)
    peak_i = np.zeros(3)
    chunk = max(1, steps // samples)
    done = 0
    while done < steps:
        eng.iterate(chunk)
        done += chunk
        peak_v = np.maximum(peak_v, eng.field_absmax(VOLT))
        peak_i = np.maximum(peak_i, eng.field_absmax(CURR))
    eng.close()
    top_v = max(peak_v.max(), 1e-30)
    top_i = max(peak_i.max(), 1e-30)
    sv = [2.0 ** round(np.log2(max(p, top_v * 1e-3) / headroom)) for p in peak_v]
    si = [2.0 ** round(np.log2(max(p, top_i * 1e-3) / headroom)) for p in peak_i]
    return sv, si


def engine_from_operator(data: OperatorData, precision="fp32", compressed=None, max_entries=65536, scales=None,
                         factored=None, ade_compression=-1):
    """Create an Engine and load the operator, extensions and probes into it.

    compressed: True -> per-cell index + table (fails if more than max_entries
    unique tuples), False -> full coefficient arrays, None -> compress when the
    table fits in max_entries.
    factored: True -> factored coefficients (float handles; fails if they do
    not fit 65536 rows), None -> factored for FP16/BF16 handles whose plain
    table does not fit (when the Operator's full build is at hand), False ->
    never.
    ade_compression: fdtd_set_ade_compression mode for the ADE orders."""
    eng = Engine(data.num_lines, data.dT, precision)
    use_cmp = compressed
    fac = None
    if factored:
        fac = data.factor()
        if fac is None:
            raise ValueError("more than 65536 factored coefficient tuples")
        use_cmp = False
    elif compressed is None or compressed:
        if data.index is None:
            res = data.compress(max_entries if compressed is None else 0)
        else:
            res = (data.index, data.table)
        use_cmp = res is not None
        if (not use_cmp and factored is None and precision in ("fp16", "bf16") and data.source is not None
                and getattr(data.source, "_built_prec", None) == data.prec and data.vv is not None):
            fac = data.factor()
    if fac is not None:
        eng.set_coeffs_factored(fac)
    elif use_cmp:
        eng.set_coeffs_compressed(data.index, data.table)
    else:
        eng.set_coeffs(data.vv, data.vi, data.ii, data.iv)
    eng.compressed = bool(use_cmp)
    eng.factored = fac is not None
    if scales is not None:
        eng.set_storage_scale(*scales)
    if data.excitation is not None and data.excitation.length:
        eng.set_excitation(data.excitation)
    for s in data.upml:
        eng.add_upml(s)
    if data.ade:
        eng.set_ade_compression(ade_compression)
    for a in data.ade:
        eng.add_ade(a)
    for m in data.mur:
        eng.add_mur(m)
    if data.rlc is not None:
        eng.set_rlc(data.rlc)
    if data.steady_state is not None:
        eng.set_steady_state(data.steady_state)
    if data.tfsf:
        for which in (0, 1):
            t = data.tfsf[which]
            eng.set_tfsf(which, t["pos"], t["comp"], t["delay"], t["delta"], t["amp"], data.tfsf["max_delay"])
    for p in data.probes:
        eng.add_probe(p)
    for kind, values in sorted(data.materials.items()):
        eng.set_material(kind, values)
    if data.dft_boxes:
        eng.set_edge_lengths(data.edge_primary, data.edge_dual)
        for b in data.dft_boxes:
            eng.add_dft_box(b)
    return eng


def run_fdtd(eng, max_ts, max_excite_step=0, end_crit=1e-6, check_steps=1000):
    """The time loop of openEMS::RunFDTD (openems.cpp:1246-1333): step until
    max_ts timesteps or until `change` falls to end_crit.  Without a
    steady-state detection `change` is the domain energy (CalcFastEnergy)
    over the largest energy seen, sampled at the step of maximum excitation
    (the end-criterion processing's only step, :1253-1258) and at every check;
    with one it is GetLastDiff (:1313-1318).  The reference checks every 4 s
    of wall time; this checks every `check_steps` timesteps, so the result
    does not depend on speed.  Returns (timesteps, change)."""
    steady = hasattr(eng, "steady_state_period") and eng.steady_state_period() > 0
    change, max_e = 1.0, 0.0
    n = eng.GetNumberOfTimesteps()
    while n < max_ts and change > end_crit:
        nxt = min(int(max_ts), n + int(check_steps))
        if not steady and n < max_excite_step <= nxt:
            eng.IterateTS(max_excite_step - n)
            n = max_excite_step
            max_e = max(max_e, eng.energy())
        if nxt > n:
            eng.IterateTS(nxt - n)
            n = nxt
        if steady:
            change = eng.GetLastDiff()
        else:
            e = eng.energy()
            max_e = max(max_e, e)
            if max_e:
                change = e / max_e
    return n, change


class Batch:
    """Many independent domains advanced with one launch per update phase
    (fdtd_batch_* in include/wipems_fdtd.h).  Results are read from the member
    Engines after synchronize()."""

    def __init__(self, engines):
        self._lib = engine_lib()
        self.engines = list(engines)
        arr = (c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = c_void_p()
        check(self._lib, self._lib.fdtd_batch_create(arr, len(self.engines), byref(h)), "fdtd_last_error")
        self._h = h

    def _chk(self, rc):
        return check(self._lib, rc, "fdtd_last_error")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fdtd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_block(self, threads):
        self._chk(self._lib.fdtd_batch_set_block(self._h, int(threads)))

    def set_use_graph(self, on):
        self._chk(self._lib.fdtd_batch_set_use_graph(self._h, int(bool(on))))

    def set_split(self, mode):
        """1/-1: interior and shell launches for members whose UPML forms a shell, 0: one UPML-form launch."""
        self._chk(self._lib.fdtd_batch_set_split(self._h, int(mode)))

    def iterate(self, n):
        self._chk(self._lib.fdtd_batch_iterate(self._h, int(n)))

    def synchronize(self):
        self._chk(self._lib.fdtd_batch_synchronize(self._h))

    def set_kernel_timing(self, on):
        self._chk(self._lib.fdtd_batch_set_kernel_timing(self._h, int(bool(on))))

    def kernel_times(self):
        a, b = c_double(), c_double()
        n = c_uint32()
        self._chk(self._lib.fdtd_batch_get_kernel_times(self._h, byref(a), byref(b), byref(n)))
        return a.value, b.value, n.value
