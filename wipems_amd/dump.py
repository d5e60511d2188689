"""HDF5 field dumps in the reference's on-disk layout (include/wipems_dump.h).

* ``H5Dump``: one dump file (ProcessFields::InitProcess writes the mesh of the
  box and the version attribute, Common/processfields.cpp:126-158).
* ``TdDump``: a time-domain dump box on an Engine (ProcessFieldsTD): the engine
  evaluates the dump type (E, H, J, rot H, D, B) with its interpolation on the
  device and hands every sample to this writer while stepping continues.
* ``write_fd_dump``: the frequency-domain result of a field box
  (ProcessFieldsFD::DumpFDData) after the run.
* ``write_sar_dump``: local SAR or SAR raw data of an E box with cell
  interpolation (ProcessFieldsSAR::DumpFDData).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

from ._native import FdtdError, check, dump_lib

TD_SINK = ctypes.CFUNCTYPE(None, c_void_p, c_int, c_uint32, c_double, POINTER(c_float), c_uint64)


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


class H5Dump:
    """One HDF5 dump file for a box of mesh nodes.  lines: the box's mesh lines
    per axis in drawing units; unit: metres per drawing unit."""

    def __init__(self, path, lines, unit):
        self._lib = dump_lib()
        self.path = str(path)
        ls = [np.ascontiguousarray(l, dtype=np.float64) for l in lines]
        self.shape = tuple(len(l) for l in ls)
        h = c_void_p()
        self._chk(self._lib.wdump_create(self.path.encode(), _dp(ls[0]), len(ls[0]), _dp(ls[1]), len(ls[1]),
                                         _dp(ls[2]), len(ls[2]), float(unit), byref(h)))
        self._h = h

    def _chk(self, rc):
        return check(self._lib, rc, "wdump_last_error")

    def write_td(self, ts, time, field):
        """field: [3][nx][ny][nz] float32 array, or a ctypes float pointer to one."""
        if isinstance(field, np.ndarray):
            f = np.ascontiguousarray(field, dtype=np.float32)
            assert f.shape == (3,) + self.shape, (f.shape, self.shape)
            ptr = f.ctypes.data_as(c_void_p)
        else:
            ptr = ctypes.cast(field, c_void_p)
        self._chk(self._lib.wdump_write_td(self._h, int(ts), float(time), ptr))

    def write_fd(self, freqs, acc):
        """acc: complex [n_freq][3][nx][ny][nz] (Engine.read_dft)."""
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        a = np.ascontiguousarray(acc, dtype=np.complex128)
        assert a.shape == (len(fr), 3) + self.shape, (a.shape, self.shape)
        pairs = a.view(np.float64)
        self._chk(self._lib.wdump_write_fd(self._h, len(fr), _dp(fr), _dp(pairs)))

    def write_sar(self, freqs, acc, kappa, density, volume, raw=False):
        """Local SAR (raw=False, dump type 20) or SAR raw data (raw=True,
        type 29) of the cell-interpolated E accumulators acc (complex
        [n_freq][3][nx][ny][nz]) with the cell materials of
        Operator.cell_materials (include/wipems_dump.h, wdump_write_sar)."""
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        a = np.ascontiguousarray(acc, dtype=np.complex128)
        assert a.shape == (len(fr), 3) + self.shape, (a.shape, self.shape)
        cells = [np.ascontiguousarray(c, dtype=np.float32) for c in (kappa, density, volume)]
        for c in cells:
            assert c.shape == self.shape, (c.shape, self.shape)
        fp = lambda c: c.ctypes.data_as(POINTER(c_float))
        self._chk(self._lib.wdump_write_sar(self._h, len(fr), _dp(fr), _dp(a.view(np.float64)), fp(cells[0]), fp(cells[1]),
                                            fp(cells[2]), 1 if raw else 0))

    def close(self):
        if self._h:
            h, self._h = self._h, None
            self._chk(self._lib.wdump_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def box_lines(lines, start, stop, dual=False):
    """The dump file's mesh of nodes start..stop: the primary lines, or with
    dual=True the dual nodes of Operator::GetDiscLine (FDTD/operator.cpp:136-149),
    half way to the next line and half a cell past the last one."""
    out = []
    for n in range(3):
        l = np.asarray(lines[n], dtype=np.float64)
        if not dual:
            out.append(l[start[n]:stop[n] + 1].copy())
            continue
        idx = np.arange(start[n], stop[n] + 1)
        nxt = np.where(idx + 1 < len(l), l[np.minimum(idx + 1, len(l) - 1)], 0.0)
        last = l[idx] + 0.5 * (l[idx] - l[np.maximum(idx - 1, 0)])
        out.append(np.where(idx + 1 < len(l), 0.5 * (l[idx] + nxt), last))
    return out


def dump_on_dual_mesh(dump_type, interp):
    """Whether a dump's file mesh is the dual one: H dumps start on the dual
    mesh (openems.cpp:595-600), cell interpolation moves any dump there and
    node interpolation to the primary mesh (ProcessFields::SetDumpMode,
    Common/processfields.cpp:53-61)."""
    if interp == 2:
        return True
    if interp == 1:
        return False
    return dump_type == 1


class TdDump:
    """Time-domain dump box writing every sample of `engine` to `path`
    (dump type `field`: 0 E, 1 H, 2 J, 3 rot H, 4 D, 5 B; interpolation
    `interp`: 0 none, 1 node, 2 cell; nodes start..stop; every `interval`
    steps).  `lines` is the main mesh; the file holds the mesh the dump is on."""

    def __init__(self, engine, path, field, start, stop, interval, lines, unit, interp=1):
        self.file = H5Dump(path, box_lines(lines, start, stop, dump_on_dual_mesh(field, interp)), unit)
        self.samples = 0
        self.id = engine.add_td_box(field, start, stop, interval, self._sink, interp=interp)

    def _sink(self, ts, time, ptr, count):
        self.file.write_td(ts, time, ptr)
        self.samples += 1

    def close(self):
        self.file.close()


def write_fd_dump(path, engine, box_id, lines, unit):
    """The accumulators of frequency-domain box `box_id` (fdtd_read_dft) as a
    /FieldData/FD dump file."""
    b = engine.dft_boxes[box_id]
    f = H5Dump(path, box_lines(lines, b.start, b.stop, dump_on_dual_mesh(b.field, getattr(b, "interp", 1))), unit)
    try:
        f.write_fd(b.freqs, engine.read_dft(box_id))
    finally:
        f.close()


def write_sar_dump(path, engine, box_id, op, raw=False):
    """Local SAR (or, with raw=True, SAR raw data) of frequency-domain box
    `box_id`, an E box with cell interpolation, after the run
    (ProcessFieldsSAR::DumpFDData, Common/processfields_sar.cpp:179-335).
    op: the Operator the engine was built from, holding the materials'
    conductivity and density.  Mass-averaged SAR is not built."""
    b = engine.dft_boxes[box_id]
    if b.field != 0 or getattr(b, "interp", 1) != 2:
        raise ValueError("a SAR dump needs an E box with cell interpolation")
    kappa, density, volume = op.cell_materials(b.start, b.stop)
    f = H5Dump(path, box_lines(op.lines, b.start, b.stop, True), op.unit)
    try:
        f.write_sar(b.freqs, engine.read_dft(box_id), kappa, density, volume, raw=raw)
    finally:
        f.close()
