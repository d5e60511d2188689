"""HDF5 field dumps in the reference's on-disk layout (include/wipems_dump.h).

* ``H5Dump``: one dump file (ProcessFields::InitProcess writes the mesh of the
  box and the version attribute, Common/processfields.cpp:126-158).
* ``TdDump``: a time-domain dump box on an Engine (ProcessFieldsTD): the engine
  snapshots the node-interpolated field on the device and hands every sample
  to this writer while stepping continues.
* ``write_fd_dump``: the frequency-domain result of a field box
  (ProcessFieldsFD::DumpFDData) after the run.
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

    def close(self):
        if self._h:
            h, self._h = self._h, None
            self._chk(self._lib.wdump_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def box_lines(lines, start, stop):
    return [np.asarray(lines[n], dtype=np.float64)[start[n]:stop[n] + 1] for n in range(3)]


class TdDump:
    """Time-domain dump box writing every sample of `engine` to `path`
    (field 0 = E, 1 = H; nodes start..stop; every `interval` steps)."""

    def __init__(self, engine, path, field, start, stop, interval, lines, unit):
        self.file = H5Dump(path, box_lines(lines, start, stop), unit)
        self.samples = 0
        self.id = engine.add_td_box(field, start, stop, interval, self._sink)

    def _sink(self, ts, time, ptr, count):
        self.file.write_td(ts, time, ptr)
        self.samples += 1

    def close(self):
        self.file.close()


def write_fd_dump(path, engine, box_id, lines, unit):
    """The accumulators of frequency-domain box `box_id` (fdtd_read_dft) as a
    /FieldData/FD dump file."""
    b = engine.dft_boxes[box_id]
    f = H5Dump(path, box_lines(lines, b.start, b.stop), unit)
    try:
        f.write_fd(b.freqs, engine.read_dft(box_id))
    finally:
        f.close()
