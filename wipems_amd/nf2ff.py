"""Near-field to far-field transform on the GPU (include/wipems_nf2ff.h).

``NF2FF`` mirrors nf2ff_calc (nf2ff/nf2ff_calc.cpp) for Cartesian meshes; ``far_field`` feeds it the six faces of an nf2ff box
recorded by an Engine (Operator.nf2ff_box), the data the reference's nf2ff
reads back from the boxes' HDF5 dumps (python/openEMS/nf2ff.py:79-139).
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_double, c_void_p

import numpy as np

from ._native import check, engine_lib


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


class NF2FF:
    def __init__(self, freq, theta, phi, center=(0.0, 0.0, 0.0), radius=1.0, eps_r=1.0, mue_r=1.0):
        self._lib = engine_lib()
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)
        self.phi = np.ascontiguousarray(phi, dtype=np.float64)
        c = np.ascontiguousarray(center, dtype=np.float64)
        h = c_void_p()
        self._chk(self._lib.nf2ff_create(float(freq), _dp(self.theta), len(self.theta), _dp(self.phi), len(self.phi),
                                         _dp(c), float(radius), float(eps_r), float(mue_r), byref(h)))
        self._h = h

    def _chk(self, rc):
        return check(self._lib, rc, "fdtd_last_error")

    MIRROR_OFF, MIRROR_PEC, MIRROR_PMC = 0, 1, 2

    def set_mirror(self, kind, axis, pos):
        """nf2ff_calc::SetMirror: a PEC (1) or PMC (2) symmetry plane at pos (m)."""
        self._chk(self._lib.nf2ff_set_mirror(self._h, int(kind), int(axis), float(pos)))

    def add_plane(self, lines, E, H):
        """lines: node coordinates (m) per axis, one axis with a single line;
        E, H: complex [3][nx][ny][nz]."""
        ls = [np.ascontiguousarray(l, dtype=np.float64).reshape(-1) for l in lines]
        shape = (3,) + tuple(len(l) for l in ls)
        e = np.ascontiguousarray(E, dtype=np.complex128).reshape(shape)
        hh = np.ascontiguousarray(H, dtype=np.complex128).reshape(shape)
        self._chk(self._lib.nf2ff_add_plane(self._h, _dp(ls[0]), len(ls[0]), _dp(ls[1]), len(ls[1]), _dp(ls[2]), len(ls[2]),
                                            _dp(e.view(np.float64)), _dp(hh.view(np.float64))))

    def results(self):
        """dict: E_theta, E_phi, H_theta, H_phi (complex [n_theta][n_phi]),
        P_rad, rad_power, max_dir."""
        sh = (len(self.theta), len(self.phi))
        out = {k: np.zeros(sh, np.complex128) for k in ("E_theta", "E_phi", "H_theta", "H_phi")}
        prad = np.zeros(sh)
        pw, md = c_double(), c_double()
        self._chk(self._lib.nf2ff_get(self._h, _dp(out["E_theta"].view(np.float64)), _dp(out["E_phi"].view(np.float64)),
                                      _dp(out["H_theta"].view(np.float64)), _dp(out["H_phi"].view(np.float64)), _dp(prad),
                                      byref(pw), byref(md)))
        out.update(P_rad=prad, rad_power=pw.value, max_dir=md.value)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nf2ff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def far_field(engine, lines, unit, freq_index, theta, phi, center=(0.0, 0.0, 0.0), radius=1.0, box_ids=None):
    """Far field at the box frequency `freq_index` from the engine's nf2ff
    faces: E and H frequency-domain boxes on the same nodes, paired in the
    order Operator.nf2ff_box appends them.  lines: mesh lines per axis in
    drawing units, unit: metres per drawing unit; center in metres."""
    ids = list(range(len(engine.dft_boxes))) if box_ids is None else list(box_ids)
    by_nodes = {}
    for i in ids:
        b = engine.dft_boxes[i]
        by_nodes.setdefault((tuple(b.start), tuple(b.stop)), {})[b.field] = i
    freq = engine.dft_boxes[ids[0]].freqs[freq_index]
    nf = NF2FF(freq, theta, phi, center, radius)
    try:
        for (a, b), pair in by_nodes.items():
            if 0 not in pair or 1 not in pair:
                continue
            pl = [np.asarray(lines[n], dtype=np.float64)[a[n]:b[n] + 1] * unit for n in range(3)]
            E = engine.read_dft(pair[0])[freq_index]
            H = engine.read_dft(pair[1])[freq_index]
            nf.add_plane(pl, E, H)
        return nf.results()
    finally:
        nf.close()
