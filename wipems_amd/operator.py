"""Python mirror of the openEMS Operator interface over libwipems_operator.so.

``Operator`` collects the mesh, boundary conditions, materials and sources of
a simulation (what openEMS reads from CSXCAD + the FDTD XML), runs the native
coefficient build and returns an ``OperatorData`` with everything an engine
needs: the coefficient arrays (Operator::GetVV/VI/II/IV), the excitation
lists (Operator_Ext_Excitation), UPML slabs (Operator_Ext_UPML) and ADE
orders (Operator_Ext_LorentzMaterial).
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, byref, c_double, c_float, c_int, c_uint32, c_void_p
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._native import FdtdError, check, operator_lib

BC_PEC, BC_PMC, BC_MUR, BC_PML = 0, 1, 2, 3
EXC_GAUSS, EXC_SINE, EXC_DIRAC, EXC_STEP = 0, 1, 2, 3


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _up(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def _vp(a):
    return a.ctypes.data_as(c_void_p)


def _fp(a):
    return a.ctypes.data_as(POINTER(c_float))


@dataclass
class FactoredCoeffs:
    """Coefficients as vi = m_E * ((fv[n][x] * fv[n][Nx+y]) * fv[n][Nx+Ny+z])
    (float arithmetic, that order; iv likewise with fc and m_H): a per-cell
    index into rows {vv0..2, m_E0..2, ii0..2, m_H0..2} plus per-line mesh
    factors.  max_rel_err bounds |rebuilt - built| / |built| over all cells."""
    index: np.ndarray              # uint32 [Nx][Ny][Nz]
    table: np.ndarray              # float32 [n][12]
    fv: np.ndarray                 # float32 [3][Nx+Ny+Nz]
    fc: np.ndarray                 # float32 [3][Nx+Ny+Nz]
    bits: int
    max_rel_err: float


@dataclass
class UpmlSlab:
    start: np.ndarray              # uint32[3]
    num_lines: np.ndarray          # uint32[3]
    arrays: List[np.ndarray]       # vv, vvfo, vvfn, ii, iifo, iifn, each [3][nx][ny][nz]


@dataclass
class MurFace:
    """One Mur ABC face (Operator_Ext_Mur_ABC): face = 2*axis + top, first-order
    coefficients of the two tangential components over the boundary plane
    [N(axis+1)][N(axis+2)], and the timestep from which it acts."""
    face: int
    k_nyP: np.ndarray
    k_nyPP: np.ndarray
    start_ts: int


@dataclass
class RlcCells:
    """Lumped RLC cells (Operator_Ext_LumpedRLC): pos uint32 [3][count], dir
    [count], coef [9][count] = ilv, i2v, vv2, vj1, vj2, vvd, ib0, b1, b2."""
    pos: np.ndarray
    dir: np.ndarray
    coef: np.ndarray


@dataclass
class SteadyState:
    """Steady-state detection of a periodic excitation (Operator_Ext_SteadyState,
    FDTD/extensions/operator_ext_steadystate.cpp:21-25,64-76): the signal period
    in seconds and one voltage probe per column of pos [3][n] / dir [n]."""
    period: float
    pos: np.ndarray                # uint32 [3][n]
    dir: np.ndarray                # uint32 [n]


def default_steady_state_probes(num_lines):
    """The probe set openems.cpp:1062-1089 gives the detection: the three
    components at the centre node, then for each axis n the three components
    at the centre with pos[n] scaled by 1/4 and then by 3/4.  Both factors are
    integer divisions there (1/4 == 3/4 == 0), so each of those nodes has
    pos[n] = 0 and appears twice: 21 probes in all."""
    c = [int(n) // 2 for n in num_lines]
    cols = [(tuple(c), d) for d in range(3)]
    for n in range(3):
        q = list(c)
        q[n] *= 1 // 4
        cols += [(tuple(q), d) for d in range(3)]
        q[n] *= 3 // 4
        cols += [(tuple(q), d) for d in range(3)]
    pos = np.array([p for p, _ in cols], dtype=np.uint32).T.copy()
    return pos, np.array([d for _, d in cols], dtype=np.uint32)


@dataclass
class AdeOrder:
    pos: np.ndarray                # uint32 [3][count]
    lists: List[np.ndarray]        # v_int, v_ext, v_lor, i_int, i_ext, i_lor, each [3][count]
    flags: int


@dataclass
class Excitation:
    length: int
    period: int
    signal_volt: np.ndarray
    signal_curr: np.ndarray
    volt_pos: np.ndarray           # uint32 [3][n]
    volt_dir: np.ndarray
    volt_amp: np.ndarray           # float64 (values are float-rounded for a float build)
    volt_delay: np.ndarray
    curr_pos: np.ndarray
    curr_dir: np.ndarray
    curr_amp: np.ndarray
    curr_delay: np.ndarray


@dataclass
class ProbeDef:
    kind: int                      # 0 voltage, 1 current
    start: tuple
    stop: tuple
    weight: float = 1.0
    norm_dir: int = -1
    inside_mask: int = 0x3F
    td_interval: int = 0
    fd_interval: int = 0
    freqs: tuple = ()
    name: str = ""


@dataclass
class DftBoxDef:
    """Frequency-domain field dump over a box of mesh nodes (ProcessFieldsFD).
    field: the dump type of ProcessFields (Common/processfields.h:43), 0 E,
    1 H, 2 J, 3 rot H, 4 D, 5 B; interp: 0 none, 1 node (the nf2ff boxes'
    mode), 2 cell (Common/engine_interface_base.h:33)."""
    field: int
    start: tuple
    stop: tuple
    freqs: tuple
    fd_interval: int
    name: str = ""
    interp: int = 1


def fd_interval_for(freqs, dT, nyquist):
    """Processing::AddFrequency (Common/processing.cpp:150-170): the FD
    interval starts at the excitation's Nyquist step count and shrinks to
    CalcNyquistNum(f, dT) = floor(1/f/2/dT) of the highest frequency."""
    interval = 0
    for f in freqs:
        nts = int(math.floor(1.0 / f / 2.0 / dT)) if f != 0 else 2 ** 32 - 1
        if nts == 0:
            continue
        if interval == 0:
            interval = nyquist
        if interval > nts:
            interval = nts
    if interval == 0:
        raise ValueError("no frequency below the time-step limit 1/(2 dT)")
    return interval


@dataclass
class OperatorData:
    num_lines: tuple
    dT: float
    prec: int                      # 0 float build, 1 double build
    vv: np.ndarray                 # [3][Nx][Ny][Nz]
    vi: np.ndarray
    ii: np.ndarray
    iv: np.ndarray
    excitation: Optional[Excitation]
    upml: List[UpmlSlab] = field(default_factory=list)
    ade: List[AdeOrder] = field(default_factory=list)
    nyquist: int = 0
    max_excite_step: int = 0
    probes: List[ProbeDef] = field(default_factory=list)
    tfsf: Optional[dict] = None
    index: Optional[np.ndarray] = None      # compressed form, filled by compress()
    table: Optional[np.ndarray] = None
    source: object = None                   # the Operator that built this data (native compression)
    edge_primary: Optional[np.ndarray] = None   # [Nx+Ny+Nz] primary edge lengths (m)
    edge_dual: Optional[np.ndarray] = None      # [Nx+Ny+Nz] dual edge lengths (m)
    dft_boxes: List[DftBoxDef] = field(default_factory=list)
    mur: List[MurFace] = field(default_factory=list)
    rlc: Optional[RlcCells] = None
    signal_period: float = 0.0             # Excitation::GetSignalPeriod (s), 0 when not periodic
    steady_state: Optional[SteadyState] = None
    # averaged material per edge kept for dumps (Operator::Calc_ECPos, FDTD/operator.cpp:1176-1188):
    # kind 0 epsilon, 1 kappa, 2 mu -> float [3][Nx][Ny][Nz]
    materials: dict = field(default_factory=dict)

    @property
    def cells(self):
        n = self.num_lines
        return int(n[0]) * int(n[1]) * int(n[2])

    def compress(self, max_entries=0):
        """Per-cell coefficient dedup (Operator_SSE_Compressed::CompressOperator)."""
        return compress_coeffs(self, max_entries)

    def factor(self, bits=16, max_entries=65536):
        """Factored coefficients (FactoredCoeffs) from the Operator that built
        this data; None if more than max_entries tuples."""
        if self.source is None or getattr(self.source, "_built_prec", None) != self.prec or self.vv is None:
            raise ValueError("factoring needs the full build of the Operator that made this data")
        return self.source.factor(bits, max_entries)

    def expanded(self):
        """A copy with the full [3][Nx][Ny][Nz] coefficient arrays rebuilt from
        the compressed index and table (the streamed build keeps only those),
        in the build's float type."""
        import copy
        if self.vv is not None:
            return self
        if self.index is None:
            raise ValueError("no coefficients to expand")
        dt = np.float64 if self.prec == 1 else np.float32
        tab = self.table.astype(dt)
        idx = self.index.reshape(-1)
        out = copy.copy(self)
        arrs = []
        for k in range(4):
            a = np.empty((3,) + tuple(self.num_lines), dtype=dt)
            for n in range(3):
                a[n] = tab[idx, 3 * k + n].reshape(self.num_lines)
            arrs.append(a)
        out.vv, out.vi, out.ii, out.iv = arrs
        return out


def compress_coeffs(data: OperatorData, max_entries=0):
    """Deduplicate the per-cell 12-tuples {vv,vi,ii,iv}x3 bitwise, first-seen order.

    Returns (index uint32 [Nx][Ny][Nz], table float64 [n][12]) or None when the
    number of unique tuples exceeds max_entries (> 0)."""
    if data.source is not None and getattr(data.source, "_built_prec", None) == data.prec:
        res = data.source.compress(max_entries)
        if res is not None:
            data.index, data.table = res
        return res
    nd = data.cells
    dt = np.float64 if data.prec == 1 else np.float32
    stack = np.empty((nd, 12), dtype=dt)
    for k, arr in enumerate((data.vv, data.vi, data.ii, data.iv)):
        stack[:, 3 * k:3 * k + 3] = arr.reshape(3, nd).T
    view = stack.view(np.dtype((np.void, stack.dtype.itemsize * 12)))
    uniq, first, inv = np.unique(view, return_index=True, return_inverse=True)
    if max_entries and len(uniq) > max_entries:
        return None
    # renumber by first occurrence (the reference's insertion order)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    index = rank[inv.reshape(-1)].astype(np.uint32).reshape(tuple(data.num_lines))
    table = stack[first[order]].astype(np.float64)
    data.index, data.table = index, table
    return index, table


class Operator:
    """Scenario description + native EC operator build."""

    def __init__(self):
        self._lib = operator_lib()
        h = c_void_p()
        self._chk(self._lib.wop_create(byref(h)))
        self._h = h
        self.num_lines = None
        self._store_mat = [False, False, False]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.wop_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _chk(self, rc):
        return check(self._lib, rc, "wop_last_error")

    # ---- setup ----
    def set_mesh(self, lines_xyz, unit):
        n = []
        for ax, lines in enumerate(lines_xyz):
            a = np.ascontiguousarray(lines, dtype=np.float64)
            self._chk(self._lib.wop_set_mesh(self._h, ax, _dp(a), len(a)))
            n.append(len(a))
        self._chk(self._lib.wop_set_grid_delta(self._h, float(unit)))
        self.num_lines = tuple(n)
        self.lines = [np.asarray(l, dtype=np.float64) for l in lines_xyz]
        self.unit = unit

    def set_boundary(self, bc, pml_size=(8,) * 6):
        b = (c_int * 6)(*bc)
        s = (c_uint32 * 6)(*pml_size)
        self._chk(self._lib.wop_set_boundary(self._h, b, s))

    def set_timestep_factor(self, f):
        self._chk(self._lib.wop_set_timestep_factor(self._h, float(f)))

    def set_timestep(self, dT):
        self._chk(self._lib.wop_set_timestep(self._h, float(dT)))

    def set_threads(self, n):
        self._chk(self._lib.wop_set_threads(self._h, int(n)))

    def set_verbose(self, level):
        self._chk(self._lib.wop_set_verbose(self._h, int(level)))

    def set_row_reuse(self, on=True):
        """Copy provably equal cells along z in the builds (default) or
        evaluate every cell; the result is identical (wop_set_row_reuse)."""
        self._chk(self._lib.wop_set_row_reuse(self._h, 1 if on else 0))

    def add_material(self, eps_r=1.0, kappa=0.0, mue_r=1.0, sigma=0.0):
        i = c_int()
        self._chk(self._lib.wop_add_material(self._h, eps_r, kappa, mue_r, sigma, byref(i)))
        return i.value

    def set_density(self, prop, density):
        """Mass density of material `prop` in kg/m^3 for SAR dumps
        (CSPropMaterial's Density; 0, the default, marks cells SAR skips)."""
        self._chk(self._lib.wop_set_density(self._h, int(prop), float(density)))

    def cell_materials(self, start, stop):
        """(kappa, density, volume) of every cell of the nodes start..stop
        (inclusive), each float32 [nx][ny][nz]: the conductivity and density of
        the material at the cell's centre and the cell volume in m^3, as
        ProcessFieldsSAR::DumpFDData gathers them (Common/processfields_sar.cpp:200-230)."""
        shape = tuple(int(stop[n]) - int(start[n]) + 1 for n in range(3))
        if min(shape) < 1:
            raise ValueError("stop must not be below start")
        out = [np.empty(shape, np.float32) for _ in range(3)]
        s, e = (c_uint32 * 3)(*start), (c_uint32 * 3)(*stop)
        fp = lambda a: a.ctypes.data_as(POINTER(c_float))
        self._chk(self._lib.wop_get_cell_materials(self._h, s, e, fp(out[0]), fp(out[1]), fp(out[2])))
        return tuple(out)

    def add_metal(self):
        i = c_int()
        self._chk(self._lib.wop_add_metal(self._h, byref(i)))
        return i.value

    def add_lorentz_material(self, eps_r=1.0, kappa=0.0, mue_r=1.0, sigma=0.0, plasma_freq=0.0, relax_time=0.0, lor_pole_freq=0.0):
        i = c_int()
        self._chk(self._lib.wop_add_lorentz_material(self._h, eps_r, kappa, mue_r, sigma, plasma_freq, relax_time, lor_pole_freq, byref(i)))
        return i.value

    def add_lumped_element(self, ny, R=-1.0, C=-1.0, caps=True):
        i = c_int()
        self._chk(self._lib.wop_add_lumped_element(self._h, int(ny), float(R), float(C), int(bool(caps)), byref(i)))
        return i.value

    def add_lumped_rlc(self, ny, rlc_type, R=-1.0, L=-1.0, C=-1.0, caps=True):
        """Lumped RLC element: rlc_type 0 parallel (with L > 0), 1 series; values < 0 are unset."""
        i = c_int()
        self._chk(self._lib.wop_add_lumped_rlc(self._h, int(ny), int(rlc_type), float(R), float(L), float(C),
                                               int(bool(caps)), byref(i)))
        return i.value

    def add_excitation(self, exc_type, value, delay=0.0):
        i = c_int()
        v = (c_double * 3)(*value)
        self._chk(self._lib.wop_add_excitation(self._h, int(exc_type), v, float(delay), byref(i)))
        return i.value

    def set_plane_wave(self, start, stop, prop_dir, e_vec, freq=0.0):
        """Plane-wave (TFSF) excitation on the faces of a box (openEMS
        AddPlaneWaveExcite + one box; Operator_Ext_TFSF)."""
        a = (c_double * 3)(*start)
        b = (c_double * 3)(*stop)
        k = (c_double * 3)(*prop_dir)
        e = (c_double * 3)(*e_vec)
        self._chk(self._lib.wop_set_plane_wave(self._h, a, b, k, e, float(freq)))

    def add_box(self, prop, start, stop, priority=0):
        a = (c_double * 3)(*start)
        b = (c_double * 3)(*stop)
        self._chk(self._lib.wop_add_box(self._h, int(prop), a, b, int(priority)))

    def add_sphere(self, prop, center, radius, priority=0):
        c = (c_double * 3)(*center)
        self._chk(self._lib.wop_add_sphere(self._h, int(prop), c, float(radius), int(priority)))

    def set_gauss_excite(self, f0, fc):
        self._chk(self._lib.wop_set_excite_signal(self._h, EXC_GAUSS, float(f0), float(fc)))

    def set_excite_signal(self, kind, f0=0.0, fc=0.0):
        self._chk(self._lib.wop_set_excite_signal(self._h, int(kind), float(f0), float(fc)))

    def snap(self, axis, coord, dual=False):
        """Operator::SnapToMeshLine -> (index, inside)."""
        idx = c_uint32()
        ins = c_int()
        self._chk(self._lib.wop_snap_to_mesh_line(self._h, int(axis), float(coord), int(bool(dual)), byref(idx), byref(ins)))
        return idx.value, bool(ins.value)

    # ---- build ----
    def build(self, prec=0, nr_ts=1_000_000_000) -> OperatorData:
        lib, h = self._lib, self._h
        self._chk(lib.wop_build(h, int(prec), int(nr_ts)))
        n = (c_uint32 * 3)()
        self._chk(lib.wop_get_num_lines(h, n))
        shape = (3, n[0], n[1], n[2])
        dt = np.float64 if prec == 1 else np.float32
        arrs = [np.empty(shape, dtype=dt) for _ in range(4)]
        self._chk(lib.wop_get_coeffs(h, 1 if prec == 1 else 0, *[_vp(a) for a in arrs]))
        return self._collect(prec, arrs)

    def build_streamed(self, nr_ts=1_000_000_000, max_entries=65536) -> OperatorData:
        """Float build that never holds full per-cell arrays (wop_build_streamed):
        returns OperatorData with the compressed index/table and no vv..iv."""
        lib, h = self._lib, self._h
        n = (c_uint32 * 3)()
        self._chk(lib.wop_get_num_lines(h, n))
        index = np.empty((n[0], n[1], n[2]), dtype=np.uint32)
        nu = c_uint32()
        self._chk(lib.wop_build_streamed(h, int(nr_ts), _up(index), int(max_entries), byref(nu)))
        table = np.empty((nu.value, 12), dtype=np.float64)
        self._chk(lib.wop_get_streamed_table(h, _dp(table)))
        data = self._collect(0, [None] * 4)
        data.index, data.table = index, table
        data.source = None
        return data

    def _collect(self, prec, arrs) -> OperatorData:
        lib, h = self._lib, self._h
        n = (c_uint32 * 3)()
        self._chk(lib.wop_get_num_lines(h, n))
        dt = np.float64 if prec == 1 else np.float32
        dT = c_double()
        self._chk(lib.wop_get_timestep(h, byref(dT)))
        length, nyq, mstep = c_uint32(), c_uint32(), c_uint32()
        self._chk(lib.wop_get_excite_info(h, byref(length), byref(nyq), byref(mstep)))
        sv = np.empty(length.value, dtype=np.float64)
        si = np.empty(length.value, dtype=np.float64)
        if length.value:
            self._chk(lib.wop_get_signal(h, _dp(sv), _dp(si)))
        nv, nc = c_uint32(), c_uint32()
        self._chk(lib.wop_get_excitation_counts(h, byref(nv), byref(nc)))
        nv, nc = nv.value, nc.value
        vpos = np.zeros((3, max(nv, 1)), np.uint32); vdir = np.zeros(max(nv, 1), np.uint32)
        vamp = np.zeros(max(nv, 1)); vdel = np.zeros(max(nv, 1), np.uint32)
        cpos = np.zeros((3, max(nc, 1)), np.uint32); cdir = np.zeros(max(nc, 1), np.uint32)
        camp = np.zeros(max(nc, 1)); cdel = np.zeros(max(nc, 1), np.uint32)
        self._chk(lib.wop_get_excitation(h, _up(vpos), _up(vdir), _dp(vamp), _up(vdel), _up(cpos), _up(cdir), _dp(camp), _up(cdel)))
        per, per_ts = c_double(), c_uint32()
        self._chk(lib.wop_get_signal_period(h, byref(per), byref(per_ts)))
        exc = Excitation(length.value, per_ts.value, sv, si, np.ascontiguousarray(vpos[:, :nv]), vdir[:nv].copy(), vamp[:nv].copy(), vdel[:nv].copy(),
                         np.ascontiguousarray(cpos[:, :nc]), cdir[:nc].copy(), camp[:nc].copy(), cdel[:nc].copy())
        nu = c_uint32()
        self._chk(lib.wop_get_num_upml(h, byref(nu)))
        slabs = []
        for k in range(nu.value):
            st = (c_uint32 * 3)(); nl = (c_uint32 * 3)()
            self._chk(lib.wop_get_upml_range(h, k, st, nl))
            sh = (3, nl[0], nl[1], nl[2])
            ua = [np.empty(sh, dtype=dt) for _ in range(6)]
            self._chk(lib.wop_get_upml(h, k, 1 if prec == 1 else 0, *[_vp(a) for a in ua]))
            slabs.append(UpmlSlab(np.array(list(st), np.uint32), np.array(list(nl), np.uint32), ua))
        na = c_uint32()
        self._chk(lib.wop_get_num_ade(h, byref(na)))
        ades = []
        for o in range(na.value):
            cnt, fl = c_uint32(), c_uint32()
            self._chk(lib.wop_get_ade_info(h, o, byref(cnt), byref(fl)))
            pos = np.empty((3, cnt.value), np.uint32)
            lists = [np.empty((3, cnt.value), dtype=dt) for _ in range(6)]
            self._chk(lib.wop_get_ade(h, o, _up(pos), 1 if prec == 1 else 0, *[_vp(a) for a in lists]))
            ades.append(AdeOrder(pos, lists, fl.value))
        nv, nc, md = c_uint32(), c_uint32(), c_uint32()
        self._chk(lib.wop_get_tfsf_info(h, byref(nv), byref(nc), byref(md)))
        tfsf = None
        if nv.value or nc.value:
            tfsf = {"max_delay": md.value}
            for which, cnt in ((0, nv.value), (1, nc.value)):
                t = {"pos": np.empty((3, cnt), np.uint32), "comp": np.empty(cnt, np.uint32),
                     "delay": np.empty(cnt, np.uint32), "delta": np.empty(cnt), "amp": np.empty(cnt)}
                self._chk(lib.wop_get_tfsf(h, which, _up(t["pos"]), _up(t["comp"]), _up(t["delay"]),
                                           _dp(t["delta"]), _dp(t["amp"])))
                tfsf[which] = t
        mur = self._mur_faces(n, prec)
        rlc = self._rlc_cells(prec)
        self._built_prec = prec
        tot = int(n[0]) + int(n[1]) + int(n[2])
        edges = []
        for dual in (0, 1):
            ebuf = np.empty(tot, dtype=np.float64)
            self._chk(lib.wop_get_edge_lengths(h, dual, _dp(ebuf)))
            edges.append(ebuf)
        self._edges = edges
        nl = tuple(int(x) for x in n)
        materials = {}
        for kind in range(3):
            if self._store_mat[kind]:
                m = np.empty((3,) + nl, dtype=np.float32)
                self._chk(lib.wop_get_material(h, kind, _fp(m)))
                materials[kind] = m
        # openems.cpp:1060-1091: a periodic excitation gets a steady-state detection
        ss = None
        if per.value > 0:
            spos, sdir = default_steady_state_probes(nl)
            ss = SteadyState(per.value, spos, sdir)
        return OperatorData(nl, dT.value, prec, *arrs, exc, slabs, ades, nyq.value, mstep.value,
                            tfsf=tfsf, source=self, edge_primary=edges[0], edge_dual=edges[1], mur=mur, rlc=rlc,
                            signal_period=per.value, steady_state=ss, materials=materials)

    def _rlc_cells(self, prec):
        cnt = c_uint32()
        self._chk(self._lib.wop_get_rlc_count(self._h, byref(cnt)))
        n = cnt.value
        if not n:
            return None
        pos = np.empty((3, n), np.uint32)
        d = np.empty(n, np.uint32)
        coef = np.empty((9, n), np.float64)
        self._chk(self._lib.wop_get_rlc(self._h, _up(pos), _up(d), 1, _vp(coef)))
        return RlcCells(pos, d, coef.astype(np.float64 if prec == 1 else np.float32))

    def _mur_faces(self, n, prec):
        lib, h = self._lib, self._h
        cnt = c_uint32()
        self._chk(lib.wop_get_num_mur(h, byref(cnt)))
        faces = []
        for k in range(cnt.value):
            face, start = c_int(), c_uint32()
            self._chk(lib.wop_get_mur(h, k, byref(face), 1, None, None, byref(start)))
            ny = face.value // 2
            shape = (int(n[(ny + 1) % 3]), int(n[(ny + 2) % 3]))
            kp = np.empty(shape, np.float64)
            kpp = np.empty(shape, np.float64)
            self._chk(lib.wop_get_mur(h, k, byref(face), 1, _vp(kp), _vp(kpp), byref(start)))
            dt = np.float64 if prec == 1 else np.float32
            faces.append(MurFace(face.value, kp.astype(dt), kpp.astype(dt), start.value))
        return faces

    def set_mur_phase_velocity(self, face, v_phase):
        """Phase velocity of one Mur face (face = 2*axis + top); 0 = the material's own."""
        self._chk(self._lib.wop_set_mur_phase_velocity(self._h, int(face), float(v_phase)))

    def compress(self, max_entries=0):
        """Native per-cell compression of the last build (wop_compress); None if
        more than max_entries unique tuples exist."""
        nu = c_uint32()
        rc = self._lib.wop_compress(self._h, byref(nu), None, None, int(max_entries))
        if rc != 0:
            if max_entries and nu.value > max_entries:
                return None
            self._chk(rc)
        n = self.num_lines
        index = np.empty(n, dtype=np.uint32)
        table = np.empty((nu.value, 12), dtype=np.float64)
        self._chk(self._lib.wop_compress(self._h, byref(nu), _up(index), _dp(table), 0))
        return index, table

    def factor(self, bits=16, max_entries=65536):
        """Factored coefficients of the last build (wop_factor) with m rounded
        to `bits` significant bits; None if more than max_entries tuples."""
        n = self.num_lines
        L = int(n[0]) + int(n[1]) + int(n[2])
        index = np.empty(n, dtype=np.uint32)
        table = np.empty((int(max_entries), 12), dtype=np.float32)
        fv = np.empty((3, L), dtype=np.float32)
        fc = np.empty((3, L), dtype=np.float32)
        nu = c_uint32()
        err = c_double()
        rc = self._lib.wop_factor(self._h, int(bits), int(max_entries), _up(index), _fp(table), byref(nu), _fp(fv), _fp(fc),
                                  byref(err))
        if rc == -2:
            return None
        self._chk(rc)
        return FactoredCoeffs(index, table[:nu.value].copy(), fv, fc, int(bits), err.value)

    def set_material_store_flags(self, kind, val=True):
        """Keep the averaged material of every edge for J, D or B dumps
        (Operator_Base::SetMaterialStoreFlags): kind 0 epsilon, 1 kappa,
        2 mu.  The next build fills OperatorData.materials[kind]."""
        if kind not in (0, 1, 2):
            raise ValueError("material kind must be 0 (epsilon), 1 (kappa) or 2 (mu)")
        self._chk(self._lib.wop_set_material_store_flags(self._h, int(kind), 1 if val else 0))
        self._store_mat[kind] = bool(val)

    def get_ec(self):
        n = self.num_lines
        shape = (3,) + tuple(n)
        out = [np.empty(shape) for _ in range(4)]
        self._chk(self._lib.wop_get_ec(self._h, *[_dp(a) for a in out]))
        return out

    # ---- field-box helpers ----
    def nf2ff_box(self, data, start, stop, freqs, directions=(True,) * 6, name="nf2ff"):
        """The six faces of an nf2ff recording box (python/openEMS/nf2ff.py:61-77):
        one E and one H frequency-domain dump per enabled face, node
        interpolated (dump_mode 1), coordinates snapped to the main mesh.
        Appends the boxes to data.dft_boxes and returns them."""
        s = [self.snap(n, min(start[n], stop[n]), False)[0] for n in range(3)]
        e = [self.snap(n, max(start[n], stop[n]), False)[0] for n in range(3)]
        interval = fd_interval_for(freqs, data.dT, data.nyquist)
        out = []
        for ny in range(3):
            for side in (0, 1):
                if not directions[2 * ny + side]:
                    continue
                a, b = list(s), list(e)
                if side == 0:
                    b[ny] = a[ny]
                else:
                    a[ny] = b[ny]
                for fld, tag in ((0, "E"), (1, "H")):
                    out.append(DftBoxDef(fld, tuple(a), tuple(b), tuple(float(f) for f in freqs), interval,
                                         "%s_%s_%s%s" % (name, tag, "xyz"[ny], "-+"[side])))
        data.dft_boxes.extend(out)
        return out

    # ---- probe helpers (ProcessIntegral::DefineStartStopCoord snapping) ----
    def voltage_probe(self, start, stop, weight=1.0, **kw):
        """Voltage probe on a line between two coordinates (snapped to the main mesh)."""
        s = tuple(self.snap(n, start[n], False)[0] for n in range(3))
        e = tuple(self.snap(n, stop[n], False)[0] for n in range(3))
        return ProbeDef(0, s, e, weight=weight, **kw)

    def current_probe(self, start, stop, norm_dir, weight=1.0, **kw):
        """Current probe on a (possibly degenerate) box; snapped to the dual mesh
        and widened to a loop as in ProcessCurrent::DefineStartStopCoord
        (Common/processcurrent.cpp:39-90)."""
        s, e, ins_s, ins_e = [], [], [], []
        for n in range(3):
            a, ia = self.snap(n, min(start[n], stop[n]), True)
            b, ib = self.snap(n, max(start[n], stop[n]), True)
            s.append(a); e.append(b); ins_s.append(ia); ins_e.append(ib)
        for n in range(3):
            if n == norm_dir:
                continue
            if start[n] == stop[n]:
                dual = lambda i: 0.5 * (self.lines[n][i] + self.lines[n][i + 1]) if i < len(self.lines[n]) - 1 else \
                    self.lines[n][i] + 0.5 * (self.lines[n][i] - self.lines[n][i - 1])
                if dual(s[n]) > start[n] and s[n] > 0:
                    s[n] -= 1
                if dual(s[n]) < start[n] and e[n] < len(self.lines[n]) - 1:
                    e[n] += 1
        mask = sum((1 << n) for n in range(3) if ins_s[n]) | sum((1 << (3 + n)) for n in range(3) if ins_e[n])
        return ProbeDef(1, tuple(s), tuple(e), weight=weight, norm_dir=norm_dir, inside_mask=mask, **kw)
