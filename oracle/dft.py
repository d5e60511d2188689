"""CPU restatement of the frequency-domain field dump over a box of mesh nodes.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
field-box accumulation; the product never routes through it.

Follows
* ProcessFieldsFD::Process (Common/processfields_fd.cpp:72-107): at every
  timestep divisible by the FD interval, X_f += F(T) * exp(-2 pi i f T) * 2 * dT
  * FD_Interval, T = numTS*dT (E) or (numTS + 0.5)*dT (H, dual time,
  openems.cpp:594-598 and FDTD/engine_interface_fdtd.h:52);
* node interpolation (NODE_INTERPOLATE, the nf2ff box's dump_mode 1,
  python/openEMS/nf2ff.py:61-77) of
  - E: Engine_Interface_FDTD::GetRawInterpolatedField
    (FDTD/engine_interface_fdtd.cpp:72-97), raw E = V / primary edge length
    (:257-262);
  - H: Engine_Interface_FDTD::GetRawInterpolatedDualField (:134-158),
    raw H = I / dual edge length (:115-124).
The reference stores the interpolated values as float and accumulates in
complex<float>; this restatement, like the device code, keeps FP64 throughout.
dump_value / dump_box restate the other dump types (J, rot H, D, B) and the
other interpolations (none, cell) node by node.
"""
from __future__ import annotations

import numpy as np


def _axis_view(shape_len, n):
    sh = [1, 1, 1]
    sh[n] = shape_len
    return sh


def interp_e(V, lens, start, stop):
    """Node-interpolated E on the nodes start..stop (inclusive): [3][nx][ny][nz] float64.
    V is the [3][Nx][Ny][Nz] voltage array, lens[n] the primary edge lengths of axis n."""
    N = V.shape[1:]
    out = np.empty((3,) + tuple(stop[k] - start[k] + 1 for k in range(3)))
    for n in range(3):
        if N[n] < 2:
            out[n] = 0.0
            continue
        sub = [slice(start[k], stop[k] + 1) for k in range(3)]
        sub[n] = slice(None)
        Fn = V[n][tuple(sub)].astype(np.float64)     # whole axis n, box range across it
        L = np.asarray(lens[n], dtype=np.float64)
        i = np.arange(start[n], stop[n] + 1)
        view = _axis_view(len(i), n)

        def raw(idx):
            a = np.take(Fn, idx, axis=n)
            d = L[idx].reshape(view)
            with np.errstate(divide="ignore", invalid="ignore"):
                r = a / d
            return np.where(d == 0.0, 0.0, r)
        upper = i == N[n] - 1
        ic = np.where(upper, i - 1, i)               # upper bound: the lower value only
        im = np.maximum(ic - 1, 0)
        here, below = raw(ic), raw(im)
        delta, deltaDown = L[ic], L[im]
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = (delta / (delta + deltaDown)).reshape(view)
        mixed = here * (1.0 - rel) + below * rel
        use_mix = ((~upper) & (i > 0) & (delta != 0.0)).reshape(view)
        zero = ((~upper) & (delta == 0.0)).reshape(view)
        out[n] = np.where(zero, 0.0, np.where(use_mix, mixed, here))
    return out


def interp_h(I, lens, start, stop):
    """Node-interpolated H on the nodes start..stop: [3][nx][ny][nz] float64.
    lens[n] are the dual edge lengths of axis n."""
    N = I.shape[1:]
    shape = tuple(stop[k] - start[k] + 1 for k in range(3))
    out = np.zeros((3,) + shape)
    pos = np.meshgrid(*[np.arange(start[k], stop[k] + 1) for k in range(3)], indexing="ij")
    edge = (pos[0] == N[0] - 1) | (pos[1] == N[1] - 1) | (pos[2] == N[2] - 1)
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        zero = edge | (pos[nP] == 0) | (pos[nPP] == 0)
        Fn = I[n].astype(np.float64)
        d = lens[n].reshape(_axis_view(N[n], n))
        with np.errstate(divide="ignore", invalid="ignore"):
            raw = np.where(d == 0.0, 0.0, Fn / d)

        def at(dP, dPP):
            q = [p.copy() for p in pos]
            q[nP] = np.maximum(q[nP] - dP, 0)
            q[nPP] = np.maximum(q[nPP] - dPP, 0)
            return raw[q[0], q[1], q[2]]
        o = at(0, 0)
        o = o + at(1, 0)
        o = o + at(1, 1)
        o = o + at(0, 1)
        out[n] = np.where(zero, 0.0, o / 4.0)
    return out


# ---------------------------------------------------------------------------
# Every dump type and interpolation, node by node, in the reference's loop
# order and operation order (pure Python: small boxes only).
#   raw primary-grid values: Engine_Interface_FDTD::GetRawField
#     (FDTD/engine_interface_fdtd.cpp:257-288): type 0 E = V/delta,
#     1 J = V*kappa/delta, 2 rot H = discrete circulation of I over the dual
#     face (area = dual lengths along nP and nPP, Operator::GetEdgeArea ->
#     GetNodeArea, FDTD/operator.cpp:227-232), 3 D = V*eps/delta;
#   raw dual-grid values: GetRawDualField (:132-142): 0 H = I/delta_dual,
#     1 B = I*mu/delta_dual;
#   interpolations: GetRawInterpolatedField (:60-116) and
#     GetRawInterpolatedDualField (:143-198).
# Dump types (Common/processfields.h:43): 0 E, 1 H, 2 J, 3 rot H, 4 D, 5 B.
# Materials are float [3][Nx][Ny][Nz] arrays (epsilon, kappa, mu), as
# Operator::Calc_ECPos stores them (FDTD/operator.cpp:1176-1188); a missing
# one gives 0, as the reference does without stored materials.
# ---------------------------------------------------------------------------
def _raw_primary(V, I, ep, ed, mats, rt, n, q):
    if rt == 2:
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        area = float(ed[nP][q[nP]]) * float(ed[nPP][q[nPP]])
        value = float(I[nPP][q[0], q[1], q[2]])
        value -= float(I[nP][q[0], q[1], q[2]])
        loc = list(q)
        if q[nPP] > 0:
            loc[nPP] -= 1
            value += float(I[nP][loc[0], loc[1], loc[2]])
            loc[nPP] += 1
        if q[nP] > 0:
            loc[nP] -= 1
            value -= float(I[nPP][loc[0], loc[1], loc[2]])
        return value / area
    value = float(V[n][q[0], q[1], q[2]])
    delta = float(ep[n][q[n]])
    if rt == 0 and delta != 0.0:
        return value / delta
    m = mats.get(1) if rt == 1 else mats.get(0) if rt == 3 else None
    if m is not None and delta != 0.0:
        return value * float(m[n][q[0], q[1], q[2]]) / delta
    return 0.0


def _raw_dual(I, ed, mats, rt, n, q):
    value = float(I[n][q[0], q[1], q[2]])
    delta = float(ed[n][q[n]])
    if rt == 0 and delta != 0.0:
        return value / delta
    m = mats.get(2)
    if rt == 1 and m is not None and delta != 0.0:
        return value * float(m[n][q[0], q[1], q[2]]) / delta
    return 0.0


def dump_value(V, I, ep, ed, mats, dump_type, interp, pos):
    """The 3 components of dump `dump_type` with interpolation `interp`
    (0 none, 1 node, 2 cell) at node pos.  ep / ed: primary / dual edge
    lengths per axis; mats: {0: eps, 1: kappa, 2: mu}."""
    N = V.shape[1:]
    out = [0.0, 0.0, 0.0]
    if dump_type not in (1, 5):
        rt = {0: 0, 2: 1, 3: 2, 4: 3}[dump_type]
        for n in range(3):
            q = list(pos)
            if interp == 0:
                out[n] = _raw_primary(V, I, ep, ed, mats, rt, n, q)
            elif interp == 1:
                if N[n] < 2:
                    out[n] = 0.0
                    continue
                if pos[n] == N[n] - 1:
                    q[n] -= 1
                    out[n] = _raw_primary(V, I, ep, ed, mats, rt, n, q)
                    continue
                delta = float(ep[n][q[n]])
                out[n] = _raw_primary(V, I, ep, ed, mats, rt, n, q)
                if delta == 0.0:
                    out[n] = 0.0
                    continue
                if pos[n] == 0:
                    continue
                q[n] -= 1
                delta_down = float(ep[n][q[n]])
                rel = delta / (delta + delta_down)
                out[n] = out[n] * (1.0 - rel) + _raw_primary(V, I, ep, ed, mats, rt, n, q) * rel
            else:
                nP, nPP = (n + 1) % 3, (n + 2) % 3
                if pos[0] == N[0] - 1 or pos[1] == N[1] - 1 or pos[2] == N[2] - 1:
                    out[n] = 0.0
                    continue
                o = _raw_primary(V, I, ep, ed, mats, rt, n, q)
                q[nP] += 1
                o += _raw_primary(V, I, ep, ed, mats, rt, n, q)
                q[nPP] += 1
                o += _raw_primary(V, I, ep, ed, mats, rt, n, q)
                q[nP] -= 1
                o += _raw_primary(V, I, ep, ed, mats, rt, n, q)
                out[n] = o / 4.0
        return out
    rt = 0 if dump_type == 1 else 1
    for n in range(3):
        q = list(pos)
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        if interp == 0:
            out[n] = _raw_dual(I, ed, mats, rt, n, q)
        elif interp == 1:
            if pos[0] == N[0] - 1 or pos[1] == N[1] - 1 or pos[2] == N[2] - 1 or pos[nP] == 0 or pos[nPP] == 0:
                out[n] = 0.0
                continue
            o = _raw_dual(I, ed, mats, rt, n, q)
            q[nP] -= 1
            o += _raw_dual(I, ed, mats, rt, n, q)
            q[nPP] -= 1
            o += _raw_dual(I, ed, mats, rt, n, q)
            q[nP] += 1
            o += _raw_dual(I, ed, mats, rt, n, q)
            out[n] = o / 4.0
        else:
            delta = float(ed[n][q[n]])
            out[n] = _raw_dual(I, ed, mats, rt, n, q)
            if pos[n] >= N[n] - 1:
                out[n] = 0.0
                continue
            q[n] += 1
            delta_up = float(ed[n][q[n]])
            rel = delta / (delta + delta_up)
            out[n] = out[n] * (1.0 - rel) + _raw_dual(I, ed, mats, rt, n, q) * rel
    return out


def float_ftz(x):
    """Values rounded to float as the reference stores dump fields
    (ProcessFields::CalcField, Common/processfields.cpp:280-400, FDTD_FLOAT
    arrays) on a thread with flush-to-zero set: Engine_sse's constructor
    calls Denormal::Disable (FDTD/engine_sse.cpp:38-39, tools/denormal.h),
    so a double that rounds to a float denormal becomes a zero of its sign."""
    y = np.asarray(x, dtype=np.float64).astype(np.float32)
    return np.where(np.abs(y) < np.finfo(np.float32).tiny, np.copysign(np.float32(0), y), y)


def split_edges(edges, num_lines):
    e = np.asarray(edges, dtype=np.float64)
    N = num_lines
    return [e[:N[0]], e[N[0]:N[0] + N[1]], e[N[0] + N[1]:]]


def dump_box(V, I, edge_primary, edge_dual, mats, dump_type, interp, start, stop):
    """dump_value over the nodes start..stop (inclusive): [3][nx][ny][nz] float64."""
    N = V.shape[1:]
    ep, ed = split_edges(edge_primary, N), split_edges(edge_dual, N)
    shape = tuple(stop[k] - start[k] + 1 for k in range(3))
    out = np.empty((3,) + shape)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                v = dump_value(V, I, ep, ed, mats, dump_type, interp, (start[0] + i, start[1] + j, start[2] + k))
                out[:, i, j, k] = v
    return out


def dump_time(dump_type, ts, dT):
    """Sample time of a dump: the dual time (ts + 1/2) dT for H only
    (openems.cpp:595-600, FDTD/engine_interface_fdtd.h:52)."""
    return (float(ts) + (0.5 if dump_type == 1 else 0.0)) * dT


class DftBox:
    """One field box: accumulators [n_freq][3][nx][ny][nz] complex128."""

    def __init__(self, field, start, stop, freqs, fd_interval, edge_primary, edge_dual, num_lines, interp=1, mats=None):
        self.field = int(field)
        self.interp = int(interp)
        self.mats = mats or {}
        self.edge_primary, self.edge_dual = edge_primary, edge_dual
        self.start = tuple(int(v) for v in start)
        self.stop = tuple(int(v) for v in stop)
        self.freqs = np.asarray(freqs, dtype=np.float64)
        self.interval = int(fd_interval)
        N = num_lines
        ep = np.asarray(edge_dual if self.field else edge_primary, dtype=np.float64)
        self.lens = [ep[:N[0]], ep[N[0]:N[0] + N[1]], ep[N[0] + N[1]:]]
        shape = tuple(self.stop[k] - self.start[k] + 1 for k in range(3))
        self.acc = np.zeros((len(self.freqs), 3) + shape, dtype=np.complex128)
        self.samples = 0

    def values(self, F):
        if self.field == 0:
            return interp_e(F, self.lens, self.start, self.stop)
        return interp_h(F, self.lens, self.start, self.stop)

    def accumulate(self, F, ts, dT, vals=None):
        """Add one sample of field array F ([3][Nx][Ny][Nz]: V for E, I for H) at timestep ts."""
        if vals is None:
            vals = self.values(F)
        t = dump_time(self.field, ts, dT)
        g = dT * float(self.interval)
        for k, f in enumerate(self.freqs):
            arg = -2.0 * np.pi * f * t
            pr = (np.cos(arg) * 2.0) * g
            pi = (np.sin(arg) * 2.0) * g
            self.acc[k].real += vals * pr
            self.acc[k].imag += vals * pi
        self.samples += 1

    def accumulate_dump(self, V, I, ts, dT):
        """Add one sample of any dump type and interpolation (dump_box) from
        the voltage and current arrays at timestep ts."""
        vals = dump_box(V, I, self.edge_primary, self.edge_dual, self.mats, self.field, self.interp, self.start, self.stop)
        self.accumulate(None, ts, dT, vals)


def run_with_boxes(o, boxes, nsteps, dT, sampled0):
    """Advance Oracle o by nsteps, accumulating every box at its sample steps
    (numTS divisible by its interval; numTS = 0 on the first call)."""
    if not sampled0:
        for b in boxes:
            b.accumulate(o.get_field(b.field), 0, dT)
    left = int(nsteps)
    while left:
        ts = o.num_timesteps()
        nxt = min((ts // b.interval + 1) * b.interval for b in boxes)
        run = min(left, nxt - ts)
        o.iterate(run)
        left -= run
        if o.num_timesteps() == nxt:
            for b in boxes:
                if nxt % b.interval == 0:
                    b.accumulate(o.get_field(b.field), nxt, dT)
    return True
