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


class DftBox:
    """One field box: accumulators [n_freq][3][nx][ny][nz] complex128."""

    def __init__(self, field, start, stop, freqs, fd_interval, edge_primary, edge_dual, num_lines):
        self.field = int(field)
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

    def accumulate(self, F, ts, dT):
        """Add one sample of field array F ([3][Nx][Ny][Nz]: V for E, I for H) at timestep ts."""
        vals = self.values(F)
        t = (float(ts) + (0.5 if self.field else 0.0)) * dT
        g = dT * float(self.interval)
        for k, f in enumerate(self.freqs):
            arg = -2.0 * np.pi * f * t
            pr = (np.cos(arg) * 2.0) * g
            pi = (np.sin(arg) * 2.0) * g
            self.acc[k].real += vals * pr
            self.acc[k].imag += vals * pi
        self.samples += 1


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
