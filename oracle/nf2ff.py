"""CPU restatement of the near-field to far-field transform (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device transform
(libwipems_fdtd.so, nf2ff_*); the product never routes through it.

nf2ff_calc (nf2ff/nf2ff_calc.cpp) on a Cartesian mesh without mirror planes:

* AddSinglePlane (:357-528) per face of nodes (one axis with a single line):
  normal sign +1 when the face lies at or above the centre on its normal
  axis, else -1; node areas from half the neighbouring line spacings
  (:381-391); radiated power += 1/2 area Re(E_nP conj(H_nPP) - E_nPP
  conj(H_nP)) * normal sign (:414-423).
* nf2ff_calc_thread (:40-163): Js = n x H, Ms = -n x E (:66-78); for each
  (theta, phi), with r.rhat measured from the centre,
    Nt = sum area e^{jk r.rhat} (Js_x ct cp + Js_y ct sp - Js_z st),
    Np = sum area e^{jk r.rhat} (Js_y cp - Js_x sp), Lt/Lp alike from Ms.
* Far field (:488-505): factor = jk/(4 pi R) e^{-jkR},
  E_theta -= factor (Lp + Z0 Nt), E_phi += factor (Lt - Z0 Np),
  H_theta += factor (Np - Lt/Z0), H_phi -= factor (Nt + Lp/Z0),
  P_rad = (|E_theta|^2 + |E_phi|^2) / (2 Z0); Dmax = 4 pi max(P_rad) / P.
The reference evaluates the terms in float with double accumulators; this
restatement runs in float64 throughout, so comparisons use tolerances.
"""
import math

import numpy as np

C0 = 299792458.0                 # __C0__, tools/constants.h:25
Z0 = 376.730313461               # __Z0__, tools/constants.h:26


def node_widths(lines):
    """Half the spacing to each neighbour (nf2ff_calc.cpp:381-391)."""
    l = np.asarray(lines, dtype=np.float64)
    w = np.empty(len(l))
    w[1:-1] = 0.5 * np.abs(l[2:] - l[:-2])
    w[0] = 0.5 * abs(l[1] - l[0])
    w[-1] = 0.5 * abs(l[-1] - l[-2])
    return w


class NF2FF:
    def __init__(self, freq, theta, phi, center=(0.0, 0.0, 0.0), radius=1.0, eps_r=1.0, mue_r=1.0):
        self.freq = float(freq)
        self.theta = np.asarray(theta, dtype=np.float64)
        self.phi = np.asarray(phi, dtype=np.float64)
        self.center = np.asarray(center, dtype=np.float64)
        self.radius = float(radius)
        self.k = 2 * math.pi * self.freq / C0 * math.sqrt(eps_r * mue_r)
        self.Z = Z0 * math.sqrt(mue_r / eps_r)
        sh = (len(self.theta), len(self.phi))
        self.E_theta = np.zeros(sh, complex)
        self.E_phi = np.zeros(sh, complex)
        self.H_theta = np.zeros(sh, complex)
        self.H_phi = np.zeros(sh, complex)
        self.P_rad = np.zeros(sh)
        self.rad_power = 0.0
        self.max_dir = 0.0

    MIRROR_OFF, MIRROR_PEC, MIRROR_PMC = 0, 1, 2

    def set_mirror(self, kind, axis, pos):
        """nf2ff_calc::SetMirror (:231-245): a PEC or PMC symmetry plane."""
        if not hasattr(self, "mirror"):
            self.mirror = [0, 0, 0]
            self.mirror_pos = [0.0, 0.0, 0.0]
        self.mirror[axis] = kind
        self.mirror_pos[axis] = float(pos)

    def _mirror_plane(self, n, lines, E, H):
        """AddMirrorPlane (:247-279): mirror the lines of axis n about the
        plane, scale the fields in place, add the image face."""
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        lines[n] = 2.0 * self.mirror_pos[n] - lines[n]
        ef, hf = np.ones(3), np.ones(3)
        if self.mirror[n] == self.MIRROR_PEC:
            hf[n] = -1.0
            ef[nP] = ef[nPP] = -1.0
        else:
            ef[n] = -1.0
            hf[nP] = hf[nPP] = -1.0
        for d in range(3):
            E[d] *= ef[d]
            H[d] *= hf[d]
        self._single(lines, E, H)

    def add_plane(self, lines, E, H):
        """AddPlane (:281-328): the face and, with mirror planes set, its
        images in the reference's order."""
        lines = [np.array(l, dtype=np.float64) for l in lines]
        E = np.array(E, dtype=complex)
        H = np.array(H, dtype=complex)
        self._single(lines, E, H)
        m = getattr(self, "mirror", [0, 0, 0])
        for n in range(3):
            nP, nPP = (n + 1) % 3, (n + 2) % 3
            if m[n] and not m[nP] and not m[nPP]:
                self._mirror_plane(n, lines, E, H)
                break
            if not m[n] and m[nP] and m[nPP]:
                self._mirror_plane(nP, lines, E, H)
                self._mirror_plane(nPP, lines, E, H)
                self._mirror_plane(nP, lines, E, H)
                break
        if m[0] and m[1] and m[2]:
            for n in (0, 1, 0, 2, 0, 1, 0):
                self._mirror_plane(n, lines, E, H)

    def _single(self, lines, E, H):
        """AddSinglePlane: lines: three arrays of node coordinates (m), one of
        length 1; E, H: complex [3][n0][n1][n2]."""
        n = [len(l) for l in lines]
        ny = [a for a in range(3) if n[a] == 1 and n[(a + 1) % 3] > 2 and n[(a + 2) % 3] > 2]
        if not ny:
            raise ValueError("cannot determine the normal direction")
        ny = ny[-1]
        nP, nPP = (ny + 1) % 3, (ny + 2) % 3
        E = np.asarray(E, complex).reshape(3, -1)
        H = np.asarray(H, complex).reshape(3, -1)
        sign = 1.0 if lines[ny][0] >= self.center[ny] else -1.0
        nd = np.zeros(3)
        nd[ny] = sign
        grid = np.meshgrid(*[np.asarray(l, dtype=np.float64) for l in lines], indexing="ij")
        pos = np.stack([g.reshape(-1) for g in grid])            # [3][nodes]
        w = [node_widths(lines[a]) if n[a] > 1 else np.ones(1) for a in range(3)]
        wg = np.meshgrid(*w, indexing="ij")
        area = (wg[nP] * wg[nPP]).reshape(-1)
        p = E[nP] * np.conj(H[nPP]) - E[nPP] * np.conj(H[nP])
        self.rad_power += float(np.sum(0.5 * area * p.real * sign))
        Js = np.stack([nd[1] * H[2] - nd[2] * H[1], nd[2] * H[0] - nd[0] * H[2], nd[0] * H[1] - nd[1] * H[0]])
        Ms = np.stack([nd[2] * E[1] - nd[1] * E[2], nd[0] * E[2] - nd[2] * E[0], nd[1] * E[0] - nd[0] * E[1]])
        st, ct = np.sin(self.theta)[:, None], np.cos(self.theta)[:, None]
        sp, cp = np.sin(self.phi)[None, :], np.cos(self.phi)[None, :]
        rel = pos - self.center[:, None]
        Nt = np.zeros(self.E_theta.shape, complex)
        Np, Lt, Lp = Nt.copy(), Nt.copy(), Nt.copy()
        for i in range(len(self.theta)):
            for j in range(len(self.phi)):
                u = np.array([cp[0, j] * st[i, 0], sp[0, j] * st[i, 0], ct[i, 0]])
                ph = area * np.exp(1j * self.k * (u @ rel))
                S = Js @ ph
                T = Ms @ ph
                Nt[i, j] = S[0] * ct[i, 0] * cp[0, j] + S[1] * ct[i, 0] * sp[0, j] - S[2] * st[i, 0]
                Np[i, j] = S[1] * cp[0, j] - S[0] * sp[0, j]
                Lt[i, j] = T[0] * ct[i, 0] * cp[0, j] + T[1] * ct[i, 0] * sp[0, j] - T[2] * st[i, 0]
                Lp[i, j] = T[1] * cp[0, j] - T[0] * sp[0, j]
        factor = 1j * self.k / (4 * math.pi * self.radius) * np.exp(-1j * self.k * self.radius)
        self.E_theta -= factor * (Lp + self.Z * Nt)
        self.E_phi += factor * (Lt - self.Z * Np)
        self.H_theta += factor * (Np - Lt / self.Z)
        self.H_phi -= factor * (Nt + Lp / self.Z)
        self.P_rad = (np.abs(self.E_theta) ** 2 + np.abs(self.E_phi) ** 2) / (2 * self.Z)
        self.max_dir = 4 * math.pi * self.P_rad.max() / self.rad_power


def box_faces(lo, hi, counts):
    """Node lines of the six faces of a box (x-, x+, y-, y+, z-, z+)."""
    axes = [np.linspace(lo[a], hi[a], counts[a]) for a in range(3)]
    faces = []
    for ny in range(3):
        for v in (lo[ny], hi[ny]):
            l = list(axes)
            l[ny] = np.array([v])
            faces.append(l)
    return faces


def hertzian_dipole_fields(points, freq, Il=1.0, at=(0.0, 0.0, 0.0)):
    """Exact E, H of a z-directed infinitesimal dipole of moment I*l at `at`
    (free space), at points [3][n] (m): complex [3][n] each."""
    k = 2 * math.pi * freq / C0
    x, y, z = (np.asarray(points, dtype=np.float64) - np.asarray(at, dtype=np.float64)[:, None])
    r = np.sqrt(x * x + y * y + z * z)
    ct = z / r
    st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
    rho = np.sqrt(x * x + y * y)
    cp = np.where(rho > 0, x / np.where(rho > 0, rho, 1), 1.0)
    sp = np.where(rho > 0, y / np.where(rho > 0, rho, 1), 0.0)
    g = np.exp(-1j * k * r)
    kr = k * r
    Er = Z0 * Il * ct / (2 * math.pi * r * r) * (1 + 1 / (1j * kr)) * g
    Et = 1j * Z0 * k * Il * st / (4 * math.pi * r) * (1 + 1 / (1j * kr) - 1 / (kr * kr)) * g
    Hp = 1j * k * Il * st / (4 * math.pi * r) * (1 + 1 / (1j * kr)) * g
    E = np.stack([Er * st * cp + Et * ct * cp, Er * st * sp + Et * ct * sp, Er * ct - Et * st])
    H = np.stack([-Hp * sp, Hp * cp, np.zeros_like(Hp)])
    return E, H


def hertzian_dipole_power(freq, Il=1.0):
    k = 2 * math.pi * freq / C0
    return Z0 * (k * Il) ** 2 / (12 * math.pi)
