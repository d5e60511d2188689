"""Near-field to far-field transform (nf2ff_calc).

Oracle (oracle/nf2ff.py) pinned by the exact near field of a z-directed
Hertzian dipole sampled on a closed box: the far field converges to
E_theta = j Z0 k Il sin(theta) e^{-jkR} / (4 pi R) at second order in the
node spacing, E_phi vanishes, the Poynting flux through the faces gives the
dipole's power Z0 (k Il)^2 / (12 pi) and the maximum directivity is 1.5.
The device transform is compared with the oracle on the same data: the
dipole box, random fields on a graded off-centre box in a dielectric, and the
six faces an FDTD run records around a soft source.
"""
import math

import numpy as np
import pytest

from oracle.nf2ff import NF2FF as OracleNF2FF
from oracle.nf2ff import C0, Z0, box_faces, hertzian_dipole_fields, hertzian_dipole_power

F = 1e9
THETA = np.linspace(0, math.pi, 19)
PHI = np.array([0.0, 0.7, math.pi / 2])


def _dipole_box(cnt, target):
    for lines in box_faces((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), (cnt, cnt, cnt)):
        g = np.meshgrid(*lines, indexing="ij")
        pts = np.stack([a.reshape(-1) for a in g])
        E, H = hertzian_dipole_fields(pts, F)
        shape = (3,) + tuple(len(l) for l in lines)
        target.add_plane(lines, E.reshape(shape), H.reshape(shape))
    return target


def test_oracle_hertzian_dipole():
    k = 2 * math.pi * F / C0
    ref = 1j * Z0 * k * np.sin(THETA) / (4 * math.pi) * np.exp(-1j * k)
    errs = []
    for cnt in (21, 41, 81):
        nf = _dipole_box(cnt, OracleNF2FF(F, THETA, PHI))
        errs.append(np.abs(nf.E_theta[:, 0] - ref).max() / np.abs(ref).max())
        # the pattern does not depend on phi
        np.testing.assert_allclose(np.abs(nf.E_theta[:, 2]), np.abs(nf.E_theta[:, 0]), rtol=2e-3, atol=1e-6 * np.abs(ref).max())
        assert np.abs(nf.E_phi).max() < 1e-3 * np.abs(ref).max()
        # H = rhat x E / Z0 in the far field
        np.testing.assert_allclose(nf.H_phi, nf.E_theta / Z0, rtol=0, atol=1e-3 * np.abs(ref).max() / Z0)
    assert errs[2] < 1e-4
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5      # second order
    assert nf.rad_power == pytest.approx(hertzian_dipole_power(F), rel=2e-4)
    assert nf.max_dir == pytest.approx(1.5, rel=1e-4)


def test_oracle_planted_errors_fail():
    """A flipped magnetic current or a lost normal sign breaks the pins."""
    k = 2 * math.pi * F / C0
    ref = 1j * Z0 * k * np.sin(THETA) / (4 * math.pi) * np.exp(-1j * k)

    class FlipMs(OracleNF2FF):
        def add_plane(self, lines, E, H):
            return super().add_plane(lines, -np.asarray(E), H)

    nf = _dipole_box(41, FlipMs(F, THETA, PHI))
    assert np.abs(nf.E_theta[:, 0] - ref).max() / np.abs(ref).max() > 0.1

    class WrongCenter(OracleNF2FF):
        pass
    nf = _dipole_box(41, WrongCenter(F, THETA, PHI, center=(0.2, 0.0, 0.0)))   # every face on one side
    assert abs(nf.rad_power - hertzian_dipole_power(F)) > 0.1 * hertzian_dipole_power(F)


# ---------------------------------------------------------------------------
# device transform
# ---------------------------------------------------------------------------
def _compare(dev, orc, rtol=1e-8):
    r = dev.results()
    scale = np.abs(orc.E_theta).max() + np.abs(orc.E_phi).max()
    for name in ("E_theta", "E_phi", "H_theta", "H_phi"):
        s = scale if name.startswith("E") else scale / Z0
        np.testing.assert_allclose(r[name], getattr(orc, name), rtol=0, atol=rtol * s, err_msg=name)
    np.testing.assert_allclose(r["P_rad"], orc.P_rad, rtol=0, atol=rtol * orc.P_rad.max() * 2)
    assert r["rad_power"] == pytest.approx(orc.rad_power, rel=1e-10)
    assert r["max_dir"] == pytest.approx(orc.max_dir, rel=rtol * 10)
    return r


@pytest.mark.gpu
def test_gpu_nf2ff_hertzian_dipole(gpu):
    from wipems_amd.nf2ff import NF2FF
    print("case: nf2ff dipole box", flush=True)
    dev = _dipole_box(41, NF2FF(F, THETA, PHI))
    orc = _dipole_box(41, OracleNF2FF(F, THETA, PHI))
    r = _compare(dev, orc)
    assert r["max_dir"] == pytest.approx(1.5, rel=2e-4)
    dev.close()


@pytest.mark.gpu
def test_gpu_nf2ff_random_graded(gpu):
    from wipems_amd.nf2ff import NF2FF
    print("case: nf2ff random graded box", flush=True)
    rng = np.random.default_rng(9)
    axes = [np.sort(rng.uniform(-0.05, 0.07, n)) for n in (9, 13, 7)]
    lo = [a[0] for a in axes]
    hi = [a[-1] for a in axes]
    theta = np.linspace(0.1, 3.0, 11)
    phi = np.linspace(0, 2 * math.pi, 8, endpoint=False)
    kw = dict(center=(0.01, -0.02, 0.005), radius=3.0, eps_r=2.2, mue_r=1.3)
    dev = NF2FF(2.5e9, theta, phi, **kw)
    orc = OracleNF2FF(2.5e9, theta, phi, **kw)
    for ny in range(3):
        for v in (lo[ny], hi[ny]):
            lines = list(axes)
            lines[ny] = np.array([v])
            shape = (3,) + tuple(len(l) for l in lines)
            E = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            H = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            dev.add_plane(lines, E, H)
            orc.add_plane(lines, E, H)
    _compare(dev, orc)
    dev.close()


@pytest.mark.gpu
def test_gpu_nf2ff_from_fdtd_run(gpu):
    """Six E/H face pairs recorded around a soft Ez source in a UPML box:
    the device transform equals the oracle on the same accumulators, and the
    short source radiates with directivity near 1.5."""
    from wipems_amd import scenarios
    from wipems_amd.engine import engine_from_operator
    from wipems_amd.nf2ff import far_field
    from wipems_amd.operator import BC_PML
    print("case: nf2ff from an FDTD run", flush=True)
    # an 8 GHz carrier (lambda = 37 cells): the 8-cell UPML absorbs it, and
    # a baseband pulse's DC content would dominate a low-frequency sample
    sc = scenarios.cavity_point_source(n_cells=48, steps=900, bc=BC_PML, f0=8e9, fc=4e9)
    op, d = sc.op, sc.data
    c = op.lines[0][24]
    op.nf2ff_box(d, (op.lines[0][14],) * 3, (op.lines[0][34],) * 3, [8e9])
    eng = engine_from_operator(d, precision="fp32")
    eng.iterate(900)
    theta = np.linspace(0, math.pi, 13)
    phi = np.array([0.0, math.pi / 2])
    center = (c * op.unit, c * op.unit, (c + 0.5) * op.unit)
    r = far_field(eng, op.lines, op.unit, 0, theta, phi, center=center)
    orc = OracleNF2FF(8e9, theta, phi, center=center)
    for i in range(0, len(eng.dft_boxes), 2):
        b = eng.dft_boxes[i]
        pl = [np.asarray(op.lines[n])[b.start[n]:b.stop[n] + 1] * op.unit for n in range(3)]
        orc.add_plane(pl, eng.read_dft(i)[0], eng.read_dft(i + 1)[0])
    np.testing.assert_allclose(r["E_theta"], orc.E_theta, rtol=0, atol=1e-8 * np.abs(orc.E_theta).max())
    assert r["rad_power"] == pytest.approx(orc.rad_power, rel=1e-10)
    assert r["max_dir"] == pytest.approx(1.5, abs=0.02)
    assert r["rad_power"] > 0


# ---------------------------------------------------------------------------
# mirror planes (image theory)
# ---------------------------------------------------------------------------
L, P0 = 0.08, (0.02, -0.015, 0.03)
MIRROR_CASES = {
    "pec_z": {2: 1},
    "pmc_z": {2: 2},
    "pec_x_pmc_y": {0: 1, 1: 2},
    "pec_xyz": {0: 1, 1: 1, 2: 1},
}


def _images(mirrors):
    """The z-dipole at P0 and its images: across a plane normal to z a PEC
    keeps its sign and a PMC flips it; across a plane normal to x or y it is
    the other way round."""
    out = [(np.array(P0), 1.0)]
    for axis, kind in mirrors.items():
        new = []
        for p, s in out:
            q = p.copy()
            q[axis] = -q[axis]
            flip = (kind == 2) if axis == 2 else (kind == 1)
            new.append((q, -s if flip else s))
        out += new
    return out


def _fields(lines, images):
    g = np.meshgrid(*lines, indexing="ij")
    pts = np.stack([a.reshape(-1) for a in g])
    shape = (3,) + tuple(len(l) for l in lines)
    E = np.zeros((3, pts.shape[1]), complex)
    H = np.zeros_like(E)
    for p, s in images:
        e, h = hertzian_dipole_fields(pts, F, at=p)
        E += s * e
        H += s * h
    return E.reshape(shape), H.reshape(shape)


def _mirror_pair(mirrors, cnt=13):
    """(half-box faces with mirrors set, full-box faces without): the half
    box spans [0, L] on mirrored axes and leaves out the faces on the planes."""
    half_axes = [np.linspace(0, L, cnt) if a in mirrors else np.linspace(-L, L, 2 * cnt - 1) for a in range(3)]
    full_axes = [np.linspace(-L, L, 2 * cnt - 1) for _ in range(3)]
    imgs = _images(mirrors)

    def faces(axes, skip_planes):
        out = []
        for ny in range(3):
            for v in (axes[ny][0], axes[ny][-1]):
                if skip_planes and ny in mirrors and v == 0.0:
                    continue
                l = list(axes)
                l[ny] = np.array([v])
                out.append((l,) + _fields(l, imgs))
        return out
    return faces(half_axes, True), faces(full_axes, False)


def _run(nf, faces, mirrors=None):
    for a, k in (mirrors or {}).items():
        nf.set_mirror(k, a, 0.0)
    for l, E, H in faces:
        nf.add_plane(l, E, H)
    return nf


@pytest.mark.parametrize("case", sorted(MIRROR_CASES))
def test_oracle_mirror_planes_equal_full_box(case):
    mirrors = MIRROR_CASES[case]
    half, full = _mirror_pair(mirrors)
    theta = np.linspace(0, math.pi, 9)
    phi = np.linspace(0, 2 * math.pi, 6, endpoint=False)
    m = _run(OracleNF2FF(F, theta, phi), half, mirrors)
    f = _run(OracleNF2FF(F, theta, phi), full)
    s = np.abs(f.E_theta).max() + np.abs(f.E_phi).max()
    np.testing.assert_allclose(m.E_theta, f.E_theta, rtol=0, atol=1e-9 * s)
    np.testing.assert_allclose(m.E_phi, f.E_phi, rtol=0, atol=1e-9 * s)
    assert m.rad_power == pytest.approx(f.rad_power, rel=1e-9)
    # without the mirrors the half box is a different (open) surface
    h = _run(OracleNF2FF(F, theta, phi), half)
    assert np.abs(h.E_theta - f.E_theta).max() > 1e-2 * s


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(MIRROR_CASES))
def test_gpu_nf2ff_mirror_planes(gpu, case):
    from wipems_amd.nf2ff import NF2FF
    print("case: nf2ff mirrors %s" % case, flush=True)
    mirrors = MIRROR_CASES[case]
    half, _ = _mirror_pair(mirrors)
    theta = np.linspace(0, math.pi, 9)
    phi = np.linspace(0, 2 * math.pi, 6, endpoint=False)
    dev = _run(NF2FF(F, theta, phi), half, mirrors)
    orc = _run(OracleNF2FF(F, theta, phi), half, mirrors)
    _compare(dev, orc)
    dev.close()
