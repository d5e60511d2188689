"""SAR dumps (ProcessFieldsSAR, Common/processfields_sar.cpp:179-335, with
SAR_Calculation, tools/sar_calculation.cpp:162-230): local SAR and SAR raw
data of an E box with cell interpolation, and the cell materials they use.

The file is read back through the HDF5 C library (tests/h5read.py).  The
checks are closed forms and invariants, not a retyped formula: a uniform
field in a uniform material gives SAR = kappa |E|^2 / (2 rho) and a total
power of kappa |E|^2 / 2 times the box volume; cells without density hold 0;
SAR is quadratic in the field and blind to its phase; the cell materials are
the ones whose boxes hold the cell centres, by priority; the volume is the
product of the primary cell widths.  On the GPU, the SAR of the engine's
accumulators equals the SAR of the oracle's.
"""
import numpy as np
import pytest

from oracle.dft import float_ftz
from tests import h5read
from wipems_amd.dump import H5Dump
from wipems_amd.operator import BC_PEC, Operator


def _lines():
    rng = np.random.default_rng(8)
    return [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) for n in (5, 4, 6)]


def _write(tmp_path, name, acc, kappa, density, volume, freqs=(1e9,), raw=False):
    lines = _lines()
    path = tmp_path / name
    w = H5Dump(path, lines, 1e-3)
    w.write_sar(np.asarray(freqs, dtype=np.float64), acc, kappa, density, volume, raw=raw)
    w.close()
    return h5read.File(path)


def _shape():
    return tuple(len(l) for l in _lines())


def test_uniform_field_closed_form(tmp_path):
    shape = _shape()
    # |E|^2 = 3^2 + 4^2 + 12^2 = 169 V^2/m^2, split over components and phases
    e = np.array([3.0, 4.0j, 12.0 * np.exp(0.7j)])
    acc = np.broadcast_to(e[None, :, None, None, None], (2, 3) + shape).copy()
    acc[1] *= 2.0                                              # second frequency: twice the field
    kappa = np.full(shape, 0.5, np.float32)
    density = np.full(shape, 1000.0, np.float32)
    density[1, 2, 3] = 0.0                                     # no mass: SAR 0, still counted in the power
    volume = np.full(shape, 2.0e-9, np.float32)
    f = _write(tmp_path, "u.h5", acc, kappa, density, volume, freqs=(1e9, 2e9))
    np.testing.assert_array_equal(f.attribute("/FieldData/FD", "frequency"), [1e9, 2e9])
    for k, amp in ((0, 1.0), (1, 2.0)):
        d = f.dataset("/FieldData/FD/f%d" % k)
        assert d.shape == shape[::-1]                          # [nz][ny][nx]
        sar = 0.5 * 0.5 * 169.0 * amp ** 2 / 1000.0
        want = np.full(shape[::-1], sar, np.float32)
        want[3, 2, 1] = 0.0
        np.testing.assert_allclose(d, want, rtol=1e-6)
        assert d[3, 2, 1] == 0.0
        power = 0.5 * 0.5 * 169.0 * amp ** 2 * 2.0e-9 * np.prod(shape)
        assert f.attribute("/FieldData/FD/f%d" % k, "power", np.float32)[0] == pytest.approx(power, rel=1e-6)
        assert f.attribute("/FieldData/FD/f%d" % k, "frequency", np.float32)[0] == np.float32((1e9, 2e9)[k])
    assert not f.exists("/CellData")
    f.close()


def test_sar_is_quadratic_phase_blind_and_placed(tmp_path):
    shape = _shape()
    rng = np.random.default_rng(9)
    acc = rng.standard_normal((1, 3) + shape) + 1j * rng.standard_normal((1, 3) + shape)
    kappa = rng.uniform(0.0, 2.0, shape).astype(np.float32)
    kappa[0] = 0.0                                             # lossless cells: no power
    density = rng.uniform(500, 2000, shape).astype(np.float32)
    density[:, 1] = 0.0
    volume = rng.uniform(1e-9, 3e-9, shape).astype(np.float32)
    f = _write(tmp_path, "a.h5", acc, kappa, density, volume)
    s1, p1 = f.dataset("/FieldData/FD/f0"), f.attribute("/FieldData/FD/f0", "power", np.float32)[0]
    f.close()
    f = _write(tmp_path, "b.h5", 2.0 * acc, kappa, density, volume)          # power of two: exact scaling
    np.testing.assert_array_equal(f.dataset("/FieldData/FD/f0"), float_ftz(4.0 * s1.astype(np.float64)))
    assert f.attribute("/FieldData/FD/f0", "power", np.float32)[0] == np.float32(4.0) * p1
    f.close()
    f = _write(tmp_path, "c.h5", acc * np.exp(1.1j), kappa, density, volume)
    np.testing.assert_allclose(f.dataset("/FieldData/FD/f0"), s1, rtol=1e-6)
    f.close()
    assert not np.any(s1[:, 1, :]) and not np.any(s1[:, :, 0])
    # one hot cell: its SAR lands at [z][y][x] and the power is that cell's alone
    hot = np.zeros_like(acc)
    hot[0, 1, 3, 2, 4] = 5.0
    f = _write(tmp_path, "d.h5", hot, kappa, density, volume)
    d = f.dataset("/FieldData/FD/f0")
    assert np.count_nonzero(d) == 1
    assert d[4, 2, 3] == pytest.approx(0.5 * kappa[3, 2, 4] * 25.0 / density[3, 2, 4], rel=1e-6)
    assert f.attribute("/FieldData/FD/f0", "power", np.float32)[0] == pytest.approx(
        0.5 * kappa[3, 2, 4] * 25.0 * volume[3, 2, 4], rel=1e-6)
    f.close()


def test_sar_raw_data_layout(tmp_path):
    shape = _shape()
    rng = np.random.default_rng(10)
    acc = rng.standard_normal((2, 3) + shape) + 1j * rng.standard_normal((2, 3) + shape)
    kappa, density, volume = (rng.uniform(0.1, 1.0, shape).astype(np.float32) for _ in range(3))
    f = _write(tmp_path, "raw.h5", acc, kappa, density, volume, freqs=(1e9, 3e9), raw=True)
    np.testing.assert_array_equal(f.attribute("/FieldData/FD", "frequency"), [1e9, 3e9])
    for k in range(2):
        assert not f.exists("/FieldData/FD/f%d" % k)
        for part, fn in (("real", np.real), ("imag", np.imag)):
            d = f.dataset("/FieldData/FD/f%d_%s" % (k, part))
            assert d.shape == (3,) + shape[::-1]
            np.testing.assert_array_equal(d, float_ftz(np.transpose(fn(acc[k]), (0, 3, 2, 1))))
    for name, src in (("Conductivity", kappa), ("Density", density), ("Volume", volume)):
        np.testing.assert_array_equal(f.dataset("/CellData/" + name), src.transpose(2, 1, 0))
    f.close()


def _material_op():
    lines = _lines()
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    tissue = op.add_material(eps_r=40.0, kappa=0.8)
    bone = op.add_material(eps_r=12.0, kappa=0.1)
    op.set_density(tissue, 1050.0)
    op.set_density(bone, 1900.0)
    # box corners on mesh lines: every cell centre is strictly inside or outside
    op.add_box(tissue, (lines[0][1], lines[1][0], lines[2][1]), (lines[0][4], lines[1][2], lines[2][5]), priority=1)
    op.add_box(bone, (lines[0][2], lines[1][1], lines[2][2]), (lines[0][3], lines[1][3], lines[2][4]), priority=5)
    return op, lines


def test_cell_materials_by_centre_and_priority():
    op, lines = _material_op()
    n = [len(l) for l in lines]
    kappa, density, volume = op.cell_materials((0, 0, 0), tuple(k - 1 for k in n))
    assert kappa.shape == tuple(n)
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                if 2 <= i < 3 and 1 <= j < 3 and 2 <= k < 4:
                    want = (0.1, 1900.0)
                elif 1 <= i < 4 and 0 <= j < 2 and 1 <= k < 5:
                    want = (0.8, 1050.0)
                else:
                    want = (0.0, 0.0)
                assert (kappa[i, j, k], density[i, j, k]) == (np.float32(want[0]), np.float32(want[1])), (i, j, k)
                w = [lines[a][p + 1] - lines[a][p] if p + 1 < n[a] else lines[a][p] - lines[a][p - 1]
                     for a, p in enumerate((i, j, k))]
                assert volume[i, j, k] == pytest.approx(np.prod(w) * 1e-9, rel=1e-6)
    # a sub-box gives the same cells
    sub = op.cell_materials((1, 1, 2), (3, 2, 4))
    for full, part in zip((kappa, density, volume), sub):
        np.testing.assert_array_equal(part, full[1:4, 1:3, 2:5])


def test_cell_material_arguments_are_checked():
    from wipems_amd._native import FdtdError
    op, lines = _material_op()
    exc = op.add_excitation(0, (0.0, 0.0, 1.0))
    with pytest.raises(FdtdError):
        op.set_density(exc, 10.0)
    with pytest.raises(FdtdError):
        op.set_density(0, -1.0)
    with pytest.raises(FdtdError):
        op.cell_materials((0, 0, 0), (len(lines[0]), 1, 1))
    with pytest.raises(ValueError):
        op.cell_materials((2, 0, 0), (1, 1, 1))


@pytest.mark.gpu
def test_gpu_sar_dump_matches_oracle(gpu, tmp_path):
    """A lossy, dense box excited by a dipole: the SAR file written from the
    engine's E box (cell interpolation) equals, to the FD accumulators'
    tolerance, the SAR of the oracle's accumulators of the same box."""
    from oracle.dft import DftBox
    from oracle.oracle import oracle_from_operator
    from wipems_amd.dump import write_sar_dump
    from wipems_amd.engine import engine_from_operator
    from wipems_amd.operator import DftBoxDef
    rng = np.random.default_rng(12)
    lines = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.4, k - 1))]) for k in (16, 14, 12)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_gauss_excite(6e9, 5e9)
    m = op.add_material(eps_r=20.0, kappa=0.6)
    op.set_density(m, 1100.0)
    op.add_box(m, (lines[0][2], lines[1][2], lines[2][2]), (lines[0][9], lines[1][8], lines[2][9]), priority=1)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (lines[0][11], lines[1][6], lines[2][4]), (lines[0][11], lines[1][6], lines[2][6]), priority=10)
    steps = 400                                                # past the pulse's peak
    d = op.build(0, nr_ts=steps)
    freqs = (4e9, 7e9)
    box = DftBoxDef(0, (1, 1, 1), (13, 11, 10), freqs, 2, "sar", interp=2)
    d.dft_boxes = [box]
    print("case: SAR dump, %s, %d steps" % (d.num_lines, steps), flush=True)
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.IterateTS(steps)
    eng.synchronize()
    path = tmp_path / "sar.h5"
    write_sar_dump(path, eng, 0, op)
    write_sar_dump(tmp_path / "raw.h5", eng, 0, op, raw=True)
    o = oracle_from_operator(d, ftz=True)
    ref = DftBox(0, box.start, box.stop, freqs, 2, d.edge_primary, d.edge_dual, d.num_lines, interp=2, mats=d.materials)
    for ts in range(steps + 1):
        if ts:
            o.iterate(1)
        if ts % 2 == 0:
            ref.accumulate_dump(o.get_field(0), o.get_field(1), ts, d.dT)
    got = eng.read_dft(0)
    assert np.abs(got - ref.acc).max() <= 1e-11 * np.abs(ref.acc).max()
    kappa, density, volume = op.cell_materials(box.start, box.stop)
    assert kappa.max() == np.float32(0.6) and density.max() == np.float32(1100.0) and kappa.min() == 0.0
    f = h5read.File(path)
    for k in range(2):
        p_cell = 0.5 * kappa * (np.abs(ref.acc[k]) ** 2).sum(axis=0)
        sar = np.where(density > 0, p_cell / np.where(density > 0, density, 1), 0.0)
        d_sar = f.dataset("/FieldData/FD/f%d" % k)
        assert d_sar.shape == tuple(kappa.shape[::-1])
        np.testing.assert_allclose(d_sar, float_ftz(sar.transpose(2, 1, 0)), rtol=1e-6, atol=1e-6 * sar.max())
        assert sar.max() > 1e-30
        assert f.attribute("/FieldData/FD/f%d" % k, "power", np.float32)[0] == pytest.approx((p_cell * volume).sum(), rel=1e-6)
    f.close()
    f = h5read.File(tmp_path / "raw.h5")
    np.testing.assert_array_equal(f.dataset("/CellData/Density"), density.transpose(2, 1, 0))
    np.testing.assert_allclose(f.dataset("/FieldData/FD/f1_real"), np.real(ref.acc[1]).transpose(0, 3, 2, 1),
                               rtol=1e-6, atol=1e-9 * np.abs(ref.acc).max())
    f.close()
    eng.close()
