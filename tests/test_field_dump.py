"""HDF5 field dumps (ProcessFieldsTD / ProcessFieldsFD with HDF5_File_Writer).

The file layout is read back through the HDF5 C library directly
(tests/h5read.py), not through the writer: mesh datasets scaled in float,
the version attribute, /FieldData/TD/<8-digit timestep> of shape
[3][nz][ny][nx] with a float time attribute, /FieldData/FD with the
frequency list and f<k>_real / f<k>_imag.  On the GPU, every time sample
equals the oracle's node interpolation (oracle/dft.py) of the oracle's
fields at that step, rounded to float, and an FD dump file holds the field
box's accumulators.
"""
import numpy as np
import pytest

from oracle.dft import float_ftz
from tests import h5read
from wipems_amd.dump import H5Dump


def _lines():
    rng = np.random.default_rng(5)
    return [np.cumsum(rng.uniform(0.3, 1.7, n)) for n in (4, 5, 3)]


def test_writer_layout(tmp_path):
    lines = _lines()
    unit = 1e-3
    path = tmp_path / "box.h5"
    w = H5Dump(path, lines, unit)
    rng = np.random.default_rng(6)
    f1 = rng.standard_normal((3, 4, 5, 3)).astype(np.float32)
    f2 = rng.standard_normal((3, 4, 5, 3)).astype(np.float32)
    w.write_td(0, 0.0, f1)
    w.write_td(1234, 3.25e-9, f2)
    freqs = np.array([1e9, 2.5e9])
    acc = rng.standard_normal((2, 3, 4, 5, 3)) + 1j * rng.standard_normal((2, 3, 4, 5, 3))
    acc[0, 1, 2, 3, 1] = 3e-40 - 2e-39j          # float denormals: flushed to zero, as the reference stores them
    w.write_fd(freqs, acc)
    w.close()

    f = h5read.File(path)
    for n, name in enumerate("xyz"):
        got = f.dataset("/Mesh/" + name)
        # WriteRectMesh: float(line) * float(scaling) in float
        np.testing.assert_array_equal(got, lines[n].astype(np.float32) * np.float32(unit))
    assert f.attribute("/", "openEMS_HDF5_version")[0] == 0.2
    assert f.exists("/FieldData/TD/00000000") and f.exists("/FieldData/TD/00001234")
    assert not f.exists("/FieldData/TD/1234")
    for name, src, t in (("00000000", f1, 0.0), ("00001234", f2, 3.25e-9)):
        d = f.dataset("/FieldData/TD/" + name)
        assert d.shape == (3, 3, 5, 4)                              # [3][nz][ny][nx]
        np.testing.assert_array_equal(d, np.transpose(src, (0, 3, 2, 1)))
        assert f.attribute("/FieldData/TD/" + name, "time", np.float32)[0] == np.float32(t)
    np.testing.assert_array_equal(f.attribute("/FieldData/FD", "frequency"), freqs)
    for k in range(2):
        for part, fn in (("real", np.real), ("imag", np.imag)):
            d = f.dataset("/FieldData/FD/f%d_%s" % (k, part))
            np.testing.assert_array_equal(d, float_ftz(np.transpose(fn(acc[k]), (0, 3, 2, 1))))
            assert f.attribute("/FieldData/FD/f%d_%s" % (k, part), "frequency", np.float32)[0] == np.float32(freqs[k])
    f.close()


def test_writer_rejects_bad_path(tmp_path):
    from wipems_amd._native import FdtdError
    with pytest.raises(FdtdError):
        H5Dump(tmp_path / "missing_dir" / "x.h5", _lines(), 1e-3)


# ---------------------------------------------------------------------------
# device engine
# ---------------------------------------------------------------------------
def _case(prec=0):
    from wipems_amd import scenarios
    sc = scenarios.graded_port_scenario(n_cells=22, steps=60, prec=prec, n_boxes=3, pml=4)
    return sc.op, sc.data


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [True, False])
def test_gpu_td_dump_matches_oracle_interpolation(gpu, tmp_path, use_graph):
    from oracle.dft import float_ftz, interp_e, interp_h
    from oracle.oracle import oracle_from_operator
    from wipems_amd.engine import engine_from_operator
    print("case: TD dump vs oracle interpolation, graph %s" % use_graph, flush=True)
    op, d = _case()
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.set_use_graph(use_graph)
    eng.set_edge_lengths(d.edge_primary, d.edge_dual)
    N = d.num_lines
    boxes = [(0, (3, 2, 4), (15, 12, 9), 7), (1, (0, 5, 0), (N[0] - 1, 5, N[2] - 1), 5)]
    captured = {0: {}, 1: {}}
    ids = []
    for field, a, b, interval in boxes:
        shape = (3,) + tuple(b[n] - a[n] + 1 for n in range(3))

        def keep(ts, t, ptr, n, field=field, shape=shape):
            arr = np.ctypeslib.as_array(ptr, shape=(int(n),)).reshape(shape).copy()
            captured[field][ts] = (t, arr)
        ids.append(eng.add_td_box(field, a, b, interval, keep))
    path = tmp_path / "e.h5"
    dump = eng.add_td_dump(str(path), 0, boxes[0][1], boxes[0][2], boxes[0][3], op.lines, op.unit)
    for c in (1, 2, 29, 1, 27):
        eng.iterate(c)
    assert sorted(captured[0]) == list(range(0, 61, 7))
    assert sorted(captured[1]) == list(range(0, 61, 5))
    assert eng.td_num_samples(ids[0]) == 9 and dump.samples == 9

    o = oracle_from_operator(d, ftz=True)
    lens_e = [d.edge_primary[sum(N[:n]):sum(N[:n + 1])] for n in range(3)]
    lens_h = [d.edge_dual[sum(N[:n]):sum(N[:n + 1])] for n in range(3)]
    for ts in range(0, 61):
        if ts:
            o.iterate(1)
        for field, a, b, interval in boxes:
            if ts % interval:
                continue
            F = o.get_field(field)
            ref = float_ftz(interp_e(F, lens_e, a, b) if field == 0 else interp_h(F, lens_h, a, b))
            t, got = captured[field][ts]
            np.testing.assert_array_equal(got, ref)
            assert t == (ts + 0.5 * field) * d.dT
    assert np.abs(captured[0][56][1]).max() > 0
    eng.close()
    f = h5read.File(path)
    for ts in range(0, 61, 7):
        name = "/FieldData/TD/%08d" % ts
        np.testing.assert_array_equal(f.dataset(name), np.transpose(captured[0][ts][1], (0, 3, 2, 1)))
        assert f.attribute(name, "time", np.float32)[0] == np.float32(ts * d.dT)
    np.testing.assert_array_equal(f.dataset("/Mesh/x"), op.lines[0][3:16].astype(np.float32) * np.float32(op.unit))
    f.close()


@pytest.mark.gpu
def test_gpu_fd_dump_file(gpu, tmp_path):
    from oracle.dft import float_ftz
    from wipems_amd.dump import write_fd_dump
    from wipems_amd.engine import engine_from_operator
    from wipems_amd.operator import DftBoxDef
    print("case: FD dump file", flush=True)
    op, d = _case()
    d.dft_boxes = [DftBoxDef(0, (2, 3, 1), (12, 9, 8), (2e9, 5e9, 7.5e9), 4, "e")]
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.iterate(60)
    acc = eng.read_dft(0)
    path = tmp_path / "fd.h5"
    write_fd_dump(str(path), eng, 0, op.lines, op.unit)
    f = h5read.File(path)
    np.testing.assert_array_equal(f.attribute("/FieldData/FD", "frequency"), [2e9, 5e9, 7.5e9])
    for k in range(3):
        re = f.dataset("/FieldData/FD/f%d_real" % k)
        im = f.dataset("/FieldData/FD/f%d_imag" % k)
        np.testing.assert_array_equal(re, float_ftz(np.transpose(acc[k].real, (0, 3, 2, 1))))
        np.testing.assert_array_equal(im, float_ftz(np.transpose(acc[k].imag, (0, 3, 2, 1))))
    assert np.abs(acc).max() > 0
    f.close()


@pytest.mark.gpu
def test_gpu_td_box_rejections(gpu):
    from wipems_amd._native import FdtdError
    from wipems_amd.engine import Batch, Engine
    eng = Engine((8, 8, 8), 1e-12, "fp32")
    cb = lambda *a: None  # noqa: E731
    with pytest.raises(FdtdError, match="set_edge_lengths"):
        eng.add_td_box(0, (0, 0, 0), (3, 3, 3), 2, cb)
    e = np.ones(24)
    eng.set_edge_lengths(e, e)
    with pytest.raises(FdtdError):
        eng.add_td_box(0, (0, 0, 0), (8, 3, 3), 2, cb)
    with pytest.raises(FdtdError):
        eng.add_td_box(0, (0, 0, 0), (3, 3, 3), 0, cb)
    eng.add_td_box(1, (0, 0, 0), (3, 3, 3), 2, cb)
    z = np.zeros((3, 8, 8, 8), np.float32)
    eng.set_coeffs(z, z, z, z)
    with pytest.raises(FdtdError):
        Batch([eng])


@pytest.mark.gpu
def test_gpu_td_sink_error_surfaces(gpu):
    from wipems_amd.engine import Engine
    eng = Engine((6, 6, 6), 1e-12, "fp32")
    e = np.ones(18)
    eng.set_edge_lengths(e, e)
    z = np.zeros((3, 6, 6, 6), np.float32)
    eng.set_coeffs(z, z, z, z)

    def bad(*a):
        raise ValueError("sink failed")
    eng.add_td_box(0, (0, 0, 0), (2, 2, 2), 3, bad)
    with pytest.raises(ValueError, match="sink failed"):
        eng.iterate(4)


@pytest.mark.gpu
def test_gpu_iterate_splits_combined(gpu):
    """Steady-state evaluations (a step launched on its own every period),
    frequency-domain samples and time-domain dump samples on one run, with
    chunks that straddle all three: fields bitwise, the steady-state value,
    FD accumulators and TD samples as the oracle gives them."""
    from oracle.dft import DftBox, float_ftz, interp_e, run_with_boxes
    from oracle.oracle import oracle_from_operator
    from tests.test_steady_state import line_op
    from wipems_amd.engine import engine_from_operator
    from wipems_amd.operator import DftBoxDef
    print("case: steady state + FD + TD boxes on one run", flush=True)
    op = line_op(5e9, nx=60)
    d = op.build(0)
    d.dft_boxes = [DftBoxDef(0, (10, 0, 0), (40, 4, 4), (5e9,), 3, "e")]
    eng = engine_from_operator(d, precision="fp32")
    period = eng.steady_state_period()
    td = {}

    def keep(ts, t, ptr, n):
        td[ts] = np.ctypeslib.as_array(ptr, shape=(int(n),)).reshape((3, 21, 5, 5)).copy()
    eng.add_td_box(0, (20, 0, 0), (40, 4, 4), 7, keep)
    o = oracle_from_operator(d)
    o.set_steady_state(d.steady_state)
    boxes = [DftBox(0, (10, 0, 0), (40, 4, 4), (5e9,), 3, d.edge_primary, d.edge_dual, d.num_lines)]
    N = d.num_lines
    lens = [d.edge_primary[sum(N[:n]):sum(N[:n + 1])] for n in range(3)]
    sampled0 = False
    ts = 0
    for c in (5, period - 3, 1, 2 * period + 7, 16, 33, period):
        eng.iterate(c)
        for _ in range(c):
            run_with_boxes(o, boxes, 1, d.dT, sampled0)
            sampled0 = True
            ts += 1
            if ts % 7 == 0:
                ref = float_ftz(interp_e(o.get_field(0), lens, (20, 0, 0), (40, 4, 4)))
                np.testing.assert_array_equal(td[ts], ref)
        assert eng.GetLastDiff() == pytest.approx(o.GetLastDiff(), rel=1e-9, abs=2e-12)
    for which in (0, 1):
        np.testing.assert_array_equal(eng.get_field(which), o.get_field(which))
    acc = eng.read_dft(0)
    np.testing.assert_allclose(acc, boxes[0].acc, rtol=0, atol=1e-11 * np.abs(boxes[0].acc).max())
    assert sorted(td) == list(range(0, ts + 1, 7))
