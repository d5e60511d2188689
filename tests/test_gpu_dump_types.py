"""Every dump type (E, H, J, rot H, D, B) with every interpolation (none,
node, cell) on the GPU, time domain and frequency domain, against the
node-by-node restatement in oracle/dft.py evaluated on the oracle's fields.

The scenario holds a lossy dielectric, magnetic box, so the stored epsilon,
kappa and mu differ between edges.  The fields are bit-identical to the
oracle's; the dump values are FP64 on both sides with the same operation
order, so the float TD samples must match bit for bit (rounded with flush to zero,
as the reference's main thread rounds them).  The FD accumulators
use the device's sincos and numpy's cos/sin for the phase, which may differ
in the last bit: they are compared at max|gpu - oracle| <= 1e-11 * max|oracle|."""
import numpy as np
import pytest

from oracle.dft import DftBox, dump_box, dump_time, float_ftz
from oracle.oracle import oracle_from_operator
from wipems_amd.engine import engine_from_operator
from wipems_amd.operator import BC_PEC, DftBoxDef, Operator

pytestmark = pytest.mark.gpu

TOL = 1e-11
STEPS = 24


def material_case(prec=0):
    rng = np.random.default_rng(41)
    lines = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.5, n - 1))]) for n in (14, 12, 11)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_gauss_excite(5e9, 5e9)
    m = op.add_material(eps_r=3.5, kappa=0.02, mue_r=1.8)
    op.add_box(m, (lines[0][3], lines[1][2], lines[2][3]), (lines[0][9], lines[1][8], lines[2][7]), priority=1)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (lines[0][6], lines[1][5], lines[2][4]), (lines[0][6], lines[1][5], lines[2][6]), priority=10)
    for k in range(3):
        op.set_material_store_flags(k)
    d = op.build(prec, nr_ts=STEPS)
    return op, d


MATS = {2: 1, 4: 0, 5: 2}     # dump type -> the material it needs


def test_td_dump_types_match_oracle():
    op, d = material_case()
    N = d.num_lines
    print("case: TD dump types x interpolations, %s, %d steps" % (N, STEPS), flush=True)
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.set_edge_lengths(d.edge_primary, d.edge_dual)
    boxes = []
    captured = {}
    a_full, b_full = (0, 0, 0), tuple(k - 1 for k in N)
    for t in range(6):
        for interp in range(3):
            a, b = (a_full, b_full) if (t + interp) % 2 == 0 else ((2, 1, 3), (9, 10, 8))
            interval = 3 + (t + interp) % 4
            shape = (3,) + tuple(b[n] - a[n] + 1 for n in range(3))
            key = (t, interp)
            captured[key] = {}

            def keep(ts, tt, ptr, n, key=key, shape=shape):
                captured[key][ts] = (tt, np.ctypeslib.as_array(ptr, shape=(int(n),)).reshape(shape).copy())
            eng.add_td_box(t, a, b, interval, keep, interp=interp)
            boxes.append((key, a, b, interval))
    for c in (1, 10, 13):
        eng.IterateTS(c)
    o = oracle_from_operator(d, ftz=True)
    seen_nonzero = set()
    for ts in range(STEPS + 1):
        if ts:
            o.iterate(1)
        V, I = o.get_field(0), o.get_field(1)
        for key, a, b, interval in boxes:
            if ts % interval:
                continue
            t, interp = key
            ref = float_ftz(dump_box(V, I, d.edge_primary, d.edge_dual, d.materials, t, interp, a, b))
            tt, got = captured[key][ts]
            np.testing.assert_array_equal(got, ref, err_msg="dump type %d interp %d ts %d" % (t, interp, ts))
            assert tt == dump_time(t, ts, d.dT)
            if np.abs(ref).max() > 0:
                seen_nonzero.add(key)
    for key, a, b, interval in boxes:
        assert sorted(captured[key]) == list(range(0, STEPS + 1, interval))
    assert len(seen_nonzero) == 18, sorted(set(captured) - seen_nonzero)
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp16", "bf16"])
def test_td_dump_types_other_storage(precision):
    """FP64 and FP16/BF16 storage: every dump type and interpolation equals
    the restatement evaluated on the engine's own fields at the sample step
    (fdtd_get_field gives the stored values, scaled, in FP64)."""
    from wipems_amd.engine import pilot_scales
    op, d = material_case(1 if precision == "fp64" else 0)
    N = d.num_lines
    print("case: TD dump types, %s storage" % precision, flush=True)
    scales = pilot_scales(d, steps=16) if precision in ("fp16", "bf16") else None
    eng = engine_from_operator(d, precision=precision, compressed=False, scales=scales)
    eng.set_edge_lengths(d.edge_primary, d.edge_dual)
    a, b = (0, 0, 0), tuple(k - 1 for k in N)
    captured = {}
    for t in range(6):
        for interp in range(3):
            captured[(t, interp)] = {}

            def keep(ts, tt, ptr, n, key=(t, interp)):
                captured[key][ts] = np.ctypeslib.as_array(ptr, shape=(int(n),)).reshape((3,) + N).copy()
            eng.add_td_box(t, a, b, 4, keep, interp=interp)
    fields = {}
    for ts in range(0, STEPS + 1, 4):
        if ts:
            eng.IterateTS(4)
        else:
            eng.IterateTS(0)
        fields[ts] = (eng.get_field(0), eng.get_field(1))
    eng.IterateTS(0)
    nonzero = 0
    for (t, interp), got in captured.items():
        assert sorted(got) == list(range(0, STEPS + 1, 4))
        for ts, arr in got.items():
            V, I = fields[ts]
            ref = float_ftz(dump_box(V, I, d.edge_primary, d.edge_dual, d.materials, t, interp, a, b))
            np.testing.assert_array_equal(arr, ref, err_msg="%s type %d interp %d ts %d" % (precision, t, interp, ts))
            nonzero += bool(np.abs(ref).max() > 0)
    assert nonzero > 50
    eng.close()


def test_fd_dump_types_match_oracle():
    op, d = material_case()
    N = d.num_lines
    freqs = (3e9, 7e9)
    defs = [DftBoxDef(2, (1, 1, 1), (10, 9, 8), freqs, 3, "J cell", interp=2),
            DftBoxDef(3, (0, 0, 0), tuple(k - 1 for k in N), freqs, 4, "rotH none", interp=0),
            DftBoxDef(4, (2, 0, 1), (8, 11, 9), freqs, 5, "D node", interp=1),
            DftBoxDef(5, (0, 2, 2), (13, 8, 6), freqs, 3, "B cell", interp=2),
            DftBoxDef(1, (1, 1, 1), (12, 10, 9), freqs, 2, "H none", interp=0)]
    d.dft_boxes = list(defs)
    print("case: FD dump types, %s, %d steps" % (N, STEPS), flush=True)
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.IterateTS(STEPS)
    eng.synchronize()
    o = oracle_from_operator(d, ftz=True)
    boxes = [DftBox(b.field, b.start, b.stop, b.freqs, b.fd_interval, d.edge_primary, d.edge_dual, N, interp=b.interp,
                    mats=d.materials) for b in defs]
    for ts in range(STEPS + 1):
        if ts:
            o.iterate(1)
        for b in boxes:
            if ts % b.interval == 0:
                b.accumulate_dump(o.get_field(0), o.get_field(1), ts, d.dT)
    for which in (0, 1):
        np.testing.assert_array_equal(eng.get_field(which), o.get_field(which))
    for k, b in enumerate(boxes):
        assert eng.dft_num_samples(k) == b.samples
        g = eng.read_dft(k)
        scale = np.abs(b.acc).max()
        assert scale > 0, defs[k].name
        assert np.abs(g - b.acc).max() <= TOL * scale, defs[k].name
    eng.close()


def test_dump_type_arguments_are_checked():
    from wipems_amd._native import FdtdError
    op, d = material_case()
    d.materials = {}
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.set_edge_lengths(d.edge_primary, d.edge_dual)
    sink = lambda *args: None
    for t in (2, 4, 5):                     # J, D and B need their material first
        with pytest.raises(FdtdError):
            eng.add_td_box(t, (0, 0, 0), (2, 2, 2), 1, sink, interp=1)
    with pytest.raises(FdtdError):
        eng.add_td_box(6, (0, 0, 0), (2, 2, 2), 1, sink)
    with pytest.raises(FdtdError):
        eng.add_td_box(0, (0, 0, 0), (2, 2, 2), 1, sink, interp=3)
    with pytest.raises(FdtdError):
        eng.add_dft_box(DftBoxDef(5, (0, 0, 0), (2, 2, 2), (1e9,), 2, interp=1))
    with pytest.raises(ValueError):
        eng.set_material(0, np.zeros((3, 2, 2, 2), np.float32))
    op2, d2 = material_case()
    eng.set_material(1, d2.materials[1])
    assert eng.add_td_box(2, (0, 0, 0), (2, 2, 2), 1, sink, interp=2) == 0
    assert eng.add_td_box(3, (0, 0, 0), (2, 2, 2), 1, sink, interp=0) == 1
    eng.close()
