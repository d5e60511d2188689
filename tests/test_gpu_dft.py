"""Frequency-domain field boxes (the six faces of an nf2ff box, E and H,
node interpolated) on the GPU against the CPU restatement in oracle/dft.py.

The fields themselves are bit-identical to the oracle's (FP32 and FP64);
interpolation and accumulation are FP64 on both sides with the same
operation order, and the phase factors come from the device's sincos and
numpy's cos/sin, which may differ in the last bit.  Accumulators are
therefore compared at max|gpu - oracle| <= 1e-11 * max|oracle| per box."""
import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.engine import engine_from_operator
from oracle.dft import DftBox, run_with_boxes
from oracle.oracle import oracle_from_operator

pytestmark = pytest.mark.gpu

TOL = 1e-11


def port_case(prec, n_cells=36, steps=700):
    sc = scenarios.graded_port_scenario(n_cells=n_cells, steps=steps, prec=prec, n_boxes=6, pml=6)
    d = sc.data
    lines = sc.op.lines
    lo = [lines[n][8] for n in range(3)]
    hi = [lines[n][-9] for n in range(3)]
    sc.op.nf2ff_box(d, lo, hi, [2e9, 5e9, 7.5e9])
    return sc


def oracle_boxes(d):
    return [DftBox(b.field, b.start, b.stop, b.freqs, b.fd_interval, d.edge_primary, d.edge_dual, d.num_lines)
            for b in d.dft_boxes]


@pytest.mark.parametrize("prec,precision", [(0, "fp32"), (1, "fp64")])
def test_nf2ff_box_matches_oracle(prec, precision):
    sc = port_case(prec)
    d = sc.data
    assert len(d.dft_boxes) == 12
    print("case: graded port 36^3 nf2ff box, %s, interval %d" % (precision, d.dft_boxes[0].fd_interval), flush=True)
    eng = engine_from_operator(d, precision=precision)
    orc = oracle_from_operator(d, ftz=True)
    boxes = oracle_boxes(d)
    done, sampled0 = 0, False
    for chunk in (97, 1, 203, 399):          # uneven chunks cross sample steps at every phase
        eng.IterateTS(chunk)
        sampled0 = run_with_boxes(orc, boxes, chunk, d.dT, sampled0)
        done += chunk
    eng.synchronize()
    for which in (0, 1):
        np.testing.assert_array_equal(eng.get_field(which), orc.get_field(which))
    for k, b in enumerate(boxes):
        assert eng.dft_num_samples(k) == b.samples == done // b.interval + 1
        g = eng.read_dft(k)
        o = b.acc
        scale = np.abs(o).max()
        assert scale > 0, "box %d recorded nothing" % k
        err = np.abs(g - o).max()
        assert err <= TOL * scale, (d.dft_boxes[k].name, err, scale)
    eng.close()


def test_dft_box_arguments_are_checked():
    sc = port_case(0, n_cells=20, steps=10)
    d = sc.data
    from wipems_amd._native import FdtdError
    from wipems_amd.operator import DftBoxDef
    d.dft_boxes = []
    eng = engine_from_operator(d, precision="fp32")
    with pytest.raises(FdtdError):      # edge lengths not set yet
        eng.add_dft_box(DftBoxDef(0, (1, 1, 1), (2, 2, 2), (1e9,), 5))
    eng.set_edge_lengths(d.edge_primary, d.edge_dual)
    with pytest.raises(FdtdError):      # outside the mesh
        eng.add_dft_box(DftBoxDef(0, (1, 1, 1), (2, 2, d.num_lines[2]), (1e9,), 5))
    with pytest.raises(FdtdError):      # zero interval
        eng.add_dft_box(DftBoxDef(1, (1, 1, 1), (2, 2, 2), (1e9,), 0))
    assert eng.add_dft_box(DftBoxDef(1, (1, 1, 1), (2, 2, 2), (1e9,), 5)) == 0
    eng.close()


def test_fp16_storage_dft_close_to_fp32():
    """FP16 storage with pilot scales: the box spectra stay within 2% of FP32's."""
    from wipems_amd.engine import pilot_scales
    sc = port_case(0)
    d = sc.data
    e32 = engine_from_operator(d, precision="fp32")
    e32.IterateTS(700)
    e32.synchronize()
    sc16 = port_case(0)
    scales = pilot_scales(sc16.data)
    e16 = engine_from_operator(sc16.data, precision="fp16", scales=scales)
    e16.IterateTS(700)
    e16.synchronize()
    for k in range(len(d.dft_boxes)):
        a, b = e32.read_dft(k), e16.read_dft(k)
        rel = np.linalg.norm(b - a) / np.linalg.norm(a)
        assert rel < 2e-2, (d.dft_boxes[k].name, rel)
    e32.close()
    e16.close()
