"""The large-dipole workload (the default benchmark: streamed operator
build -> uint16-index engine -> interior/shell UPML split -> twelve FP64 DFT
boxes -> optional FP16 storage) at a size the CPU oracle covers: 41 x 37 x 33
lines, 200 boxes, 4-cell UPML.

* FP32, compressed index from the streamed build: fields and the feed probe
  bit for bit equal to the oracle stepped on the expanded coefficient arrays
  (one UPML launch and the split launch), DFT accumulators of the nf2ff faces
  within 1e-11 of the box maximum against oracle/dft.py (FP64 on both sides;
  only the phase factors' last bits may differ);
* FP64 handle on the double full build: bit for bit equal to the double
  oracle;
* FP16 storage with pilot scales against FP32: probe and DFT rel-L2 bounded
  (the nf2ff faces of an unloaded dipole are FP16's hard case)."""
import numpy as np
import pytest

from oracle.dft import DftBox, run_with_boxes
from oracle.oracle import oracle_from_operator
from wipems_amd import scenarios
from wipems_amd.engine import engine_from_operator, pilot_scales

pytestmark = pytest.mark.gpu

TOL = 1e-11
STEPS = 400


def dipole(prec=0, streamed=True):
    return scenarios.large_dipole_scenario(num_lines=(41, 37, 33), n_boxes=200, steps=STEPS, pml=4, box_margin=3,
                                           dipole_cells=6, threads=4, prec=prec, streamed=streamed)


def oracle_boxes(d):
    return [DftBox(b.field, b.start, b.stop, b.freqs, b.fd_interval, d.edge_primary, d.edge_dual, d.num_lines)
            for b in d.dft_boxes]


@pytest.mark.parametrize("split", [0, 1])
def test_streamed_dipole_matches_oracle(gpu, split):
    sc = dipole()
    d = sc.data
    assert d.vv is None and d.index is not None and len(d.dft_boxes) == 12
    print("case: dipole 41x37x33 streamed, fp32, split %d" % split, flush=True)
    eng = engine_from_operator(d, precision="fp32")
    assert eng.compressed
    eng.set_pml_split(split)
    assert eng.pml_split() == bool(split)
    full = d.expanded()
    orc = oracle_from_operator(full, ftz=True)
    boxes = oracle_boxes(full)
    sampled0 = False
    for chunk in (33, 1, 166, 200):
        eng.IterateTS(chunk)
        sampled0 = run_with_boxes(orc, boxes, chunk, d.dT, sampled0)
    eng.synchronize()
    for w in (0, 1):
        np.testing.assert_array_equal(eng.get_field(w), orc.get_field(w))
    np.testing.assert_array_equal(eng.read_probe(0)[1], orc.read_probe(0)[1])
    assert np.abs(eng.read_probe(0)[1]).max() > 0
    for k, b in enumerate(boxes):
        g, o = eng.read_dft(k), b.acc
        scale = np.abs(o).max()
        assert scale > 0
        assert np.abs(g - o).max() <= TOL * scale, (k, np.abs(g - o).max(), scale)
    eng.close()


def test_dipole_fp64_matches_double_oracle(gpu):
    sc = dipole(prec=1, streamed=False)
    d = sc.data
    eng = engine_from_operator(d, precision="fp64", compressed=False)
    orc = oracle_from_operator(d)
    eng.IterateTS(STEPS)
    orc.iterate(STEPS)
    for w in (0, 1):
        np.testing.assert_array_equal(eng.get_field(w), orc.get_field(w))
    np.testing.assert_array_equal(eng.read_probe(0)[1], orc.read_probe(0)[1])


def test_dipole_fp16_storage_close_to_fp32(gpu):
    sc = dipole()
    d = sc.data
    e32 = engine_from_operator(d, precision="fp32")
    e32.IterateTS(STEPS)
    scales = pilot_scales(d, steps=max(64, d.excitation.length))
    e16 = engine_from_operator(d, precision="fp16", scales=scales)
    e16.IterateTS(STEPS)
    v32, v16 = e32.read_probe(0)[1], e16.read_probe(0)[1]
    rel = np.linalg.norm(v16 - v32) / np.linalg.norm(v32)
    worst = max(np.linalg.norm(e16.read_dft(k) - e32.read_dft(k)) / np.linalg.norm(e32.read_dft(k))
                for k in range(len(d.dft_boxes)))
    print("dipole fp16 vs fp32: feed probe rel-L2 %.3e, DFT faces max rel-L2 %.3e" % (rel, worst))
    assert rel < 5e-2 and worst < 1e-1
