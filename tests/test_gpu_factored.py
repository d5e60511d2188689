"""Factored coefficients on the GPU (fdtd_set_coeffs_factored).

An FP32 handle with factored coefficients must equal, bit for bit, the
oracle run on the coefficients the factored form stands for (vi, iv rebuilt
in float32 in the engine's product order): that pins the device's index,
table and mesh-factor reads exactly.  FP16 handles use the factored form on
graded meshes; their S11 error against FP64 is bounded in
test_gpu_precision.py."""
import copy

import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd._native import FdtdError
from wipems_amd.engine import Batch, Engine, engine_from_operator
from oracle.oracle import oracle_from_operator
from tests.test_factored import rebuild
from tests.test_gpu_parity import compare, _kinds

pytestmark = pytest.mark.gpu


def factored_pair(sc, steps, split=-1, chunks=(1, 5)):
    d = sc.data
    fac = d.factor(16)
    assert fac is not None
    eng = engine_from_operator(d, precision="fp32", factored=True)
    assert eng.factored
    eng.set_pml_split(split)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    rd = copy.copy(d)
    rd.vi, rd.iv = rebuild(fac, d.num_lines)
    orc = oracle_from_operator(rd, ftz=True)
    done = 0
    for c in chunks + (steps - sum(chunks),):
        eng.iterate(c)
        orc.iterate(c)
        done += c
    eng.synchronize()
    return eng, orc


@pytest.mark.parametrize("split", [0, 1])
def test_graded_port_fp32_factored_bitwise(gpu, split):
    sc = scenarios.graded_port_scenario(n_cells=40, steps=300, prec=0, n_boxes=6, pml=6)
    print("case: graded port 40^3 factored FP32, split %d" % split, flush=True)
    eng, orc = factored_pair(sc, 300, split)
    compare(eng, orc)
    eng.close()


def test_dispersive_fp32_factored_bitwise(gpu):
    sc = scenarios.dispersive_scenario(num_lines=(49, 49, 41), steps=200, material_scale=1e-3, ts_factor=0.9)
    print("case: dispersive 49x49x41 factored FP32", flush=True)
    eng, orc = factored_pair(sc, 200)
    compare(eng, orc)
    eng.close()


def test_fp16_picks_factored_on_graded_mesh(gpu):
    """49^3 graded lines: 110593 distinct plain tuples, too many for a uint16
    index, so FP16 handles take the factored form unless told otherwise."""
    sc = scenarios.graded_port_scenario(n_cells=48, steps=10, prec=0, n_boxes=6, pml=6)
    e16 = engine_from_operator(sc.data, precision="fp16", scales=([1.0] * 3, [1.0] * 3))
    assert e16.factored and not e16.compressed
    e16.close()
    e16 = engine_from_operator(sc.data, precision="fp16", scales=([1.0] * 3, [1.0] * 3), factored=False)
    assert not e16.factored
    e16.close()


def test_factored_rejections(gpu):
    sc = scenarios.graded_port_scenario(n_cells=20, steps=10, prec=0, n_boxes=4, pml=4)
    d = sc.data
    with pytest.raises(FdtdError, match="float arithmetic"):
        engine_from_operator(d, precision="fp64", factored=True)
    fac = d.factor(16)
    eng = Engine(d.num_lines, d.dT, "fp32")
    bad = copy.copy(fac)
    bad.index = fac.index.copy()
    bad.index[3, 4, 5] = fac.table.shape[0]
    with pytest.raises(FdtdError, match="n_entries"):
        eng.set_coeffs_factored(bad)
    bad = copy.copy(fac)
    bad.fv = fac.fv.copy()
    bad.fv[1, 7] = np.inf
    with pytest.raises(FdtdError, match="not finite"):
        eng.set_coeffs_factored(bad)
    eng.close()
    a = engine_from_operator(d, precision="fp32", factored=True)
    with pytest.raises(FdtdError, match="factored"):
        Batch([a])
    a.close()
