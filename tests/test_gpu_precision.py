"""Reduced-precision field storage (FP16/BF16 with per-component power-of-two
scales, FP32 update arithmetic) against the FP64 engine.

The FP16 handle's first step is checked against a float emulation of the same
storage rounding computed on the CPU (numpy), and long runs are checked for
error against FP64 on the lumped-port S11 (python/openEMS/ports.py:124-127)."""
import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.engine import engine_from_operator, pilot_scales

pytestmark = pytest.mark.gpu


def s11_run(sc, precision, steps, scales=None, freqs=None):
    d = sc.data
    eng = engine_from_operator(d, precision=precision, compressed=False, scales=scales)
    eng.iterate(steps)
    tu, u = eng.read_probe(0)
    ti, i = eng.read_probe(1)
    return scenarios.s11_from_probes(tu, u, ti, i, freqs), u, i


def test_fp16_storage_rounding_first_steps(gpu):
    """Storage rounding of one update is the FP32 result rounded to FP16 after
    scaling: after 1 step every stored value equals float16(fp32_value/scale)."""
    sc = scenarios.cavity_point_source(16, 10, prec=0)
    d = sc.data
    sv, si = [2.0 ** -10] * 3, [2.0 ** -18] * 3
    e32 = engine_from_operator(d, precision="fp32", compressed=False)
    e16 = engine_from_operator(d, precision="fp16", compressed=False, scales=(sv, si))
    e32.iterate(1)
    e16.iterate(1)
    v32 = e32.get_field(0, np.float32)
    v16 = e16.get_field(0, np.float32)
    emu = (v32 / np.float32(sv[0])).astype(np.float16).astype(np.float32) * np.float32(sv[0])
    np.testing.assert_array_equal(v16, emu)


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-4), ("fp16", 5e-2), ("bf16", 3e-1)])
def test_reduced_precision_s11_vs_fp64(gpu, precision, tol):
    sc64 = scenarios.graded_port_scenario(n_cells=48, steps=3000, prec=1, n_boxes=6, pml=8)
    sc32 = scenarios.graded_port_scenario(n_cells=48, steps=3000, prec=0, n_boxes=6, pml=8)
    freqs = np.linspace(1e9, 10e9, 91)
    s_ref, _, _ = s11_run(sc64, "fp64", 3000, freqs=freqs)
    scales = pilot_scales(sc32.data) if precision in ("fp16", "bf16") else None
    s, _, _ = s11_run(sc32, precision, 3000, scales=scales, freqs=freqs)
    err = np.max(np.abs(s - s_ref))
    print("%s max|S11-S11_fp64| = %.3e" % (precision, err))
    assert err <= tol
