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


def s11_run(sc, precision, steps, scales=None, freqs=None, factored=False):
    d = sc.data
    eng = engine_from_operator(d, precision=precision, compressed=False, scales=scales, factored=factored)
    assert eng.factored == factored
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


# measured on MI355X (48^3 cells, 3000 steps): FP32 3.1e-8, FP16 7.7e-5, BF16 1.8e-3;
# "factored": coefficients in the factored form (m rounded to 16 bits)
@pytest.mark.parametrize("precision,factored,tol", [("fp32", False, 1e-6), ("fp16", False, 1e-3), ("bf16", False, 1e-2),
                                                    ("fp32", True, 1e-4), ("fp16", True, 1e-3)])
def test_reduced_precision_s11_vs_fp64(gpu, precision, factored, tol):
    sc64 = scenarios.graded_port_scenario(n_cells=48, steps=3000, prec=1, n_boxes=6, pml=8)
    sc32 = scenarios.graded_port_scenario(n_cells=48, steps=3000, prec=0, n_boxes=6, pml=8)
    freqs = np.linspace(1e9, 10e9, 91)
    s_ref, _, _ = s11_run(sc64, "fp64", 3000, freqs=freqs)
    scales = pilot_scales(sc32.data) if precision in ("fp16", "bf16") else None
    s, _, _ = s11_run(sc32, precision, 3000, scales=scales, freqs=freqs, factored=factored)
    err = np.max(np.abs(s - s_ref))
    print("%s%s max|S11-S11_fp64| = %.3e" % (precision, " factored" if factored else "", err))
    assert err <= tol


# Drude slab + Lorentz spheres lit by a TFSF plane wave (the dispersive
# workload, material frequencies x1e-3, timestep factor 0.9) on 97 x 97 x 65
# lines for 3000 steps: probe rel-L2 below and above the slab against the
# FP64 engine.  Bounds are a few times the values measured on MI355X.
DISPERSIVE_TOL = {"fp32": 1e-4, "fp16": 5e-2, "bf16": 5e-1}


def test_dispersive_reduced_precision_vs_fp64(gpu):
    kw = dict(num_lines=(97, 97, 65), steps=3000, material_scale=1e-3, ts_factor=0.9)
    sc64 = scenarios.dispersive_scenario(prec=1, **kw)
    sc32 = scenarios.dispersive_scenario(prec=0, **kw)
    e64 = engine_from_operator(sc64.data, precision="fp64", compressed=False)
    e64.iterate(3000)
    ref = [e64.read_probe(p)[1] for p in range(2)]
    out = {}
    for precision in ("fp32", "fp16", "bf16"):
        scales = pilot_scales(sc32.data) if precision != "fp32" else None
        e = engine_from_operator(sc32.data, precision=precision, scales=scales)
        e.iterate(3000)
        out[precision] = [float(np.linalg.norm(e.read_probe(p)[1] - ref[p]) / np.linalg.norm(ref[p])) for p in range(2)]
        e.close()
    print("dispersive 97x97x65, 3000 steps, probe rel-L2 vs FP64 (below, above the slab): %s" % out)
    for precision, errs in out.items():
        assert max(errs) <= DISPERSIVE_TOL[precision], (precision, errs)
