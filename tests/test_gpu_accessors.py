"""Per-element and bulk field access of the engine (Engine::GetVolt/GetCurr/
SetVolt/SetCurr, FDTD/engine.h:52-61, "must be overloaded by any new engine
using a different storage model"; fdtd_get_value/fdtd_set_value and
fdtd_get_field/fdtd_set_field) against the CPU oracle, through the C ABI.
FP16/BF16 handles store value/scale rounded to the 2-byte type and return
stored*scale in float; that rounding is emulated in numpy."""
import numpy as np
import pytest

from oracle.oracle import oracle_from_operator
from wipems_amd import FdtdError, scenarios
from wipems_amd.engine import engine_from_operator

pytestmark = pytest.mark.gpu


def _case():
    return scenarios.graded_port_scenario(n_cells=20, steps=200, prec=0, n_boxes=4, pml=4)


def _positions(rng, shape, k=60):
    return [(int(rng.integers(3)),) + tuple(int(rng.integers(n)) for n in shape) for _ in range(k)]


def test_get_set_value_match_oracle(gpu):
    sc = _case()
    d = sc.data
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(60)
    orc.iterate(60)
    V, I = orc.get_field(0), orc.get_field(1)
    rng = np.random.default_rng(5)
    for (n, x, y, z) in _positions(rng, d.num_lines):
        assert eng.GetVolt(n, x, y, z) == V[n, x, y, z]
        assert eng.GetCurr(n, x, y, z) == I[n, x, y, z]
    # SetVolt/SetCurr on the device, the same edits on the oracle's fields
    for (n, x, y, z) in _positions(rng, d.num_lines, 25):
        v = float(np.float32(rng.standard_normal()))
        c = float(np.float32(rng.standard_normal() * 1e-2))
        eng.SetVolt(n, x, y, z, v)
        eng.SetCurr(n, x, y, z, c)
        V[n, x, y, z] = v
        I[n, x, y, z] = c
    orc.set_field(0, V)
    orc.set_field(1, I)
    eng.iterate(40)
    orc.iterate(40)
    for w in (0, 1):
        np.testing.assert_array_equal(eng.get_field(w), orc.get_field(w))


def test_set_field_matches_oracle(gpu):
    sc = _case()
    d = sc.data
    rng = np.random.default_rng(9)
    V = rng.standard_normal((3,) + tuple(d.num_lines)).astype(np.float32)
    I = (1e-2 * rng.standard_normal((3,) + tuple(d.num_lines))).astype(np.float32)
    eng = engine_from_operator(d, precision="fp32", compressed=True)
    orc = oracle_from_operator(d, ftz=True)
    eng.set_field(0, V)
    eng.set_field(1, I.astype(np.float64))     # double input is rounded to the storage type
    orc.set_field(0, V.astype(np.float64))
    orc.set_field(1, I.astype(np.float64))
    np.testing.assert_array_equal(eng.get_field(0), V)
    eng.iterate(30)
    orc.iterate(30)
    for w in (0, 1):
        np.testing.assert_array_equal(eng.get_field(w), orc.get_field(w))


def _emulate(v, scale, dt):
    s = np.float32(scale)
    inv = np.float32(1.0 / scale)
    if dt == "fp16":
        q = (np.float32(v) * inv).astype(np.float16).astype(np.float32)
    else:   # bfloat16: round to nearest even on the top 16 bits
        b = np.array([np.float32(v) * inv], dtype=np.float32).view(np.uint32)
        b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
        q = b.astype(np.uint32).view(np.float32)[0]
    return float(np.float32(q) * s)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_reduced_precision_accessors_scale_and_round(gpu, precision):
    sc = _case()
    d = sc.data
    sv, si = [2.0 ** -6, 2.0 ** -3, 2.0 ** -9], [2.0 ** -14, 2.0 ** -12, 2.0 ** -16]
    eng = engine_from_operator(d, precision=precision, compressed=False, scales=(sv, si))
    rng = np.random.default_rng(13)
    for (n, x, y, z) in _positions(rng, d.num_lines, 40):
        v = float(rng.standard_normal() * 1e-2)
        c = float(rng.standard_normal() * 1e-5)
        eng.SetVolt(n, x, y, z, v)
        eng.SetCurr(n, x, y, z, c)
        assert eng.GetVolt(n, x, y, z) == _emulate(v, sv[n], precision)
        assert eng.GetCurr(n, x, y, z) == _emulate(c, si[n], precision)
    # bulk: set_field stores value/scale rounded, get_field returns stored*scale
    V = (1e-2 * rng.standard_normal((3,) + tuple(d.num_lines))).astype(np.float32)
    eng.set_field(0, V)
    got = eng.get_field(0, np.float32)
    for n in range(3):
        flat = V[n].ravel()[:500]
        want = np.array([_emulate(float(v), sv[n], precision) for v in flat], dtype=np.float32)
        np.testing.assert_array_equal(got[n].ravel()[:500], want)


def test_accessors_reject_out_of_range(gpu):
    sc = _case()
    eng = engine_from_operator(sc.data, precision="fp32")
    nx, ny, nz = sc.data.num_lines
    for args in ((3, 0, 0, 0), (0, nx, 0, 0), (0, 0, ny, 0), (0, 0, 0, nz)):
        with pytest.raises(FdtdError):
            eng.GetVolt(*args)
        with pytest.raises(FdtdError):
            eng.SetCurr(*args, 1.0)
