"""fdtd_calc_energy_noalloc (partial sums in the library's static device
scratch array, no allocation) against fdtd_calc_energy (handle-owned
buffers) through the C ABI: bitwise equal on two calls in a row, the second
on a smaller domain, on the fields of real runs; the scratch keeps the last
call's contents, so the second call also checks that nothing stale leaks."""
import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.engine import engine_from_operator

pytestmark = pytest.mark.gpu


def _run(n_cells, steps, precision="fp32"):
    sc = scenarios.cavity_point_source(n_cells, steps, prec=1 if precision == "fp64" else 0)
    eng = engine_from_operator(sc.data, precision=precision, compressed=precision != "fp64")
    eng.iterate(steps)
    return eng


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_energy_equals_energy(gpu, precision):
    big = _run(40, 120, precision)
    small = _run(12, 30, precision)
    # noalloc first on each handle: fdtd_calc_energy has not allocated yet
    e_big_na = big.energy_noalloc()
    e_small_na = small.energy_noalloc()
    e_big = big.energy()
    e_small = small.energy()
    print("energy %s: big %.17e small %.17e" % (precision, e_big, e_small))
    assert e_big > 0 and e_small > 0 and e_small != e_big
    assert e_big_na == e_big
    assert e_small_na == e_small
    # and again after more steps, smaller domain first this time
    small.iterate(10)
    big.iterate(10)
    assert small.energy_noalloc() == small.energy()
    assert big.energy_noalloc() == big.energy()
    big.close()
    small.close()
