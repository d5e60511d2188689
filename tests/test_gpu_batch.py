"""Batched domains (fdtd_batch_*): every member gives the same bits as the
same domain stepped alone, and as the CPU oracle."""
import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.engine import Batch, engine_from_operator
from oracle.oracle import oracle_from_operator

pytestmark = pytest.mark.gpu


def small_batch(prec=0):
    return scenarios.batched_port_scenarios(count=5, lo=22, hi=34, steps=200, prec=prec, pml=6, n_boxes=3)


@pytest.mark.parametrize("block", [64, 256])
def test_batch_members_match_solo_runs(gpu, block):
    scs = small_batch()
    solo = [engine_from_operator(s.data, precision="fp32") for s in scs]
    for e in solo:
        e.iterate(200)
        e.synchronize()
    members = [engine_from_operator(s.data, precision="fp32") for s in scs]
    b = Batch(members)
    b.set_block(block)
    b.iterate(7)
    b.iterate(193)
    b.synchronize()
    for s, e, m in zip(scs, solo, members):
        assert m.GetNumberOfTimesteps() == 200
        for w in (0, 1):
            np.testing.assert_array_equal(m.get_field(w), e.get_field(w))
        for pid in range(len(s.data.probes)):
            np.testing.assert_array_equal(m.read_probe(pid)[1], e.read_probe(pid)[1])


def test_batch_member_matches_oracle(gpu):
    scs = small_batch()
    members = [engine_from_operator(s.data, precision="fp32") for s in scs]
    b = Batch(members)
    b.iterate(200)
    b.synchronize()
    o = oracle_from_operator(scs[2].data, ftz=True)
    o.iterate(200)
    for w in (0, 1):
        np.testing.assert_array_equal(members[2].get_field(w), o.get_field(w))
    np.testing.assert_array_equal(members[2].read_probe(0)[1], o.read_probe(0)[1])


def test_batch_fp64(gpu):
    scs = small_batch(prec=1)[:3]
    members = [engine_from_operator(s.data, precision="fp64", compressed=False) for s in scs]
    b = Batch(members)
    b.iterate(120)
    b.synchronize()
    o = oracle_from_operator(scs[1].data)
    o.iterate(120)
    np.testing.assert_array_equal(members[1].get_field(0), o.get_field(0))


def test_batch_rejects_mixed_precision(gpu):
    scs = small_batch()[:2]
    a = engine_from_operator(scs[0].data, precision="fp32")
    c = engine_from_operator(scs[1].data, precision="fp64")
    from wipems_amd import FdtdError
    with pytest.raises(FdtdError):
        Batch([a, c])
