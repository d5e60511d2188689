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


@pytest.mark.parametrize("block,split", [(64, 1), (256, 1), (128, 0), (256, 0)])
def test_batch_members_match_solo_runs(gpu, block, split):
    """Interior/shell launches per member (split 1) and one UPML-form launch
    (split 0) give each member the bits of its solo run."""
    scs = small_batch()
    solo = [engine_from_operator(s.data, precision="fp32") for s in scs]
    for e in solo:
        e.iterate(200)
        e.synchronize()
    members = [engine_from_operator(s.data, precision="fp32") for s in scs]
    b = Batch(members)
    b.set_block(block)
    b.set_split(split)
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


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_batch_reduced_precision_members_match_solo_runs(gpu, precision):
    """2-byte storage: split members (half-width shell vectors) and unsplit
    members give the bits of the solo runs."""
    from wipems_amd.engine import pilot_scales
    scs = small_batch()[:3]
    scales = [pilot_scales(s.data, steps=64) for s in scs]
    solo = [engine_from_operator(s.data, precision=precision, scales=sc) for s, sc in zip(scs, scales)]
    for e in solo:
        e.iterate(150)
        e.synchronize()
    for split in (1, 0):
        members = [engine_from_operator(s.data, precision=precision, scales=sc) for s, sc in zip(scs, scales)]
        b = Batch(members)
        b.set_split(split)
        b.iterate(150)
        b.synchronize()
        for e, m in zip(solo, members):
            for w in (0, 1):
                np.testing.assert_array_equal(m.get_field(w), e.get_field(w))
        b.close()


def test_batch_mixed_coefficient_modes(gpu):
    """Members with compressed and with full coefficient arrays in one batch
    (per-domain dispatch) match the oracle."""
    scs = small_batch()[:3]
    members = [engine_from_operator(s.data, precision="fp32", compressed=(k != 1)) for k, s in enumerate(scs)]
    assert members[0].compressed and not members[1].compressed
    b = Batch(members)
    b.iterate(120)
    b.synchronize()
    for k in (0, 1):
        o = oracle_from_operator(scs[k].data, ftz=True)
        o.iterate(120)
        for w in (0, 1):
            np.testing.assert_array_equal(members[k].get_field(w), o.get_field(w))
