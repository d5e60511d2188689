"""Launch-shape autotuner of the batched engine.

Runs short trials on scratch copies of a batch of domains, timing each
candidate shape with HIP events around the update launches, and returns the
fastest.  A shape is (threads per block, UPML launch form): the form is
either one UPML-form launch per phase over every member, or each member's
interior in a launch without UPML code plus one shell launch for all members
(fdtd_batch_set_split).  Shapes never change results (every shape is checked
bit-exact against the solo runs and the oracle in the GPU tests); only speed.
"""
from __future__ import annotations

from .engine import Batch, engine_from_operator

BATCH_CANDIDATES = [(128, 0), (256, 0), (128, 1), (256, 1)]


def tune_batch(datas, precision="fp32", compressed=None, scales=None, steps=10, candidates=BATCH_CANDIDATES):
    """Best (threads per block, split) of the batched kernels and the timings
    [ms per step, both update phases] of every candidate."""
    engines = [engine_from_operator(d, precision=precision, compressed=compressed, scales=scales) for d in datas]
    b = Batch(engines)
    results = {}
    for blk, split in candidates:
        b.set_block(blk)
        b.set_split(split)
        b.iterate(2)
        b.set_kernel_timing(True)
        b.iterate(steps)
        tv, tc, n = b.kernel_times()
        b.set_kernel_timing(False)
        results[(blk, split)] = (tv + tc) / max(n, 1)
    b.synchronize()
    b.close()
    for e in engines:
        e.close()
    best = min(results, key=results.get)
    return best, results
