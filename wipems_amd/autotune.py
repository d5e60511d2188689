"""Kernel-shape autotuner.

Runs short trials on scratch copies of a domain (or of a batch of domains),
timing each candidate shape with HIP events around the update kernels, and
returns the fastest.  Shapes never change results (every shape is checked
bit-exact against the oracle in the GPU tests); only speed.
"""
from __future__ import annotations

from .engine import Batch, engine_from_operator

BATCH_CANDIDATES = [64, 128, 256]


def tune_batch(datas, precision="fp32", compressed=None, scales=None, steps=10, candidates=BATCH_CANDIDATES):
    """Best threads-per-block of the batched kernels and the timings [ms/step]."""
    engines = [engine_from_operator(d, precision=precision, compressed=compressed, scales=scales) for d in datas]
    b = Batch(engines)
    results = {}
    for blk in candidates:
        b.set_block(blk)
        b.iterate(2)
        b.set_kernel_timing(True)
        b.iterate(steps)
        tv, tc, n = b.kernel_times()
        b.set_kernel_timing(False)
        results[blk] = (tv + tc) / max(n, 1)
    b.synchronize()
    b.close()
    for e in engines:
        e.close()
    best = min(results, key=results.get)
    return best, results
