"""Launch-shape autotuner for batches and single meshes.

Builds the solve once per candidate option set (cells per block, cell
ordering, XCD block order, and optionally the residual accumulation), times a
few steps of each after a warm-up, and returns the fastest.  The launch-shape
candidates give the same results bit for bit (the layout tests check that),
so that choice only moves time; the atomic-scatter accumulation changes the
summation order, so it is offered only where bit-exactness is not required
(`extra=ATOMIC`, the FP32-mixed form).
"""
from __future__ import annotations

import itertools
import time

from .lib import Batch, Solver

# block sizes 128/256 (192 leaves 64-cell chunks straddling waves), RCB order
# (RCM's long ranges give 4-9x larger halos); XCD order: runs of 8 blocks per
# XCD, none, or one contiguous range per XCD
DEFAULT_SPACE = {"block_size": [128, 256], "renumber": [1], "xcd_remap": [8, 0, 1]}
# the atomic-scatter accumulation at the batch's best launch shape
ATOMIC = [{"block_size": 128, "renumber": 1, "xcd_remap": 8, "accumulate": 1}]
# reverse Cuthill-McKee cell order (longer ranges, larger halos) at the same shape
RCM = [{"block_size": 128, "renumber": 2, "xcd_remap": 8}]


def candidates(space=None, extra=None):
    space = space or DEFAULT_SPACE
    keys = sorted(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))
    yield from extra or []


def tune_batch(cases, precision: int, space=None, warmup: int = 4, steps: int = 20, log=None,
               extra=None):
    """Fastest option set for this batch: (options, [(options, ms per step)])."""
    table = []
    for opt in candidates(space, extra):
        b = Batch(cases, precision=precision, **opt)
        b.init()
        b.step(warmup)
        t0 = time.perf_counter()
        b.step(steps)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        b.close()
        table.append((opt, ms))
        if log:
            log(f"autotune {opt}: {ms:.3f} ms/step")
    best = min(table, key=lambda t: t[1])[0]
    return best, table


def tune_single(grid, case, precision: int, space=None, warmup: int = 4, steps: int = 20,
                log=None, extra=None, **fixed):
    """Fastest option set for one mesh through fvk_ctx: (options, table)."""
    table = []
    for opt in candidates(space, extra):
        s = Solver(grid, case, precision=precision, **fixed, **opt)
        s.init_freestream()
        s.step(warmup)
        t0 = time.perf_counter()
        s.step(steps)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        s.close()
        table.append((opt, ms))
        if log:
            log(f"autotune {opt}: {ms:.3f} ms/step")
    best = min(table, key=lambda t: t[1])[0]
    return best, table
