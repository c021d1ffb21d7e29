"""Tile-shape autotuning of the stencil kernels.

A tile shape is (node rows per wave strip, node planes per z block).  More
rows re-read fewer halo rows but hold more registers (fewer waves per SIMD);
shorter z blocks give more workgroups to small or shallow boards at the cost
of re-reading one halo plane per block.  The candidates below are timed on
the device and the fastest is kept per board shape.
"""

from __future__ import annotations

import time


def apply_candidates(element_shape) -> list[tuple[int, int]]:
    """(rows, planes per z block) candidates for one board."""

    s = int(element_shape[0])
    planes = s + 1
    # z blocks hold whole plane pairs (the reduction cells); the last block
    # may be odd.
    even = lambda n: n + (n & 1)  # noqa: E731
    blocks = sorted({planes, even(-(-planes // 2)), even(-(-planes // 3)), even(-(-planes // 4)), 4, 2})
    out = []
    for rows in (4, 8):
        for zb in blocks:
            if 2 <= zb <= planes:
                out.append((rows, zb))
    return out


def tune_apply(device_operator, element_shape, timer) -> tuple[int, int]:
    """Time ``timer(launch)`` for every candidate and keep the fastest shape."""

    best = None
    for rows, zb in apply_candidates(element_shape):
        device_operator.set_shape(rows, zb)
        t = timer()
        if best is None or t < best[0]:
            best = (t, rows, zb)
    assert best is not None
    device_operator.set_shape(best[1], best[2])
    return best[1], best[2]


def tune_batch(batch, run=None, rows_candidates=(4, 8), target_candidates=(512, 1024, 2048, 4096)):
    """Time one full batch run per shape and keep the fastest (the batch's
    slots must already be set)."""

    best = None
    for rows in rows_candidates:
        for target in target_candidates:
            batch.set_shape(rows, target)
            t0 = time.perf_counter()
            (run or batch.run)()
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, rows, target)
    assert best is not None
    batch.set_shape(best[1], best[2])
    return best[1], best[2]
