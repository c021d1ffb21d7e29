"""Tile-shape autotuning of the stencil kernels, with a per-shape cache.

Two kinds of shape are tuned:

* the standalone apply's (node rows per wave strip, node planes per z
  block): more rows re-read fewer halo rows but hold more registers (fewer
  waves per SIMD); shorter z blocks give more workgroups to small or shallow
  boards at the cost of one halo plane per block;
* a batch's or server's (rows per wave strip, target workgroups per launch):
  the fused PCG pass splits a board's z extent into blocks until the launch
  holds about ``target`` workgroups, so small boards fill the chip and large
  ones keep long z sweeps.

A batch or server shape is tuned on the fused pass itself, timed with HIP
events on the batch's stream over every slot (``Batch.time_pass``), and the
winner is kept in a JSON cache under a key made of the board set's element
shapes, so a workload that comes back is set up without timing.  The shape
changes no result bit (every tile shape gives the same sums), only the time.
"""

from __future__ import annotations

import hashlib
import json
import os
import threading
from pathlib import Path

PASS_ROWS = (4, 8)
PASS_TARGETS = (512, 1024, 2048, 4096, 8192, 16384)
CACHE_ENV = "PCBHIP_TUNE_CACHE"


def apply_candidates(element_shape) -> list[tuple[int, int]]:
    """(rows, planes per z block) candidates of the standalone apply on one board."""

    s = int(element_shape[0])
    planes = s + 1
    # z blocks hold whole plane pairs (the reduction cells); the last block
    # may be odd.
    even = lambda n: n + (n & 1)  # noqa: E731
    blocks = sorted({planes, even(-(-planes // 2)), even(-(-planes // 3)), even(-(-planes // 4)), 4, 2})
    return [(rows, zb) for rows in (4, 8) for zb in blocks if 2 <= zb <= planes]


def tune_apply(device_operator, element_shape, timer) -> tuple[int, int]:
    """Time ``timer()`` for every apply candidate and keep the fastest shape."""

    best = None
    for rows, zb in apply_candidates(element_shape):
        device_operator.set_shape(rows, zb)
        t = timer()
        if best is None or t < best[0]:
            best = (t, rows, zb)
    assert best is not None
    device_operator.set_shape(best[1], best[2])
    return best[1], best[2]


def one_step_config():
    """A configuration that stops after one inner step: enough to give every
    slot live state for timing the pass."""

    from ._capi import config_c

    return config_c(1e-10, 0.0, 0.5, 1, 1)


def pass_candidates() -> list[tuple[int, int]]:
    return [(rows, target) for rows in PASS_ROWS for target in PASS_TARGETS]


def board_set_key(kind: str, element_shapes, inner_fp64: bool = False) -> str:
    """Cache key of a set of boards: the kind, the inner precision and a
    digest of the sorted element shapes (with repeats: two slots of one board
    weigh twice)."""

    shapes = sorted(tuple(int(v) for v in s) for s in element_shapes)
    digest = hashlib.sha1(json.dumps(shapes).encode()).hexdigest()[:16]
    return f"{kind}/{'f64' if inner_fp64 else 'f32'}/{len(shapes)}/{digest}"


class ShapeCache:
    """Tuned shapes by key in one JSON file (``$PCBHIP_TUNE_CACHE``, else
    ``~/.cache/pcbhip/tile_shapes.json``).  A file that cannot be read is
    treated as empty and one that cannot be written is left alone: the cache
    only saves tuning time."""

    def __init__(self, path: str | os.PathLike | None = None):
        default = Path.home() / ".cache" / "pcbhip" / "tile_shapes.json"
        self.path = Path(path or os.environ.get(CACHE_ENV) or default)
        self._lock = threading.Lock()

    def _load(self) -> dict:
        try:
            data = json.loads(self.path.read_text())
        except (OSError, ValueError):
            return {}
        return data if isinstance(data, dict) else {}

    def get(self, key: str) -> dict | None:
        with self._lock:
            entry = self._load().get(key)
        if not isinstance(entry, dict):
            return None
        if (entry.get("rows"), entry.get("target")) not in set(pass_candidates()):
            return None
        return entry

    def put(self, key: str, entry: dict) -> None:
        with self._lock:
            data = self._load()
            data[key] = entry
            try:
                self.path.parent.mkdir(parents=True, exist_ok=True)
                tmp = self.path.with_suffix(f".{os.getpid()}.tmp")
                tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
                tmp.replace(self.path)
            except OSError:
                pass


def tune_batch(batch, key: str, *, slots, cache: ShapeCache | None = None, launches: int = 20) -> dict:
    """Pick the (rows, target) of ``batch`` by timing its fused pass over
    every slot for each candidate, or take it from ``cache``.

    ``slots`` is ``[(rhs, x0, config), ...]`` of every slot, as passed to
    ``Batch.set``: the batch is first run once with one inner step per slot
    (the pass needs live slot state), then every slot is set back to its own
    right-hand side and configuration.  Returns the chosen entry (``rows``,
    ``target``, ``us`` per pass, ``cached``, the timed ``sweep``)."""

    cache = cache if cache is not None else ShapeCache()
    hit = cache.get(key)
    if hit is not None:
        batch.set_shape(hit["rows"], hit["target"])
        return dict(hit, cached=True)
    probe = one_step_config()
    for i, (rhs, x0, _) in enumerate(slots):
        batch.set(i, rhs, x0, probe)
    batch.run()
    sweep = []
    for rows, target in pass_candidates():
        batch.set_shape(rows, target)
        sweep.append(dict(rows=rows, target=target, us=batch.time_pass(launches)))
    best = min(sweep, key=lambda row: row["us"])
    batch.set_shape(best["rows"], best["target"])
    for i, (rhs, x0, config) in enumerate(slots):
        batch.set(i, rhs, x0, config)
    entry = dict(rows=best["rows"], target=best["target"], us=best["us"], sweep=sweep)
    cache.put(key, entry)
    return dict(entry, cached=False)


def tune_server(server_backend, operators, *, cache: ShapeCache | None = None, launches: int = 20,
                inner_fp64: bool = False) -> dict:
    """The (rows, target) of a server from a representative set of its
    operators: a throwaway batch with one slot per operator (a right-hand
    side of ones suffices for timing) is tuned with ``tune_batch`` under
    the key of that set, and the server takes the winner."""

    import numpy as np

    from ._capi import Batch

    ops = list(operators)
    key = board_set_key("serve", [op.element_shape for op in ops], inner_fp64)
    cache = cache if cache is not None else ShapeCache()
    hit = cache.get(key)
    if hit is None:
        batch = Batch(ops, inner_fp64=inner_fp64)
        try:
            slots = [(np.ones(op.size), None, one_step_config()) for op in ops]
            hit = tune_batch(batch, key, slots=slots, cache=cache, launches=launches)
        finally:
            batch.close()
    else:
        hit = dict(hit, cached=True)
    server_backend.set_shape(hit["rows"], hit["target"])
    return hit
