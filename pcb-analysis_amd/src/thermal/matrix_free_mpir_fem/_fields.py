"""Validation and node-grid helpers shared by the thermal front end.

Inputs of the reference's API arrive as scalars, per-slab lists, face
arrays or full element/node arrays.  ``Field`` describes one such input (the
ranks it may take, the bounds its values must respect and the message each
violation raises) and ``Field.read`` turns a value into a checked float
array, so every class states its inputs as a table instead of a sequence of
``if`` statements.  The message texts are the reference's, since callers
match on them.

``spread_to_nodes`` moves per-element (or per-face) values onto their
corner nodes: a node receives the sum of the values of the cells that touch
it along the chosen axes.  It is separable (one 2-point pass per axis), the
same sum the reference forms by adding eight (or four) shifted corner views.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Iterable

import numpy as np


@dataclass(frozen=True)
class Field:
    """One numeric input: allowed array ranks and value bounds.

    ``rank_message`` is raised for a value of a rank not in ``ranks``;
    ``value_message`` for a non-finite value, or one outside [``low``,
    ``high``] (``low_open``: the low bound itself is not allowed)."""

    ranks: tuple[int, ...]
    rank_message: str
    value_message: str
    low: float | None = None
    low_open: bool = False
    high: float | None = None

    def read(self, value: Any) -> np.ndarray:
        array = np.asarray(value, dtype=np.float64)
        if array.ndim not in self.ranks:
            raise ValueError(self.rank_message)
        bad = ~np.isfinite(array)
        if self.low is not None:
            bad |= (array <= self.low) if self.low_open else (array < self.low)
        if self.high is not None:
            bad |= array > self.high
        if np.any(bad):
            raise ValueError(self.value_message)
        return array

    def read_compact(self, value: Any) -> float | np.ndarray:
        """As ``read``, with a scalar returned as a Python float."""

        array = self.read(value)
        return float(array) if array.ndim == 0 else array.copy()


def require_shape(value: Any, shape: tuple[int, ...], message: str, *, when_rank: int | None = None) -> None:
    """Raise ``message`` unless ``value`` has ``shape``; with ``when_rank``,
    only values of that rank are checked (the others are broadcastable)."""

    array = np.asarray(value)
    if when_rank is not None and array.ndim != when_rank:
        return
    if array.shape != shape:
        raise ValueError(message)


def per_element(value: Any, shape: tuple[int, int, int], name: str) -> np.ndarray:
    """A scalar, one value per slab, or a full ``(slabs, rows, cols)`` array,
    as a full array."""

    array = np.asarray(value, dtype=np.float64)
    if array.ndim == 0 or array.shape == shape:
        return np.array(np.broadcast_to(array, shape))
    if array.shape == (shape[0],):
        return np.array(np.broadcast_to(array[:, None, None], shape))
    raise ValueError(f"{name} must be a scalar, one value per slab, or match shape {shape}")


def unique_directions(directions: Iterable[str], allowed: tuple[str, ...]) -> tuple[str, ...]:
    out = tuple(directions)
    if not out or not set(out) <= set(allowed):
        raise ValueError(f"directions must be drawn from {allowed}")
    if len(set(out)) != len(out):
        raise ValueError("directions must be unique")
    return out


def spread_to_nodes(cells: np.ndarray, axes: tuple[int, ...]) -> np.ndarray:
    """Every node gets the sum of the ``cells`` values that touch it along
    ``axes`` (each axis grows by one); other axes keep their length."""

    out = np.asarray(cells, dtype=np.float64)
    for axis in axes:
        pad = [(0, 0)] * out.ndim
        pad[axis] = (1, 1)
        padded = np.pad(out, pad)
        lower = [slice(None)] * out.ndim
        upper = [slice(None)] * out.ndim
        lower[axis] = slice(None, -1)
        upper[axis] = slice(1, None)
        out = padded[tuple(lower)] + padded[tuple(upper)]
    return out


def average_to_cells(nodes: np.ndarray, axes: tuple[int, ...]) -> np.ndarray:
    """The mean over each cell's corner nodes along ``axes`` (each axis
    shrinks by one): the transpose of ``spread_to_nodes`` divided by the
    corner count."""

    out = np.asarray(nodes, dtype=np.float64)
    for axis in axes:
        lower = [slice(None)] * out.ndim
        upper = [slice(None)] * out.ndim
        lower[axis] = slice(None, -1)
        upper[axis] = slice(1, None)
        out = 0.5 * (out[tuple(lower)] + out[tuple(upper)])
    return out
