"""Tensor-product in-plane grids: uniform or graded grid lines per axis.

The reference interface is src/electrical/matrix_free_mpir_fem/grid.py:35-221:
``TensorGrid`` (grid lines of a structured in-plane mesh and the cell
queries on them), ``merge_intervals``, ``graded_edges``, ``refined_grid`` and
``check_pitch_axis``, with the reference's arguments and messages.  A graded
grid keeps every kernel structured: only the per-element coefficients of the
unit element matrices change, which the meshes take as per-column and per-row
pitch arrays (``pitch_x_m`` / ``pitch_y_m``).

``graded_edges`` asks for cells of the fine pitch F over the refined
intervals and of the coarse pitch C far away, growing geometrically by at
most ``growth`` = g between: a target size h(d) = min(C, F + (g - 1) d) at
distance d from the nearest refined interval (a geometric progression of
ratio g starting at F has that size after covering d).  Grid lines sit where
the cumulative cell count, the integral of 1 / h, crosses whole numbers
(scaled so that the count, the rounded integral, ends exactly at the end of
the range).  Here the integral is taken in closed form -- ln(1 + (g - 1) d /
F) / (g - 1) on a growth ramp, linear on fine and coarse stretches -- over a
table of pieces on which the distance rises or falls monotonically, and the
lines come from its exact inverse, so no sampling resolution enters.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np


def check_pitch_axis(pitch, count: int, name: str) -> np.ndarray:
    """One positive finite width per cell: a scalar is repeated ``count`` times."""

    values = np.asarray(pitch, dtype=np.float64)
    if values.ndim == 0:
        values = np.full(int(count), float(values))
    elif values.shape != (int(count),):
        raise ValueError(f"{name} must be a scalar or hold one value per cell ({count})")
    if not np.all(np.isfinite(values)) or np.any(values <= 0.0):
        raise ValueError(f"{name} must be finite and positive")
    return values.copy()


def _grid_lines(edges, name: str) -> np.ndarray:
    lines = np.array(edges, dtype=np.float64).reshape(-1)
    if lines.size < 2:
        raise ValueError(f"{name} needs at least two grid lines")
    if not (np.isfinite(lines).all() and (np.diff(lines) > 0.0).all()):
        raise ValueError(f"{name} must be finite and strictly increasing")
    return lines


@dataclass(frozen=True)
class TensorGrid:
    """Grid lines of a structured in-plane mesh, in metres: ``x_edges_m``
    (``cols + 1`` lines) and ``y_edges_m`` (``rows + 1``).  ``pitch_x_m`` /
    ``pitch_y_m`` are the per-column / per-row cell widths the meshes take."""

    x_edges_m: np.ndarray
    y_edges_m: np.ndarray

    def __post_init__(self) -> None:
        object.__setattr__(self, "x_edges_m", _grid_lines(self.x_edges_m, "x_edges_m"))
        object.__setattr__(self, "y_edges_m", _grid_lines(self.y_edges_m, "y_edges_m"))

    @classmethod
    def uniform(cls, pitch_m, shape: tuple[int, int], origin_m: tuple[float, float] = (0.0, 0.0)) -> "TensorGrid":
        """``shape = (rows, cols)`` cells of one pitch, or of ``(pitch_x, pitch_y)``."""

        px, py = np.broadcast_to(np.asarray(pitch_m, dtype=np.float64), (2,))
        rows, cols = (int(v) for v in shape)
        return cls(origin_m[0] + float(px) * np.arange(cols + 1), origin_m[1] + float(py) * np.arange(rows + 1))

    @property
    def shape(self) -> tuple[int, int]:
        return self.y_edges_m.size - 1, self.x_edges_m.size - 1

    @property
    def size(self) -> int:
        rows, cols = self.shape
        return rows * cols

    @property
    def pitch_x_m(self) -> np.ndarray:
        return np.diff(self.x_edges_m)

    @property
    def pitch_y_m(self) -> np.ndarray:
        return np.diff(self.y_edges_m)

    @property
    def origin_m(self) -> tuple[float, float]:
        return float(self.x_edges_m[0]), float(self.y_edges_m[0])

    @property
    def is_uniform(self) -> bool:
        """Every cell the same square (what one scalar pitch describes); equal
        but oblong cells do not count."""

        widths = np.concatenate((self.pitch_x_m, self.pitch_y_m))
        return bool(np.allclose(widths, widths[0]))

    @property
    def cell_area_m2(self) -> np.ndarray:
        return np.multiply.outer(self.pitch_y_m, self.pitch_x_m)

    @property
    def x_centres_m(self) -> np.ndarray:
        return 0.5 * (self.x_edges_m[1:] + self.x_edges_m[:-1])

    @property
    def y_centres_m(self) -> np.ndarray:
        return 0.5 * (self.y_edges_m[1:] + self.y_edges_m[:-1])

    def cell_of(self, x_m, y_m) -> tuple[np.ndarray, np.ndarray]:
        """(row, column) of the cells holding the points, clipped to the grid."""

        rows, cols = self.shape
        col = np.searchsorted(self.x_edges_m, np.asarray(x_m), side="right") - 1
        row = np.searchsorted(self.y_edges_m, np.asarray(y_m), side="right") - 1
        return np.clip(row, 0, rows - 1), np.clip(col, 0, cols - 1)

    def cell_mask(self, x0_m: float, x1_m: float, y0_m: float, y1_m: float) -> np.ndarray:
        """``(rows, cols)``: the cells whose centre lies in the closed rectangle."""

        in_x = (self.x_centres_m >= x0_m) & (self.x_centres_m <= x1_m)
        in_y = (self.y_centres_m >= y0_m) & (self.y_centres_m <= y1_m)
        return np.logical_and.outer(in_y, in_x)

    def cell_overlap_fraction(self, x0_m: float, x1_m: float, y0_m: float, y1_m: float) -> np.ndarray:
        """``(rows, cols)``: the share of each cell's area inside the
        rectangle; weighted by ``cell_area_m2`` it sums to the rectangle's
        area on any grid."""

        def covered(edges, lo, hi):
            return np.clip(np.minimum(edges[1:], hi) - np.maximum(edges[:-1], lo), 0.0, None)

        inside = np.multiply.outer(covered(self.y_edges_m, y0_m, y1_m), covered(self.x_edges_m, x0_m, x1_m))
        return inside / self.cell_area_m2


def merge_intervals(intervals: Sequence[tuple[float, float]]) -> list[tuple[float, float]]:
    """The union of closed intervals (either end first) as sorted disjoint
    intervals; touching ones merge."""

    merged: list[list[float]] = []
    for lo, hi in sorted((min(a, b), max(a, b)) for a, b in intervals):
        if merged and lo <= merged[-1][1]:
            merged[-1][1] = max(merged[-1][1], hi)
        else:
            merged.append([lo, hi])
    return [(lo, hi) for lo, hi in merged]


class _CellCount:
    """The cumulative target cell count N(x) = integral of 1 / h and its
    inverse over the pieces of [start, end] on which the distance to the
    refined intervals is zero, rising or falling."""

    FINE, RISING, FALLING = 0, 1, 2

    def __init__(self, start: float, end: float, fine: list[tuple[float, float]], coarse_pitch: float,
                 fine_pitch: float, growth: float):
        self.f, self.c, self.k = fine_pitch, coarse_pitch, growth - 1.0
        self.ramp = (coarse_pitch - fine_pitch) / self.k  # distance at which h reaches the coarse pitch
        self.ramp_count = np.log(coarse_pitch / fine_pitch) / self.k
        pieces = []  # (kind, x0, x1, anchor): distance 0, x - anchor or anchor - x
        cursor = start
        for i, (lo, hi) in enumerate(fine):
            if lo > cursor:
                if i == 0:
                    pieces.append((self.FALLING, cursor, lo, lo))
                else:
                    mid = 0.5 * (cursor + lo)
                    pieces.append((self.RISING, cursor, mid, cursor))
                    pieces.append((self.FALLING, mid, lo, lo))
            pieces.append((self.FINE, lo, hi, lo))
            cursor = hi
        if end > cursor:
            pieces.append((self.RISING, cursor, end, cursor))
        self.pieces = pieces
        counts = [0.0]
        for kind, x0, x1, anchor in pieces:
            counts.append(counts[-1] + self._across(kind, x0, x1, anchor))
        self.starts = np.array(counts[:-1])
        self.total = counts[-1]

    def _count_to(self, d):
        """Cells from distance 0 to distance d."""

        d = np.asarray(d, dtype=np.float64)
        return np.where(d <= self.ramp, np.log1p(self.k * d / self.f) / self.k,
                        self.ramp_count + (d - self.ramp) / self.c)

    def _distance_at(self, n):
        """Inverse of _count_to."""

        n = np.asarray(n, dtype=np.float64)
        return np.where(n <= self.ramp_count, self.f * np.expm1(self.k * n) / self.k,
                        self.ramp + (n - self.ramp_count) * self.c)

    def _across(self, kind, x0, x1, anchor) -> float:
        if kind == self.FINE:
            return (x1 - x0) / self.f
        if kind == self.RISING:
            return float(self._count_to(x1 - anchor))
        return float(self._count_to(anchor - x0) - self._count_to(anchor - x1))

    def inverse(self, counts: np.ndarray) -> np.ndarray:
        """Positions where the cumulative count reaches ``counts``."""

        which = np.clip(np.searchsorted(self.starts, counts, side="right") - 1, 0, len(self.pieces) - 1)
        out = np.empty_like(counts)
        for p, (kind, x0, x1, anchor) in enumerate(self.pieces):
            sel = which == p
            if not sel.any():
                continue
            local = counts[sel] - self.starts[p]
            if kind == self.FINE:
                x = x0 + local * self.f
            elif kind == self.RISING:
                x = anchor + self._distance_at(local)
            else:
                x = anchor - self._distance_at(self._count_to(anchor - x0) - local)
            out[sel] = np.clip(x, x0, x1)
        return out


def graded_edges(start_m: float, end_m: float, *, coarse_pitch_m: float, fine_pitch_m: float,
                 refine_m: Sequence[tuple[float, float]] = (), margin_m: float = 0.0,
                 growth: float = 1.5) -> np.ndarray:
    """Grid lines from ``start_m`` to ``end_m``: cells of ``fine_pitch_m``
    over every interval of ``refine_m`` widened by ``margin_m`` each side,
    ``coarse_pitch_m`` far away, growing by at most ``growth`` per cell in
    between; the refined intervals need not start on a grid line.  With no
    refined interval the lines are uniform, as near the coarse pitch as a
    whole count allows."""

    if not (np.isfinite(start_m) and np.isfinite(end_m)) or end_m <= start_m:
        raise ValueError("end_m must exceed start_m")
    if coarse_pitch_m <= 0.0 or fine_pitch_m <= 0.0 or fine_pitch_m > coarse_pitch_m:
        raise ValueError("pitches must be positive with fine_pitch_m <= coarse_pitch_m")
    if growth <= 1.0:
        raise ValueError("growth must exceed one")
    if margin_m < 0.0:
        raise ValueError("margin_m must be non-negative")
    widened = [(max(start_m, min(a, b) - margin_m), min(end_m, max(a, b) + margin_m)) for a, b in refine_m]
    fine = [(lo, hi) for lo, hi in merge_intervals(widened) if hi > lo]
    if not fine:
        return np.linspace(start_m, end_m, max(1, int(round((end_m - start_m) / coarse_pitch_m))) + 1)
    cells = _CellCount(float(start_m), float(end_m), fine, float(coarse_pitch_m), float(fine_pitch_m), float(growth))
    count = max(1, int(round(cells.total)))
    edges = cells.inverse(np.linspace(0.0, cells.total, count + 1))
    edges[0], edges[-1] = start_m, end_m
    return edges


def refined_grid(x_range_m: tuple[float, float], y_range_m: tuple[float, float], *, coarse_pitch_m: float,
                 fine_pitch_m: float, refine_boxes_m: Sequence[tuple[float, float, float, float]] = (),
                 margin_m: float = 0.0, growth: float = 1.5) -> TensorGrid:
    """A tensor grid fine over the ``(x0, x1, y0, y1)`` boxes and coarse
    elsewhere: each box refines its x strip and its y strip."""

    boxes = np.asarray(refine_boxes_m, dtype=np.float64).reshape(-1, 4)
    options = dict(coarse_pitch_m=coarse_pitch_m, fine_pitch_m=fine_pitch_m, margin_m=margin_m, growth=growth)
    return TensorGrid(graded_edges(*x_range_m, refine_m=[tuple(b) for b in boxes[:, 0:2]], **options),
                      graded_edges(*y_range_m, refine_m=[tuple(b) for b in boxes[:, 2:4]], **options))


__all__ = ["TensorGrid", "check_pitch_axis", "graded_edges", "merge_intervals", "refined_grid"]
