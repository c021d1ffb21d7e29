"""Graded tensor grids on the CPU (the reference's tests/test_tensor_grid.py:
35-62 properties, plus a quadrature check of the closed-form cell count)."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.integrate import quad

from electrical.matrix_free_mpir_fem.grid import TensorGrid, graded_edges, merge_intervals, refined_grid


def test_graded_edges_are_fine_where_asked_and_grow_gently():
    edges = graded_edges(0.0, 60e-3, coarse_pitch_m=0.5e-3, fine_pitch_m=0.1e-3, refine_m=[(20e-3, 24e-3)],
                         margin_m=1e-3, growth=1.5)
    widths = np.diff(edges)
    assert edges[0] == 0.0 and edges[-1] == 60e-3 and (widths > 0.0).all()
    inside = widths[(edges[:-1] >= 19e-3) & (edges[1:] <= 25e-3)]
    assert inside.size >= 55 and np.allclose(inside, 0.1e-3, rtol=2e-3)
    assert widths.max() <= 0.5e-3 * 1.01
    ratio = np.maximum(widths[1:] / widths[:-1], widths[:-1] / widths[1:])
    assert ratio.max() < 1.75
    assert widths.size < 600 / 3
    np.testing.assert_allclose(graded_edges(0.0, 10e-3, coarse_pitch_m=0.5e-3, fine_pitch_m=0.1e-3),
                               np.linspace(0.0, 10e-3, 21))


def _target_size(x, fine, coarse_pitch, fine_pitch, growth):
    d = min(max(0.0, a - x, x - b) for a, b in fine)
    return min(coarse_pitch, fine_pitch + (growth - 1.0) * d)


@pytest.mark.parametrize("refine,margin,growth", [([(20e-3, 24e-3)], 1e-3, 1.5),
                                                  ([(2e-3, 3e-3), (9e-3, 9.5e-3), (30e-3, 41e-3)], 0.2e-3, 1.3),
                                                  ([(0.0, 1e-3), (59e-3, 60e-3)], 0.0, 2.0),
                                                  # gaps shorter than two ramps: partial ramps meet mid-gap
                                                  ([(10e-3, 11e-3), (11.8e-3, 12.5e-3), (14e-3, 15e-3)], 0.0, 1.2)])
def test_every_cell_holds_the_same_share_of_the_target_cell_count(refine, margin, growth):
    # Grid lines sit where the integral of 1 / h crosses equal steps: checked
    # cell by cell with adaptive quadrature of the target size itself, so a
    # wrong closed form (log term, ramp length, piece order) fails here.
    coarse, fine_pitch = 0.5e-3, 0.1e-3
    edges = graded_edges(0.0, 60e-3, coarse_pitch_m=coarse, fine_pitch_m=fine_pitch, refine_m=refine,
                         margin_m=margin, growth=growth)
    fine = merge_intervals([(max(0.0, a - margin), min(60e-3, b + margin)) for a, b in refine])
    ramp = (coarse - fine_pitch) / (growth - 1.0)  # distance where the target size reaches the coarse pitch
    breaks = sorted({p for a, b in fine for p in (a, b, a - ramp, b + ramp)}
                    | {0.5 * (b1 + a2) for (_, b1), (a2, _) in zip(fine, fine[1:])})

    def count(lo, hi):
        points = [p for p in breaks if lo < p < hi]
        return quad(lambda x: 1.0 / _target_size(x, fine, coarse, fine_pitch, growth), lo, hi, points=points or None,
                    epsabs=0.0, epsrel=1e-12, limit=200)[0]

    total = count(0.0, 60e-3)
    assert edges.size - 1 == max(1, round(total))
    cells = np.array([count(a, b) for a, b in zip(edges[:-1], edges[1:])])
    np.testing.assert_allclose(cells, total / (edges.size - 1), rtol=1e-8)


def test_refined_grids_and_cell_queries():
    grid = refined_grid((0.0, 60e-3), (0.0, 40e-3), coarse_pitch_m=0.5e-3, fine_pitch_m=0.1e-3,
                        refine_boxes_m=[(20e-3, 24e-3, 10e-3, 14e-3)], margin_m=1e-3)
    assert grid.size < 0.15 * 600 * 400 and not grid.is_uniform
    uniform = TensorGrid.uniform(0.5e-3, (4, 6), origin_m=(1.0, 2.0))
    assert uniform.is_uniform and uniform.shape == (4, 6) and uniform.origin_m == (1.0, 2.0)
    np.testing.assert_allclose(uniform.cell_area_m2, 0.25e-6)
    assert not TensorGrid.uniform((0.2e-3, 0.4e-3), (4, 6)).is_uniform
    row, col = grid.cell_of([21e-3], [12e-3])
    assert grid.x_edges_m[col[0]] <= 21e-3 < grid.x_edges_m[col[0] + 1]
    assert grid.y_edges_m[row[0]] <= 12e-3 < grid.y_edges_m[row[0] + 1]
    assert grid.cell_of([-1.0], [99.0]) == (np.array([grid.shape[0] - 1]), np.array([0]))
    assert grid.cell_mask(20e-3, 24e-3, 10e-3, 14e-3).sum() == pytest.approx(40 * 40, rel=0.1)
    # Overlap fractions weighted by area give the rectangle's area exactly,
    # whether or not grid lines follow it.
    for box in ((20.03e-3, 23.97e-3, 10.01e-3, 13.2e-3), (0.0, 60e-3, 0.0, 40e-3), (5e-3, 5.001e-3, 7e-3, 7.3e-3)):
        area = np.sum(grid.cell_overlap_fraction(*box) * grid.cell_area_m2)
        assert area == pytest.approx((box[1] - box[0]) * (box[3] - box[2]), rel=1e-12)
    np.testing.assert_allclose(grid.pitch_x_m.sum(), 60e-3, rtol=1e-15)


def test_merge_and_validation():
    assert merge_intervals([(3, 4), (1, 2), (2, 2.5), (5, 4.5), (4, 4.2)]) == [(1, 2.5), (3, 4.2), (4.5, 5)]
    with pytest.raises(ValueError, match="fine_pitch_m <= coarse_pitch_m"):
        graded_edges(0.0, 1.0, coarse_pitch_m=0.1, fine_pitch_m=0.2)
    with pytest.raises(ValueError, match="growth"):
        graded_edges(0.0, 1.0, coarse_pitch_m=0.2, fine_pitch_m=0.1, growth=1.0)
    with pytest.raises(ValueError, match="exceed start_m"):
        graded_edges(1.0, 1.0, coarse_pitch_m=0.2, fine_pitch_m=0.1)
    with pytest.raises(ValueError, match="margin_m"):
        graded_edges(0.0, 1.0, coarse_pitch_m=0.2, fine_pitch_m=0.1, margin_m=-1.0)
    with pytest.raises(ValueError, match="strictly increasing"):
        TensorGrid([0.0, 1.0, 1.0], [0.0, 1.0])
    with pytest.raises(ValueError, match="at least two"):
        TensorGrid([0.0], [0.0, 1.0])
