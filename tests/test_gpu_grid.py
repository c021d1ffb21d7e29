"""Graded tensor grids through the device solve: the reference's hot-spot
comparison (tests/test_tensor_grid.py:93-127: a graded grid lands within
1e-3 of the rise of the uniform fine grid with a fifth of its cells, closer
than the uniform coarse grid) and a refined board against the oracle's
direct FP64 solve of the same graded operator."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import MPIRConfig
from electrical.matrix_free_mpir_fem.grid import TensorGrid, refined_grid
from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi
from tests.helpers import to_oracle
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    LayeredThermalMesh,
    ThermalConductionProblem,
    solve_thermal_conduction,
)

pytestmark = pytest.mark.gpu

AMBIENT = 300.0
PATCH = (14e-3, 16e-3, 9e-3, 11e-3)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _hot_spot(grid: TensorGrid) -> ThermalConductionProblem:
    """1 W over a 2 x 2 mm patch (by cell overlap, so the area is exact on
    every grid) on a copper / 1.5 mm FR-4 / copper board, films on both faces."""

    mesh = LayeredThermalMesh((35e-6, 1.5e-3, 35e-6), grid.pitch_x_m, grid.pitch_y_m, (385.0, 0.8, 385.0),
                              through_plane_conductivity_w_per_m_k=(385.0, 0.3, 385.0))
    overlap = grid.cell_overlap_fraction(*PATCH) * grid.cell_area_m2
    assert np.sum(overlap) == pytest.approx(4e-6, rel=1e-9)
    heat = np.zeros(mesh.element_grid_shape)
    heat[2] = overlap / np.sum(overlap)
    return ThermalConductionProblem(
        mesh, convection=(ConvectionBoundary("top", 10.0, AMBIENT), ConvectionBoundary("bottom", 10.0, AMBIENT)),
        element_heat_w=heat)


def test_graded_grid_matches_the_uniform_fine_grid_at_a_hot_spot():
    coarse, fine = 0.5e-3, 0.125e-3
    grids = {"fine": TensorGrid.uniform(fine, (160, 240)), "coarse": TensorGrid.uniform(coarse, (40, 60)),
             "graded": refined_grid((0.0, 30e-3), (0.0, 20e-3), coarse_pitch_m=coarse, fine_pitch_m=fine,
                                    refine_boxes_m=[PATCH], margin_m=1e-3, growth=1.4)}
    peak = {}
    for name, grid in grids.items():
        print(f"grid hot spot: {name} {grid.shape}", flush=True)
        solution = solve_thermal_conduction(_hot_spot(grid))
        assert solution.solve.converged and abs(solution.heat_balance_error_w) < 1e-9
        peak[name] = solution.max_temperature_k
    rise = peak["fine"] - AMBIENT
    graded_error = abs(peak["graded"] - peak["fine"])
    assert rise > 5.0
    assert grids["graded"].size < 0.2 * grids["fine"].size
    assert graded_error < abs(peak["coarse"] - peak["fine"])
    assert graded_error < 1e-3 * rise


@pytest.mark.parametrize("preconditioner", ["jacobi", "two-level"])
def test_refined_board_matches_the_oracle_direct_solve(preconditioner):
    grid = refined_grid((0.0, 12e-3), (0.0, 8e-3), coarse_pitch_m=1e-3, fine_pitch_m=0.25e-3,
                        refine_boxes_m=[(5e-3, 7e-3, 3e-3, 5e-3), (1e-3, 1.5e-3, 6e-3, 7e-3)], margin_m=0.5e-3)
    problem = _hot_spot_board(grid)
    print(f"grid refined board {grid.shape} {preconditioner}", flush=True)
    solution = solve_thermal_conduction(problem, preconditioner=preconditioner,
                                        config=MPIRConfig(relative_tolerance=1e-12, max_outer_iterations=40,
                                                          max_inner_iterations=20000))
    op = to_oracle(problem)
    ops = P.operator_arrays(op)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], op.element_shape)
    reference = P.default_reference(ops, op)
    exact = A.solve_direct(a, P.build_rhs(ops, reference, lambda v: k @ v)).reshape(problem.mesh.node_shape)
    rise = solution.temperature_k - reference
    assert solution.solve.converged
    assert np.max(np.abs(rise - exact)) <= 1e-9 * np.max(np.abs(exact))


def _hot_spot_board(grid: TensorGrid) -> ThermalConductionProblem:
    mesh = LayeredThermalMesh((35e-6, 0.8e-3, 35e-6), grid.pitch_x_m, grid.pitch_y_m, (385.0, 0.3, 385.0),
                              through_plane_conductivity_w_per_m_k=(385.0, 0.3, 385.0))
    heat = np.zeros(mesh.element_grid_shape)
    heat[2] = grid.cell_overlap_fraction(5e-3, 7e-3, 3e-3, 5e-3) * grid.cell_area_m2 / 4e-6 * 0.5
    return ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 25.0, AMBIENT),),
                                    fixed_temperature_mask=_bottom(mesh), fixed_temperature_k=AMBIENT,
                                    element_heat_w=heat)


def _bottom(mesh):
    mask = np.zeros(mesh.node_shape, dtype=bool)
    mask[0] = True
    return mask
