"""Shared builders for the tests: reference-style problems and their oracle twins."""

from __future__ import annotations

import numpy as np

from oracle import problem as oracle_problem
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    ExposedFaceConvection,
    HeatSource,
    LayeredThermalMesh,
    ThermalConductionProblem,
)

DOUBLE_EPS = float(np.finfo(np.float64).eps)
SINGLE_EPS = float(np.finfo(np.float32).eps)


def to_oracle(problem: ThermalConductionProblem) -> oracle_problem.OracleProblem:
    """The same problem as plain arrays for the oracle."""

    return oracle_problem.from_problem(problem)


def stack(rows: int, cols: int, *, fixed: bool = False, slabs: tuple = (35.0e-6, 1.5e-3, 35.0e-6)):
    """A copper/FR-4/copper stack with a heated trace, a point source, both
    faces cooled and optionally scattered fixed nodes (the reference's
    tests/test_native_thermal.py:26-52 fixture shape)."""

    n = len(slabs)
    k_in = [385.0 if i in (0, n - 1) else 0.8 for i in range(n)]
    k_z = [385.0 if i in (0, n - 1) else 0.3 for i in range(n)]
    mesh = LayeredThermalMesh(
        slab_thickness_m=slabs,
        pitch_x_m=0.2e-3,
        pitch_y_m=0.2e-3,
        conductivity_w_per_m_k=k_in,
        through_plane_conductivity_w_per_m_k=k_z,
        element_shape=(rows, cols),
    )
    heat = np.zeros(mesh.element_grid_shape)
    heat[-1, rows // 2, cols // 4 : 3 * cols // 4] = 0.5 / max(1, cols // 2)
    fixed_mask = np.zeros(mesh.node_shape, dtype=bool)
    fixed_k = np.full(mesh.node_shape, 298.15)
    if fixed:
        fixed_mask[0, ::3, 1::2] = True
        fixed_k[0] = 310.0
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", 10.0, 298.15), ConvectionBoundary("bottom", 10.0, 298.15)),
        element_heat_w=heat,
        fixed_temperature_mask=fixed_mask if fixed else None,
        fixed_temperature_k=fixed_k if fixed else None,
        heat_sources=(HeatSource(((1, rows // 3, cols // 3),), 0.05),),
    )


def copper_stack(rows: int, cols: int) -> ThermalConductionProblem:
    """The reference's two-level fixture (tests/test_thermal_conduction.py:
    238-257): copper/FR-4/copper, a heated trace on the top slab, both faces
    cooled, no fixed node."""

    mesh = LayeredThermalMesh(
        slab_thickness_m=(35.0e-6, 1.5e-3, 35.0e-6),
        pitch_x_m=0.2e-3,
        pitch_y_m=0.2e-3,
        conductivity_w_per_m_k=(385.0, 0.8, 385.0),
        through_plane_conductivity_w_per_m_k=(385.0, 0.3, 385.0),
        element_shape=(rows, cols),
    )
    heat = np.zeros(mesh.element_grid_shape)
    heat[2, rows // 2, cols // 4 : 3 * cols // 4] = 0.5 / (cols // 2)
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", 10.0, 298.15), ConvectionBoundary("bottom", 10.0, 298.15)),
        element_heat_w=heat,
    )


def random_body(shape, seed: int = 0, *, void: bool = False):
    """Per-element anisotropic conductivity over 5 decades, graded pitches,
    optional void elements, both face kinds of convection and fixed nodes."""

    rng = np.random.default_rng(seed)
    s, r, c = shape
    k = np.exp(rng.uniform(np.log(0.3), np.log(400.0), shape))
    active = None
    if void:
        active = rng.random(shape) > 0.2
        active[0, 0, 0] = True
    mesh = LayeredThermalMesh(
        slab_thickness_m=rng.uniform(30e-6, 1e-3, s),
        pitch_x_m=rng.uniform(0.1e-3, 0.5e-3, c),
        pitch_y_m=rng.uniform(0.1e-3, 0.5e-3, r),
        conductivity_w_per_m_k=k,
        through_plane_conductivity_w_per_m_k=k * rng.uniform(0.2, 1.0, shape),
        active=active,
    )
    fixed = np.zeros(mesh.node_shape, dtype=bool)
    fixed[0, ::2, ::3] = True
    fixed &= mesh.active_nodes
    if not fixed.any():
        fixed[0, 0, 0] = True
    heat = np.where(mesh.active, rng.uniform(0, 1e-3, shape), 0.0)
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", 25.0, 300.0), ExposedFaceConvection(5.0, 295.0, ("-x", "+y"))),
        fixed_temperature_mask=fixed,
        fixed_temperature_k=310.0,
        element_heat_w=heat,
    )
