"""Transient schedule and argument validation on the CPU (the reference's
tests/test_thermal_transient.py:24-35 and 124-140; nothing here touches the
device: every check fires before an operator is built)."""

from __future__ import annotations

import numpy as np
import pytest

from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    LayeredThermalMesh,
    ThermalConductionProblem,
    TimeSchedule,
    solve_thermal_transient,
)


def test_time_schedules():
    uniform = TimeSchedule.uniform(0.5, 2.0)
    assert uniform.times_s == (0.0, 0.5, 1.0, 1.5, 2.0)
    geometric = TimeSchedule.geometric(1.0, 100.0, growth=2.0)
    assert geometric.times_s[:4] == (0.0, 1.0, 3.0, 7.0) and geometric.end_s == 100.0
    assert np.all(np.diff(geometric.times_s) > 0.0)
    capped = TimeSchedule.geometric(1.0, 20.0, growth=2.0, max_step_s=4.0)
    assert max(capped.steps_s) == 4.0
    assert TimeSchedule.uniform(0.3, 1.0).end_s == pytest.approx(1.2)
    with pytest.raises(ValueError, match="start at 0"):
        TimeSchedule((1.0, 2.0))
    with pytest.raises(ValueError, match="increasing"):
        TimeSchedule((0.0, 2.0, 1.0))
    with pytest.raises(ValueError, match="positive"):
        TimeSchedule.uniform(0.0, 1.0)
    with pytest.raises(ValueError, match="growth"):
        TimeSchedule.geometric(1.0, 10.0, growth=0.5)


def test_transient_validation():
    mesh = LayeredThermalMesh((1e-3,), 1e-3, 1e-3, 200.0, element_shape=(2, 2))
    problem = ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 5.0, 300.0),))
    with pytest.raises(ValueError, match="volumetric_heat_capacity"):
        solve_thermal_transient(problem, TimeSchedule.uniform(1.0, 2.0))
    with pytest.raises(ValueError, match="positive on active"):
        LayeredThermalMesh((1e-3,), 1e-3, 1e-3, 200.0, element_shape=(2, 2), volumetric_heat_capacity_j_per_m3_k=0.0)
    with_capacity = LayeredThermalMesh((1e-3,), 1e-3, 1e-3, 200.0, element_shape=(2, 2),
                                       volumetric_heat_capacity_j_per_m3_k=1e6)
    np.testing.assert_allclose(with_capacity.nodal_heat_capacity_j_per_k().sum(), 1e6 * 4e-9)
    capacitive = ThermalConductionProblem(with_capacity, convection=(ConvectionBoundary("top", 5.0, 300.0),))
    with pytest.raises(ValueError, match="store"):
        solve_thermal_transient(capacitive, TimeSchedule.uniform(1.0, 2.0), store="none")
    with pytest.raises(ValueError, match="steady_tolerance"):
        solve_thermal_transient(capacitive, TimeSchedule.uniform(1.0, 2.0), steady_tolerance_k_per_s=0.0)
