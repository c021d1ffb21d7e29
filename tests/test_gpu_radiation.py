"""Radiating faces on the device (the reference's
tests/test_thermal_radiation.py:37-150 and tests/test_thermal_transient.py:98-121):
the closed-form surface temperature of a uniformly heated plate, fast Newton
convergence, the shared convection/radiation budget, a masked body radiating
from its exposed faces, and a transient relaxing to the radiating steady state."""

from __future__ import annotations

import numpy as np
import pytest

from pcbhip import _capi
from thermal.matrix_free_mpir_fem import (
    STEFAN_BOLTZMANN_W_PER_M2_K4 as SIGMA,
    ConvectionBoundary,
    ExposedFaceConvection,
    ExposedFaceRadiation,
    LayeredThermalMesh,
    RadiationBoundary,
    ThermalConductionProblem,
    TimeSchedule,
    solve_thermal_conduction,
    solve_thermal_transient,
)

pytestmark = pytest.mark.gpu

ROWS, COLS, PITCH, THICKNESS, K = 8, 10, 1.0e-3, 1.0e-3, 200.0
AREA = ROWS * COLS * PITCH * PITCH
AMBIENT = 300.0


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _plate(power_w):
    mesh = LayeredThermalMesh((THICKNESS,), PITCH, PITCH, K, element_shape=(ROWS, COLS))
    return mesh, np.full((1, ROWS, COLS), power_w / (ROWS * COLS))


def _solve(problem, **kw):
    return solve_thermal_conduction(problem, preconditioner="jacobi", **kw)


def test_two_level_newton_lands_on_the_jacobi_answer():
    print("radiation two-level vs jacobi", flush=True)
    mesh, heat = _plate(0.6)
    problem = ThermalConductionProblem(
        mesh, radiation=(RadiationBoundary("top", 0.8, AMBIENT),), element_heat_w=heat,
        convection=(ConvectionBoundary("bottom", 5.0, AMBIENT),))
    jacobi = _solve(problem)
    two = solve_thermal_conduction(problem)
    assert two.radiation_converged and jacobi.radiation_converged
    np.testing.assert_allclose(two.temperature_k, jacobi.temperature_k, rtol=1e-9, atol=1e-7)
    assert two.heat_balance_error_w == pytest.approx(0.0, abs=1e-9)


@pytest.mark.parametrize("power_w, emissivity", [(0.05, 0.9), (0.6, 0.8), (3.0, 0.8)])
@pytest.mark.parametrize("kind", ["top", "exposed"])
def test_heated_plate_reaches_the_closed_form_surface_temperature(kind, power_w, emissivity):
    print(f"radiation plate {kind} {power_w} {emissivity}", flush=True)
    mesh, heat = _plate(power_w)
    radiation = (RadiationBoundary("top", emissivity, AMBIENT) if kind == "top"
                 else ExposedFaceRadiation(emissivity, AMBIENT, directions=("+z",)))
    solution = _solve(ThermalConductionProblem(mesh, radiation=(radiation,), element_heat_w=heat))
    surface = (AMBIENT**4 + power_w / (emissivity * SIGMA * AREA)) ** 0.25
    drop = power_w / AREA * THICKNESS / (2.0 * K)
    assert solution.radiation_converged and solution.solve.converged
    assert solution.radiation_iterations <= 15
    np.testing.assert_allclose(solution.temperature_k[1], surface, rtol=1e-7)
    np.testing.assert_allclose(solution.temperature_k[0], surface + drop, rtol=1e-7)
    assert solution.radiative_heat_w.shape == (1,) and solution.convective_heat_w.shape == (0,)
    assert solution.radiative_heat_w[0] == pytest.approx(power_w, rel=1e-9)
    assert abs(solution.heat_balance_error_w) < 1e-9 * max(power_w, 1.0)


def test_newton_converges_in_a_few_steps_and_in_one_from_the_answer():
    print("radiation newton", flush=True)
    mesh, heat = _plate(0.05)
    problem = ThermalConductionProblem(mesh, radiation=(RadiationBoundary("top", 0.8, AMBIENT),), element_heat_w=heat)
    solution = _solve(problem)
    assert 40.0 < solution.max_temperature_k - AMBIENT < 120.0
    assert solution.radiation_iterations <= 6
    again = _solve(problem, initial_temperature_k=solution.temperature_k)
    assert again.radiation_iterations == 1


def test_convection_and_radiation_share_the_budget():
    print("radiation + convection", flush=True)
    mesh, heat = _plate(0.6)
    h, emissivity = 8.0, 0.85
    solution = _solve(ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", h, AMBIENT),),
                                               radiation=(RadiationBoundary("top", emissivity, AMBIENT),),
                                               element_heat_w=heat))
    surface = float(np.mean(solution.temperature_k[1]))
    convective = h * AREA * (surface - AMBIENT)
    radiative = emissivity * SIGMA * AREA * (surface**4 - AMBIENT**4)
    assert solution.convective_heat_w[0] == pytest.approx(convective, rel=1e-6)
    assert solution.radiative_heat_w[0] == pytest.approx(radiative, rel=1e-6)
    assert convective + radiative == pytest.approx(0.6, rel=1e-6)
    only = _solve(ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", h, AMBIENT),),
                                           element_heat_w=heat))
    assert AMBIENT < solution.max_temperature_k < only.max_temperature_k


def test_exposed_face_convection_takes_per_element_values():
    print("exposed-face convection per element", flush=True)
    mesh, heat = _plate(0.6)
    coefficient = np.zeros(mesh.element_grid_shape)
    coefficient[0, :4] = 10.0
    ambient = np.full(mesh.element_grid_shape, AMBIENT)
    ambient[0, :4] = 290.0
    per_element = ExposedFaceConvection(coefficient, ambient)
    weights, rhs = per_element.lumped_nodal_weights(mesh)
    all_weights, _ = ExposedFaceConvection(10.0, 290.0).lumped_nodal_weights(mesh)
    assert per_element.cools and per_element.mean_ambient_k() == pytest.approx(295.0)
    assert 0.0 < float(np.sum(weights)) < float(np.sum(all_weights))
    np.testing.assert_allclose(rhs, weights * 290.0)
    solution = _solve(ThermalConductionProblem(mesh, convection=(per_element,), element_heat_w=heat))
    assert solution.solve.converged and abs(solution.heat_balance_error_w) < 1e-9


def test_masked_body_radiates_from_its_exposed_faces():
    print("radiation masked body", flush=True)
    active = np.zeros((4, 6, 6), dtype=bool)
    active[1:3, 2:4, 2:4] = True
    mesh = LayeredThermalMesh((1e-3,) * 4, 1e-3, 1e-3, np.where(active, 200.0, 1.0), active=active)
    heat = np.where(active, 0.5 / 8, 0.0)
    solution = _solve(ThermalConductionProblem(mesh, radiation=(ExposedFaceRadiation(0.9, AMBIENT),),
                                               element_heat_w=heat))
    area = 6 * (2e-3) ** 2
    surface = (AMBIENT**4 + 0.5 / (0.9 * SIGMA * area)) ** 0.25
    assert solution.radiation_converged
    assert solution.radiative_heat_w[0] == pytest.approx(0.5, rel=1e-9)
    assert np.nanmin(solution.temperature_k) > surface - 0.5
    assert np.nanmax(solution.temperature_k) < surface + 0.5
    assert np.all(np.isnan(solution.temperature_k[0])) and np.all(np.isnan(solution.temperature_k[:, 0]))


def test_radiating_masked_body_relaxes_to_its_steady_state():
    print("radiation transient", flush=True)
    active = np.zeros((3, 5, 5), dtype=bool)
    active[:, 1:4, 1:4] = True
    mesh = LayeredThermalMesh((1e-3,) * 3, 1e-3, 1e-3, np.where(active, 200.0, 1.0), active=active,
                              volumetric_heat_capacity_j_per_m3_k=np.where(active, 2.4e6, 1.0))
    heat = np.where(active, 0.3 / 27, 0.0)
    problem = ThermalConductionProblem(mesh, convection=(ExposedFaceConvection(5.0, AMBIENT),),
                                       radiation=(RadiationBoundary("top", 0.9, AMBIENT),), element_heat_w=heat)
    steady = _solve(problem)
    transient = solve_thermal_transient(problem, TimeSchedule.geometric(0.1, 20000.0, growth=1.7), until_steady=True,
                                        steady_tolerance_k_per_s=1e-5, preconditioner="jacobi")
    assert transient.reached_steady
    assert np.all(np.isnan(transient.final_temperature_k[:, 0]))
    assert np.nanmax(np.abs(transient.final_temperature_k - steady.temperature_k)) < 1e-2
    last = transient.history[-1]
    assert last.radiation_iterations >= 1 and last.radiative_heat_w > 0.0
    assert last.stored_heat_w + last.convective_heat_w + last.radiative_heat_w == pytest.approx(0.3, rel=1e-9)
