"""Backward-Euler transient conduction on the device: the reference's
closed-form checks (tests/test_thermal_transient.py:38-95: the discrete law
of a lumped block, a heated plate marching to its steady state) and every
step of a layered board against the oracle's FP64 direct solve of
(K + R + C/dt) T_{n+1} = q + R T_amb + (C/dt) T_n."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi
from tests.helpers import to_oracle
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    LayeredThermalMesh,
    ThermalConductionProblem,
    TimeSchedule,
    TransientThermalSolution,
    solve_thermal_conduction,
    solve_thermal_transient,
)

pytestmark = pytest.mark.gpu

AMBIENT = 300.0
COPPER_RHO_C = 3.45e6


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def test_lumped_block_cooling_follows_the_discrete_backward_euler_law():
    print("transient: lumped block", flush=True)
    thickness, h, start = 2.0e-3, 20.0, 350.0
    mesh = LayeredThermalMesh((thickness,), 1e-3, 1e-3, 4000.0, element_shape=(4, 5),
                              volumetric_heat_capacity_j_per_m3_k=COPPER_RHO_C)
    problem = ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", h, AMBIENT),))
    schedule = TimeSchedule.uniform(20.0, 400.0)
    solution = solve_thermal_transient(problem, schedule, initial_temperature_k=start, preconditioner="jacobi")
    assert isinstance(solution, TransientThermalSolution)
    assert solution.temperature_k.shape == (21, *mesh.node_shape)
    np.testing.assert_allclose(solution.temperature_k[0], start)
    tau = COPPER_RHO_C * thickness / h
    expected = start
    for step in solution.history:
        expected = AMBIENT + (expected - AMBIENT) / (1.0 + step.step_s / tau)
        assert step.max_temperature_k == pytest.approx(expected, abs=2.0e-4)
        assert step.converged
        assert step.stored_heat_w == pytest.approx(-step.convective_heat_w, rel=1e-9)
        assert abs(step.heat_balance_error_w) < 1e-9
    analytic = AMBIENT + (start - AMBIENT) * np.exp(-schedule.end_s / tau)
    assert abs(solution.history[-1].max_temperature_k - analytic) < 0.6
    finer = solve_thermal_transient(problem, TimeSchedule.uniform(5.0, 400.0), initial_temperature_k=start,
                                    store="final", preconditioner="jacobi")
    assert finer.temperature_k.shape == (1, *mesh.node_shape)
    assert abs(finer.history[-1].max_temperature_k - analytic) < 0.16


def test_heated_plate_marches_to_the_steady_solution_and_stops():
    print("transient: heated plate", flush=True)
    mesh = LayeredThermalMesh(
        (35e-6, 1.5e-3, 35e-6), 0.5e-3, 0.5e-3, (385.0, 0.8, 385.0), element_shape=(8, 12),
        through_plane_conductivity_w_per_m_k=(385.0, 0.3, 385.0),
        volumetric_heat_capacity_j_per_m3_k=(COPPER_RHO_C, 1.8e6, COPPER_RHO_C),
    )
    heat = np.zeros((3, 8, 12))
    heat[2, 3:5, 5:7] = 0.1
    problem = ThermalConductionProblem(
        mesh, convection=(ConvectionBoundary("top", 10.0, AMBIENT), ConvectionBoundary("bottom", 10.0, AMBIENT)),
        element_heat_w=heat,
    )
    steady = solve_thermal_conduction(problem, preconditioner="jacobi")
    schedule = TimeSchedule.geometric(0.01, 20000.0, growth=1.6)
    transient = solve_thermal_transient(problem, schedule, until_steady=True, steady_tolerance_k_per_s=2e-5,
                                        preconditioner="jacobi")
    assert transient.reached_steady
    assert len(transient.history) < len(schedule.steps_s)
    rise = steady.max_temperature_k - AMBIENT
    assert rise > 5.0
    assert np.nanmax(np.abs(transient.final_temperature_k - steady.temperature_k)) < 2e-3 * rise
    peaks = transient.max_temperature_k
    assert peaks[0] > AMBIENT and np.all(np.diff(peaks) >= -1e-9)
    for step in transient.history:
        assert abs(step.heat_balance_error_w) < 1e-9
        assert step.stored_heat_w + step.convective_heat_w == pytest.approx(0.4, rel=1e-9)


@pytest.mark.parametrize("preconditioner", ["jacobi", "two-level"])
def test_every_step_matches_the_oracle_direct_solve(preconditioner):
    print(f"transient: steps vs oracle ({preconditioner})", flush=True)
    rng = np.random.default_rng(31)
    k = np.where(rng.random((3, 10, 14)) < 0.4, 385.0, 0.3)
    mesh = LayeredThermalMesh((0.5e-3,) * 3, 0.25e-3, 0.25e-3, k,
                              volumetric_heat_capacity_j_per_m3_k=np.where(k > 1.0, COPPER_RHO_C, 1.8e6))
    heat = np.zeros((3, 10, 14))
    heat[2, 4:7, 5:9] = 0.05
    fixed = np.zeros(mesh.node_shape, dtype=bool)
    fixed[0] = True
    problem = ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 25.0, AMBIENT),),
                                       fixed_temperature_mask=fixed, fixed_temperature_k=AMBIENT,
                                       element_heat_w=heat)
    schedule = TimeSchedule((0.0, 0.05, 0.1, 0.3, 0.7, 1.5))
    run = solve_thermal_transient(problem, schedule, preconditioner=preconditioner)
    oracle = to_oracle(problem)
    ops = P.operator_arrays(oracle)
    capacity = mesh.nodal_heat_capacity_j_per_k().reshape(-1)
    load = P.nodal_load(oracle)
    free = ops["free"]
    previous = run.temperature_k[0].reshape(-1)
    for n, dt in enumerate(schedule.steps_s, start=1):
        robin = ops["robin_total"] + np.where(free, capacity / dt, 0.0)
        a = A.system_matrix(ops["coef"], ops["unit"], robin, free, oracle.element_shape)
        # Rise above the ambient: the Robin and fixed terms vanish, the
        # capacity term carries the previous step.
        rhs = np.where(free, load + capacity / dt * (previous - AMBIENT), 0.0)
        exact = A.solve_direct(a, rhs) + AMBIENT
        got = run.temperature_k[n].reshape(-1)
        scale = np.max(np.abs(exact - AMBIENT))
        assert np.max(np.abs(got - exact)) <= 1e-9 * scale, n
        previous = exact
