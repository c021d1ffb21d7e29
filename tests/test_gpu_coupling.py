"""The electro-thermal loop on device solves, on the reference's own checks
(tests/test_staggered_coupling.py:120-258 and
tests/test_electrothermal_coupling.py:61-121): closed-form isothermal
current and voltage drives, settling below the cold loss under a voltage
drive, one pass without a temperature coefficient, Aitken saving iterations
at a self-consistent state, a reported non-convergence, and the heat budget
of the coupled thermal solve."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import (
    CurrentTerminal,
    LayeredPCBMesh,
    PCBConductionProblem,
    ViaConnection,
    VoltageTerminal,
    solve_pcb_dc,
)
from multiphysics.staggered_coupling import (
    CouplingConfig,
    ElectroThermalScenario,
    conductivity_at_temperature,
    electrical_layer_temperature_k,
    run_electro_thermal,
)
from pcbhip import _capi
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    LayeredThermalMesh,
    ThermalConductionProblem,
    element_joule_heat_w,
    solve_thermal_conduction,
    via_joule_heat_sources,
)

pytestmark = pytest.mark.gpu

AMBIENT = 298.15
PITCH = 0.5e-3
T_REF = 293.15
ALPHA = 3.93e-3


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _loop(current_a: float = 2.0, rows: int = 6, cols: int = 24) -> PCBConductionProblem:
    """Two copper layers joined by vias at the far end: a current loop."""

    active = np.zeros((2, rows, cols), dtype=bool)
    active[:, 1:5, :] = True
    mesh = LayeredPCBMesh(active, (35e-6, 35e-6), PITCH, PITCH)
    pad = tuple(range(1, 6))
    vias = tuple(ViaConnection((0, r, c), (1, r, c), 1.5e-3) for r in pad[1:-1] for c in (cols - 1, cols))
    return PCBConductionProblem(
        mesh, (CurrentTerminal(tuple((1, r, 0) for r in pad), current_a, "source"),
               CurrentTerminal(tuple((0, r, 0) for r in pad), -current_a, "return")),
        reference_node=(0, 1, 0), vias=vias)


def _voltage_loop(voltage_v: float) -> PCBConductionProblem:
    current = _loop()
    source, sink = current.terminals
    return PCBConductionProblem(current.mesh, voltage_terminals=(VoltageTerminal(source.nodes, voltage_v, "source"),
                                                                 VoltageTerminal(sink.nodes, 0.0, "return")),
                                vias=current.vias)


def _cold_voltage(current_a: float) -> float:
    probe = solve_pcb_dc(_voltage_loop(1.0))
    return current_a / float(probe.voltage_terminal_current_a[0])


def _thermal(problem: PCBConductionProblem) -> LayeredThermalMesh:
    active = problem.mesh.element_active
    _, rows, cols = active.shape
    k = np.full((3, rows, cols), 0.8)
    k[0] = np.where(active[0], 385.0, 0.8)
    k[2] = np.where(active[1], 385.0, 0.8)
    kz = k.copy()
    kz[1] = 0.3
    return LayeredThermalMesh((35e-6, 1.5e-3, 35e-6), PITCH, PITCH, k, through_plane_conductivity_w_per_m_k=kz)


FILMS = (ConvectionBoundary("top", 10.0, AMBIENT), ConvectionBoundary("bottom", 10.0, AMBIENT))


def _scenario(cold_current_a: float = 2.0, **kw) -> ElectroThermalScenario:
    problem = _voltage_loop(_cold_voltage(cold_current_a))
    return ElectroThermalScenario(problem, _thermal(problem), (0, 2), convection=FILMS,
                                  conductivity_reference_temperature_k=T_REF, **kw)


def _current_scenario(current_a: float) -> ElectroThermalScenario:
    problem = _loop(current_a)
    with pytest.warns(DeprecationWarning, match="current-driven"):
        return ElectroThermalScenario(problem, _thermal(problem), (0, 2), convection=FILMS,
                                      conductivity_reference_temperature_k=T_REF)


def test_isothermal_current_drive_follows_the_resistivity_law():
    print("coupling isothermal current", flush=True)
    problem = _loop()
    thermal = _thermal(problem)
    hot = 353.15
    with pytest.warns(DeprecationWarning, match="current-driven"):
        scenario = ElectroThermalScenario(problem, thermal, (0, 2), fixed_temperature_mask=np.ones(thermal.node_shape,
                                                                                                   dtype=bool),
                                          fixed_temperature_k=hot, conductivity_reference_temperature_k=T_REF)
    result = run_electro_thermal(scenario)
    factor = 1.0 + ALPHA * (hot - T_REF)
    cold = solve_pcb_dc(problem)
    assert result.converged and result.iterations == 3  # cold pass, hot pass, confirming pass
    assert result.electrical.joule_loss_w == pytest.approx(cold.joule_loss_w * factor, rel=1e-9)
    assert result.loss_increase_ratio == pytest.approx(factor, rel=1e-9)
    np.testing.assert_allclose(result.element_temperature_k, hot)


def test_isothermal_voltage_drive_loses_v_squared_over_r_hot():
    print("coupling isothermal voltage", flush=True)
    problem = _voltage_loop(20e-3)
    thermal = _thermal(problem)
    hot = 353.15
    scenario = ElectroThermalScenario(problem, thermal, (0, 2),
                                      fixed_temperature_mask=np.ones(thermal.node_shape, dtype=bool),
                                      fixed_temperature_k=hot, conductivity_reference_temperature_k=T_REF)
    result = run_electro_thermal(scenario)
    factor = 1.0 + ALPHA * (hot - T_REF)
    cold = solve_pcb_dc(problem)
    assert result.converged
    assert result.electrical.joule_loss_w == pytest.approx(cold.joule_loss_w / factor, rel=1e-9)
    np.testing.assert_allclose(result.electrical.voltage_terminal_current_a, cold.voltage_terminal_current_a / factor,
                               rtol=1e-9)


def test_voltage_drive_settles_below_the_cold_loss():
    print("coupling voltage settles", flush=True)
    cold_current = 2.0
    voltage = _cold_voltage(cold_current)
    result = run_electro_thermal(_scenario(cold_current))
    runaway = run_electro_thermal(_current_scenario(cold_current))
    assert result.converged
    assert result.cold_joule_loss_w == pytest.approx(runaway.cold_joule_loss_w, rel=1e-2)
    assert result.loss_increase_ratio < 0.97 and runaway.loss_increase_ratio > 1.05
    assert result.thermal.max_temperature_k > AMBIENT + 5.0
    assert result.thermal.total_heat_input_w == pytest.approx(result.electrical.joule_loss_w, rel=1e-12)
    supplied = float(result.electrical.voltage_terminal_current_a[0])
    assert supplied < cold_current
    # KCL of the two pad currents holds to the solve's 1e-10 relative residual
    # of the lifted system: a few nA of the 1.8 A drive.
    assert float(np.sum(result.electrical.voltage_terminal_current_a)) == pytest.approx(0.0, abs=1e-8)
    assert result.electrical.joule_loss_w == pytest.approx(voltage * supplied, rel=1e-8)


def test_zero_temperature_coefficient_needs_one_pass():
    print("coupling zero coefficient", flush=True)
    result = run_electro_thermal(_scenario(temperature_coefficient_per_k=0.0))
    assert result.converged and result.iterations == 1 and result.loss_increase_ratio == 1.0
    np.testing.assert_allclose(result.conductivity_s_per_m, _voltage_loop(1e-3).mesh.conductivity_s_per_m)


def test_coupled_state_is_self_consistent_and_aitken_saves_iterations():
    print("coupling aitken", flush=True)
    scenario = _scenario(3.0)
    fast = run_electro_thermal(scenario)
    plain = run_electro_thermal(scenario, config=CouplingConfig(aitken=False))
    assert fast.converged and plain.converged
    assert fast.iterations < plain.iterations
    assert fast.loss_increase_ratio < 0.9 and fast.thermal.max_temperature_k > AMBIENT + 10.0
    np.testing.assert_allclose(fast.electrical.joule_loss_w, plain.electrical.joule_loss_w, rtol=1e-6)
    layer_t = electrical_layer_temperature_k(fast.thermal.temperature_k, scenario.thermal_mesh, scenario.layer_slabs)
    expected = conductivity_at_temperature(scenario.electrical.mesh.conductivity_s_per_m, layer_t,
                                           reference_temperature_k=T_REF, coefficient_per_k=ALPHA)
    np.testing.assert_allclose(fast.conductivity_s_per_m, expected, rtol=1e-6)
    assert fast.thermal.total_heat_input_w == pytest.approx(fast.electrical.joule_loss_w, rel=1e-12)
    assert fast.history[-1].temperature_change_k <= 1e-3
    assert all(step.relaxation > 0.0 for step in fast.history)


def test_iteration_limit_is_reported_not_raised():
    print("coupling iteration limit", flush=True)
    result = run_electro_thermal(_current_scenario(6.0), config=CouplingConfig(max_iterations=3))
    assert not result.converged and result.iterations == 3
    assert result.electrical.joule_loss_w > result.cold_joule_loss_w


def test_voltage_drive_converges_where_current_drive_runs_away():
    print("coupling 6 A", flush=True)
    result = run_electro_thermal(_scenario(6.0))
    assert result.converged and result.iterations <= 10
    assert result.loss_increase_ratio < 1.0
    assert float(result.electrical.voltage_terminal_current_a[0]) < 6.0


def _two_layer():
    rows, cols = 2, 6
    active = np.zeros((2, rows, cols), dtype=bool)
    active[0, :, :4] = True
    active[1, :, 2:] = True
    mesh = LayeredPCBMesh(active, (35e-6, 35e-6), 0.5e-3, 0.5e-3)
    problem = PCBConductionProblem(
        mesh, terminals=(CurrentTerminal(tuple((0, r, 0) for r in range(rows + 1)), 3.0, "in"),
                         CurrentTerminal(tuple((1, r, cols) for r in range(rows + 1)), -3.0, "out")),
        reference_node=(1, 0, cols),
        vias=(ViaConnection((0, 0, 3), (1, 0, 3), 2e-3), ViaConnection((0, 2, 3), (1, 2, 3), 2e-3)))
    k = np.full((3, rows, cols), 0.8)
    k[0] = np.where(active[0], 385.0, 0.8)
    k[2] = np.where(active[1], 385.0, 0.8)
    kz = k.copy()
    kz[1] = 0.3
    return problem, LayeredThermalMesh((35e-6, 1.5e-3, 35e-6), 0.5e-3, 0.5e-3, k, through_plane_conductivity_w_per_m_k=kz)


def test_coupled_thermal_solve_removes_exactly_the_electrical_loss():
    print("coupling heat budget", flush=True)
    problem, thermal = _two_layer()
    electrical = solve_pcb_dc(problem)
    assert electrical.joule_loss_w == pytest.approx(float(electrical.element_joule_loss_w.sum())
                                                    + float(electrical.via_joule_loss_w.sum()))
    # By symmetry each via passes 1.5 A through 2 mOhm.
    np.testing.assert_allclose(electrical.via_joule_loss_w, 1.5 ** 2 * 2e-3, rtol=1e-6)
    films = (ConvectionBoundary("top", 15.0, AMBIENT), ConvectionBoundary("bottom", 15.0, AMBIENT))
    heated = ThermalConductionProblem(thermal, convection=films,
                                      element_heat_w=element_joule_heat_w(electrical, thermal, (0, 2)),
                                      heat_sources=via_joule_heat_sources(problem, electrical, thermal, (0, 2)))
    solution = solve_thermal_conduction(heated, initial_temperature_k=AMBIENT)
    assert solution.solve.converged
    assert solution.total_heat_input_w == pytest.approx(electrical.joule_loss_w)
    assert float(np.sum(solution.convective_heat_w)) == pytest.approx(electrical.joule_loss_w, rel=1e-8)
    assert solution.min_temperature_k > AMBIENT


def test_many_loops_served_together_equal_each_loop_alone():
    print("coupling served loops", flush=True)
    from multiphysics.staggered_coupling import run_electro_thermal_many

    scenarios = [_scenario(current) for current in (1.0, 2.5, 4.0)]
    alone = [run_electro_thermal(s) for s in scenarios]
    served = run_electro_thermal_many(scenarios)
    for a, b in zip(alone, served):
        assert a.converged and b.converged and a.iterations == b.iterations
        assert np.array_equal(a.electrical.potential_v, b.electrical.potential_v, equal_nan=True)
        assert np.array_equal(a.thermal.temperature_k, b.thermal.temperature_k, equal_nan=True)
        assert [s.thermal_inner_iterations for s in a.history] == [s.thermal_inner_iterations for s in b.history]
        assert a.electrical.joule_loss_w == b.electrical.joule_loss_w
