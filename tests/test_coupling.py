"""Host-side parts of the electro-thermal coupling, on the CPU: the loss
mapping onto thermal slabs and via ends (the reference's
tests/test_electrothermal_coupling.py:78-131 on a stand-in solution), the
resistivity law, slab and via temperatures, scenario and configuration
checks (tests/test_staggered_coupling.py:260-271), and the Aitken-relaxed
fixed point against closed forms."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import (
    CurrentTerminal,
    LayeredPCBMesh,
    PCBConductionProblem,
    ViaConnection,
    VoltageTerminal,
)
from multiphysics.staggered_coupling import (
    CouplingConfig,
    ElectroThermalScenario,
    TemperatureFixedPoint,
    conductivity_at_temperature,
    electrical_layer_temperature_k,
    slab_element_temperature_k,
    via_temperature_k,
)
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    LayeredThermalMesh,
    element_heat_from_losses,
    via_joule_heat_sources,
)

ROWS, COLS = 2, 6


def _board():
    active = np.zeros((2, ROWS, COLS), dtype=bool)
    active[0, :, :4] = True
    active[1, :, 2:] = True
    mesh = LayeredPCBMesh(active, (35e-6, 35e-6), 0.5e-3, 0.5e-3)
    vias = (ViaConnection((0, 0, 3), (1, 0, 3), 2e-3), ViaConnection((0, 2, 3), (1, 2, 3), 2e-3))
    problem = PCBConductionProblem(
        mesh, terminals=(CurrentTerminal(tuple((0, r, 0) for r in range(ROWS + 1)), 3.0, "in"),
                         CurrentTerminal(tuple((1, r, COLS) for r in range(ROWS + 1)), -3.0, "out")),
        reference_node=(1, 0, COLS), vias=vias)
    thermal = LayeredThermalMesh((35e-6, 1.5e-3, 35e-6), 0.5e-3, 0.5e-3, (385.0, 0.3, 385.0),
                                 element_shape=(ROWS, COLS))
    return problem, thermal


def test_element_losses_land_on_their_copper_slabs():
    _, thermal = _board()
    loss = np.arange(2 * ROWS * COLS, dtype=float).reshape(2, ROWS, COLS)
    heat = element_heat_from_losses(loss, thermal, (0, 2))
    np.testing.assert_array_equal(heat[0], loss[0])
    np.testing.assert_array_equal(heat[1], 0.0)
    np.testing.assert_array_equal(heat[2], loss[1])
    assert heat.sum() == loss.sum()
    # Two layers in one slab add.
    np.testing.assert_array_equal(element_heat_from_losses(loss, thermal, (1, 1))[1], loss[0] + loss[1])
    with pytest.raises(ValueError, match="layer_slabs has"):
        element_heat_from_losses(loss, thermal, (0,))
    with pytest.raises(ValueError, match="must lie in"):
        element_heat_from_losses(loss, thermal, (0, 3))
    with pytest.raises(ValueError, match="footprint"):
        element_heat_from_losses(loss[:, :1], thermal, (0, 2))


def test_via_losses_split_between_the_faces_of_both_ends():
    problem, thermal = _board()
    solution = SimpleNamespace(via_joule_loss_w=np.array([0.4, 1.0]))
    sources = via_joule_heat_sources(problem, solution, thermal, (0, 2))
    assert [s.nodes for s in sources] == [((0, 0, 3), (1, 0, 3)), ((2, 0, 3), (3, 0, 3)),
                                         ((0, 2, 3), (1, 2, 3)), ((2, 2, 3), (3, 2, 3))]
    np.testing.assert_allclose([s.power_w for s in sources], [0.2, 0.2, 0.5, 0.5])
    with pytest.raises(ValueError, match="must lie in"):
        via_joule_heat_sources(problem, solution, thermal, (0, 3))


def test_resistivity_law_and_its_valid_range():
    sigma = conductivity_at_temperature(5.8e7, np.array([293.15, 393.15]), reference_temperature_k=293.15,
                                        coefficient_per_k=3.93e-3)
    np.testing.assert_allclose(sigma, [5.8e7, 5.8e7 / 1.393])
    with pytest.raises(ValueError, match="valid range"):
        conductivity_at_temperature(5.8e7, 0.0, reference_temperature_k=293.15, coefficient_per_k=3.93e-3)


def test_slab_and_via_temperatures_average_the_right_nodes():
    _, thermal = _board()
    # A field linear in (slab face, row, col): every average is the field at the centroid.
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float) for n in thermal.node_shape), indexing="ij")
    field = 300.0 + 5.0 * z + 2.0 * y + 0.5 * x
    centre = 300.0 + 5.0 * (2 + 0.5) + 2.0 * (np.arange(ROWS)[:, None] + 0.5) + 0.5 * (np.arange(COLS) + 0.5)
    np.testing.assert_allclose(slab_element_temperature_k(field, thermal, 2), centre)
    layers = electrical_layer_temperature_k(field, thermal, (0, 2))
    assert layers.shape == (2, ROWS, COLS)
    np.testing.assert_allclose(layers[1], centre)
    via = ViaConnection((0, 1, 3), (1, 1, 3), 1e-3)
    # Ends at faces (0.5) and (2.5) of layer slabs 0 and 2: mean face 1.5.
    np.testing.assert_allclose(via_temperature_k(field, thermal, (0, 2), (via,)), 300.0 + 5 * 1.5 + 2 + 1.5)


def test_scenario_and_config_are_validated():
    problem, thermal = _board()
    voltage = PCBConductionProblem(problem.mesh, voltage_terminals=(
        VoltageTerminal(problem.terminals[0].nodes, 1e-3, "in"), VoltageTerminal(problem.terminals[1].nodes, 0.0)),
        vias=problem.vias)
    film = (ConvectionBoundary("top", 10.0, 298.15),)
    ElectroThermalScenario(voltage, thermal, (0, 2), convection=film)
    with pytest.warns(DeprecationWarning, match="current-driven"):
        ElectroThermalScenario(problem, thermal, (0, 2), convection=film)
    with pytest.raises(ValueError, match="one thermal slab per electrical layer"):
        ElectroThermalScenario(voltage, thermal, (0,), convection=film)
    with pytest.raises(ValueError, match="share"):
        other = LayeredThermalMesh((35e-6,), 0.5e-3, 0.5e-3, 385.0, element_shape=(ROWS, COLS + 1))
        ElectroThermalScenario(voltage, other, (0, 0), convection=film)
    with pytest.raises(ValueError, match="non-negative"):
        ElectroThermalScenario(voltage, thermal, (0, 2), convection=film, temperature_coefficient_per_k=-1.0)
    for bad in (dict(max_iterations=0), dict(temperature_tolerance_k=0.0), dict(relaxation=5.0),
                dict(max_relaxation=0.5, relaxation=0.5)):
        with pytest.raises(ValueError):
            CouplingConfig(**bad)


def _fixed_point(aitken=True):
    scenario = SimpleNamespace(temperature_coefficient_per_k=3.93e-3, conductivity_reference_temperature_k=293.15)
    return TemperatureFixedPoint(scenario, CouplingConfig(aitken=aitken))


def test_aitken_solves_a_linear_contraction_in_one_relaxed_step():
    # T -> g(T) = a T + b: increments d_k shrink by (a - 1) w each step; the
    # Aitken factor is 1 / (1 - a), the step that lands on b / (1 - a).
    a, b = 0.6, 140.0  # factor 1 / (1 - a) = 2.5, inside the clip at 4
    exact = b / (1.0 - a)
    fp = _fixed_point()
    t = np.full(4, 300.0)
    assert fp.update(t) == float("inf")
    for _ in range(2):
        fp.update(a * fp.temperature + b)
    np.testing.assert_allclose(fp.temperature, exact, rtol=1e-12)
    assert fp.relaxation == pytest.approx(1.0 / (1.0 - a))
    # The plain iteration only approaches it.
    plain = _fixed_point(aitken=False)
    plain.update(t)
    for _ in range(2):
        plain.update(a * plain.temperature + b)
    assert np.all(np.abs(plain.temperature - exact) > 1.0)


def test_relaxed_steps_stay_inside_the_resistivity_law():
    fp = _fixed_point()
    fp.update(np.full(3, 400.0))
    fp.relaxation = 4.0
    fp.update(np.full(3, 300.0))   # 400 - 4 * 100 = 0 K would leave the law's range
    np.testing.assert_allclose(fp.temperature, 300.0)
    assert fp.relaxation == 1.0
