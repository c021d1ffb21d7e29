"""The layered-sheet DC operator on the device: its actions against the
oracle's assembled matrix, and solve_pcb_dc on the reference's own checks
(tests/test_matrix_free_mpir_fem.py:84-226): closed-form strip resistance
and current density, the only via current of a two-layer stack, the
voltage-driven strip, the dense lifted system, warm starts and rejected
drives; plus many excitations of one sheet operator in one batch."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import (
    CurrentTerminal,
    LayeredPCBMesh,
    MatrixFreePCBOperator,
    MPIRConfig,
    PCBConductionProblem,
    ViaConnection,
    VoltageTerminal,
    solve_pcb_dc,
)
from oracle import sheet as O
from pcbhip import _capi
from tests.helpers import DOUBLE_EPS, SINGLE_EPS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _strip_problem(columns=4):
    mesh = LayeredPCBMesh(np.ones((1, 1, columns), dtype=bool), (35.0e-6,), 1.0e-3, 1.0e-3)
    return PCBConductionProblem(
        mesh=mesh,
        terminals=(CurrentTerminal(((0, 0, 0), (0, 1, 0)), 1.0, "source"),
                   CurrentTerminal(((0, 0, columns), (0, 1, columns)), -1.0, "sink")),
        reference_node=(0, 0, columns),
    )


def _voltage_strip_problem(columns=4, voltage_v=1.0e-3):
    mesh = _strip_problem(columns).mesh
    return PCBConductionProblem(mesh=mesh, voltage_terminals=(
        VoltageTerminal(((0, 0, 0), (0, 1, 0)), voltage_v, "source"),
        VoltageTerminal(((0, 0, columns), (0, 1, columns)), 0.0, "sink")))


SHAPES = [(1, 1, 1), (2, 3, 5), (1, 2, 61), (2, 5, 62), (3, 9, 63), (2, 17, 130)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_sheet_actions_match_the_assembled_operator(shape):
    print(f"sheet parity {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    layers, rows, cols = shape
    active = rng.random(shape) < 0.8
    mesh = LayeredPCBMesh(active, tuple(rng.uniform(18e-6, 70e-6, layers)), rng.uniform(0.3e-3, 1e-3, cols),
                          rng.uniform(0.3e-3, 1e-3, rows), rng.uniform(3e7, 6e7, shape))
    node_shape = mesh.node_shape
    n = mesh.size
    vias = []
    if layers > 1:
        for _ in range(min(8, rows * cols)):
            y, x = int(rng.integers(rows + 1)), int(rng.integers(cols + 1))
            l = int(rng.integers(layers - 1))
            via = ViaConnection((l, y, x), (l + 1, y, x), float(rng.uniform(1e-4, 1e-2)))
            if all((v.lower, v.upper) != (via.lower, via.upper) for v in vias):
                vias.append(via)
    operator = MatrixFreePCBOperator(mesh, reference_node=(0, 0, 0) if active[0, 0, 0] else None,
                                     dirichlet_nodes=() if active[0, 0, 0] else
                                     [tuple(int(i) for i in np.argwhere(active)[0])],
                                     vias=vias, preconditioner="jacobi")
    coef = np.stack(mesh.element_coefficients())
    via_a = [np.ravel_multi_index(v.lower, node_shape) for v in vias]
    via_b = [np.ravel_multi_index(v.upper, node_shape) for v in vias]
    via_g = [v.conductance_s for v in vias]
    k = O.stiffness(coef, shape, via_a, via_b, via_g)
    free = operator.free_nodes
    a = O.system_matrix(k, free)
    x = rng.standard_normal(n)
    scale = float((abs(a) @ np.abs(x)).max())
    y64 = operator.apply_high(x)
    np.testing.assert_allclose(y64, a @ x, rtol=0, atol=64 * DOUBLE_EPS * scale)
    x32 = x.astype(np.float32)
    y32 = operator.apply_low(x32)
    assert np.max(np.abs(y32 - a @ x32.astype(np.float64))) <= 32 * SINGLE_EPS * scale
    fixed = ~free.reshape(-1)
    np.testing.assert_array_equal(y32[fixed], x32[fixed])
    kx = operator.apply_full_high(x)
    np.testing.assert_allclose(kx, k @ x, rtol=0, atol=64 * DOUBLE_EPS * float((abs(k) @ np.abs(x)).max()))
    np.testing.assert_allclose(operator.device_operator.diagonal(), a.diagonal(), rtol=4 * DOUBLE_EPS, atol=0)
    assert operator.apply_low(x32).tobytes() == y32.tobytes()


def test_uniform_strip_matches_closed_form_resistance_and_current_density():
    print("pcb strip", flush=True)
    problem = _strip_problem(columns=4)
    solution = solve_pcb_dc(problem, preconditioner="jacobi")
    sigma = float(problem.mesh.conductivity_s_per_m[0, 0, 0])
    t = problem.mesh.layer_thickness_m[0]
    resistance = 4.0e-3 / (sigma * 1.0e-3 * t)
    assert solution.solve.converged
    assert solution.joule_loss_w == pytest.approx(resistance, rel=2.0e-10)
    assert solution.max_current_density_a_per_m2 == pytest.approx(1.0 / (1.0e-3 * t), rel=2.0e-10)
    np.testing.assert_allclose(solution.current_density_a_per_m2[..., 1], 0.0, atol=1.0e-9 / (1.0e-3 * t))


def test_layered_problem_reports_the_only_via_current():
    print("pcb via", flush=True)
    mesh = LayeredPCBMesh(np.ones((2, 1, 1), dtype=bool), (35.0e-6, 35.0e-6), 1.0e-3, 1.0e-3)
    problem = PCBConductionProblem(
        mesh=mesh, terminals=(CurrentTerminal(((0, 0, 0),), 1.0, "source"),
                              CurrentTerminal(((1, 0, 0),), -1.0, "sink")),
        reference_node=(1, 0, 0), vias=(ViaConnection((0, 0, 1), (1, 0, 1), 2.0e-3),))
    solution = solve_pcb_dc(problem, preconditioner="jacobi")
    assert solution.solve.converged
    assert abs(solution.via_current_a[0]) == pytest.approx(1.0, rel=2.0e-9)
    assert solution.via_joule_loss_w[0] == pytest.approx(2.0e-3, rel=2.0e-9)


def test_voltage_driven_strip_matches_closed_form_current():
    print("pcb voltage strip", flush=True)
    voltage = 1.0e-3
    problem = _voltage_strip_problem(columns=4, voltage_v=voltage)
    solution = solve_pcb_dc(problem, preconditioner="jacobi")
    sigma = float(problem.mesh.conductivity_s_per_m[0, 0, 0])
    current = voltage / (4.0e-3 / (sigma * 1.0e-3 * problem.mesh.layer_thickness_m[0]))
    assert solution.solve.converged
    np.testing.assert_allclose(solution.voltage_terminal_current_a, (current, -current), rtol=2.0e-10)
    assert solution.joule_loss_w == pytest.approx(voltage * current, rel=2.0e-10)
    np.testing.assert_allclose(solution.potential_v[0, :, 0], voltage)
    np.testing.assert_allclose(solution.potential_v[0, :, 4], 0.0)
    np.testing.assert_allclose(solution.potential_v[0, 0], voltage * np.linspace(1.0, 0.0, 5), atol=1.0e-12)


def test_voltage_driven_solve_matches_the_dense_lifted_system():
    print("pcb lifted system", flush=True)
    mesh = LayeredPCBMesh(np.ones((2, 3, 4), dtype=bool), (35.0e-6, 18.0e-6), (0.4e-3, 0.5e-3, 0.6e-3, 0.4e-3),
                          0.5e-3)
    problem = PCBConductionProblem(
        mesh=mesh, terminals=(CurrentTerminal(((0, 3, 2),), -0.25, "load"),),
        voltage_terminals=(VoltageTerminal(((0, 0, 0), (0, 1, 0)), 1.2, "vrm"),
                           VoltageTerminal(((1, 3, 4), (1, 2, 4)), 0.0, "return")),
        vias=(ViaConnection((0, 1, 2), (1, 1, 2), 1.0e-3),))
    solution = solve_pcb_dc(problem, config=MPIRConfig(relative_tolerance=1.0e-12), preconditioner="jacobi")
    operator = MatrixFreePCBOperator(mesh, dirichlet_nodes=[n for t in problem.voltage_terminals for n in t.nodes],
                                     vias=problem.vias, preconditioner="jacobi")
    identity = np.eye(operator.size)
    full = np.column_stack([operator.apply_full_high(c) for c in identity.T])
    fixed = operator.dirichlet_potential(problem.voltage_terminals)
    known = ~operator.free_nodes.reshape(-1)
    free = ~known
    injection = np.zeros(operator.size)
    injection[np.ravel_multi_index((0, 3, 2), mesh.node_shape)] = -0.25
    expected = fixed.copy()
    expected[free] = np.linalg.solve(full[np.ix_(free, free)], injection[free] - full[np.ix_(free, known)] @ fixed[known])
    assert solution.solve.converged
    np.testing.assert_allclose(solution.potential_v.reshape(-1), expected, rtol=1.0e-10, atol=1.0e-13)
    assert float(np.sum(solution.voltage_terminal_current_a)) == pytest.approx(0.25, rel=1.0e-9)
    supplied = float(np.dot(solution.voltage_terminal_current_a, (1.2, 0.0)))
    assert solution.joule_loss_w == pytest.approx(supplied - 0.25 * float(solution.potential_v[0, 3, 2]), rel=1.0e-9)


def test_voltage_driven_solve_warm_starts_from_its_own_potential():
    print("pcb warm start", flush=True)
    problem = _voltage_strip_problem(columns=8)
    cold = solve_pcb_dc(problem, preconditioner="jacobi")
    warm = solve_pcb_dc(problem, initial_potential_v=cold.potential_v, preconditioner="jacobi")
    assert warm.solve.converged
    assert warm.solve.inner_iterations < cold.solve.inner_iterations
    np.testing.assert_allclose(warm.potential_v, cold.potential_v, rtol=1.0e-10, atol=1.0e-15)


def test_inconsistent_drives_and_floating_nodes_are_rejected():
    mesh = _strip_problem(columns=1).mesh
    source = VoltageTerminal(((0, 0, 0),), 1.0, "source")
    with pytest.raises(ValueError, match="reference_node must be None"):
        PCBConductionProblem(mesh=mesh, voltage_terminals=(source,), reference_node=(0, 0, 1))
    with pytest.raises(ValueError, match="belongs to terminals"):
        PCBConductionProblem(mesh=mesh, terminals=(CurrentTerminal(((0, 0, 0),), 1.0, "load"),),
                             voltage_terminals=(source,))
    with pytest.raises(ValueError, match="needs a reference_node"):
        PCBConductionProblem(mesh=mesh, terminals=(CurrentTerminal(((0, 0, 0),), 1.0),
                                                   CurrentTerminal(((0, 0, 1),), -1.0)))
    with pytest.raises(ValueError, match="sum to zero"):
        PCBConductionProblem(mesh=mesh, terminals=(CurrentTerminal(((0, 0, 0),), 1.0),), reference_node=(0, 0, 0))
    # Node (0, 0, 3) touches no active element and no via: it is not copper.
    isolated = LayeredPCBMesh(np.array([[[True, False, False]]]), (35e-6,), 1e-3, 1e-3)
    with pytest.raises(ValueError, match="reference_node must lie on active copper"):
        MatrixFreePCBOperator(isolated, reference_node=(0, 0, 3), preconditioner="jacobi")


def test_many_excitations_of_one_sheet_operator_in_one_batch():
    print("pcb batch", flush=True)
    rng = np.random.default_rng(5)
    active = np.ones((2, 12, 30), dtype=bool)
    active[:, 4:8, 10:20] = False
    mesh = LayeredPCBMesh(active, (35e-6, 35e-6), 0.5e-3, 0.5e-3)
    vias = tuple(ViaConnection((0, y, x), (1, y, x), 5e-4) for y, x in ((2, 3), (10, 25), (2, 28)))
    operator = MatrixFreePCBOperator(mesh, reference_node=(1, 0, 0), vias=vias, preconditioner="jacobi")
    rhss = []
    for _ in range(5):
        a = (0, int(rng.integers(0, 3)), int(rng.integers(0, 9)))
        b = (1, int(rng.integers(9, 13)), int(rng.integers(21, 31)))
        rhss.append(operator.build_rhs((CurrentTerminal((a,), 0.3), CurrentTerminal((b,), -0.3))))
    config = _capi.config_c(1e-10, 0.0, 2e-3, 30, 20000)
    batch = _capi.Batch([operator.device_operator] * len(rhss))
    for i, b in enumerate(rhss):
        batch.set(i, b, None, config)
    batch.run()
    coef = np.stack(mesh.element_coefficients())
    shape = mesh.node_shape
    k = O.stiffness(coef, mesh.element_shape, [np.ravel_multi_index(v.lower, shape) for v in vias],
                    [np.ravel_multi_index(v.upper, shape) for v in vias], [v.conductance_s for v in vias])
    for i, b in enumerate(rhss):
        x, res = batch.result(i)
        assert res["converged"]
        exact = O.solve(k, operator.free_nodes, b)
        assert np.max(np.abs(x - exact)) <= 1e-9 * np.max(np.abs(exact))
        alone = _capi.Batch([operator.device_operator])
        alone.set(0, b, None, config)
        alone.run()
        x1, r1 = alone.result(0)
        np.testing.assert_array_equal(x1, x)
        assert r1 == res
