"""DC conduction on the hexahedral operator: closed-form bar resistance
(the reference's tests/test_matrix_free_mpir_fem.py:84-101 on a 3-D bar),
voltage-terminal currents, and many excitations of a copper/dielectric
board against the oracle's FP64 direct solve on the copper cluster."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import MPIRConfig
from electrical.matrix_free_mpir_fem.hex_dc import (
    CurrentTerminal,
    HexConductionProblem,
    VoltageTerminal,
    solve_hex_dc,
    solve_hex_dc_many,
)
from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi, workloads
from thermal.matrix_free_mpir_fem import LayeredThermalMesh

pytestmark = pytest.mark.gpu

SIGMA = 5.8e7


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _bar(length_cells=4):
    return LayeredThermalMesh((35e-6,), 1.0e-3, 1.0e-3, SIGMA, element_shape=(1, length_cells))


def _face(mesh, x):
    return tuple((z, y, x) for z in range(mesh.node_shape[0]) for y in range(mesh.node_shape[1]))


def test_bar_matches_the_closed_form_resistance():
    print("dc bar", flush=True)
    mesh = _bar()
    source, sink = _face(mesh, 0), _face(mesh, 4)
    # One element across: equal shares at the four end nodes are the exact
    # consistent load of a uniform current, so the field is exactly linear.
    problem = HexConductionProblem(mesh, terminals=(CurrentTerminal(source, 1.0), CurrentTerminal(sink, -1.0)),
                                   reference_node=(0, 0, 2))
    solution = solve_hex_dc(problem)
    resistance = 4.0e-3 / (SIGMA * 1.0e-3 * 35e-6)
    assert solution.solve.converged
    assert solution.joule_loss_w == pytest.approx(resistance, rel=1e-9)
    v = solution.potential_v
    np.testing.assert_allclose(v[..., 0] - v[..., 4], resistance, rtol=1e-9)
    np.testing.assert_allclose(v[..., 2], 0.0, atol=1e-9 * resistance)


def test_voltage_terminals_drive_the_closed_form_current():
    print("dc voltage terminals", flush=True)
    mesh = _bar()
    problem = HexConductionProblem(mesh, voltage_terminals=(VoltageTerminal(_face(mesh, 0), 1.0e-3),
                                                            VoltageTerminal(_face(mesh, 4), 0.0)))
    solution = solve_hex_dc(problem)
    resistance = 4.0e-3 / (SIGMA * 1.0e-3 * 35e-6)
    assert solution.solve.converged
    assert solution.voltage_terminal_current_a[0] == pytest.approx(1.0e-3 / resistance, rel=1e-9)
    assert solution.voltage_terminal_current_a[1] == pytest.approx(-1.0e-3 / resistance, rel=1e-9)


def test_unbalanced_terminals_are_rejected():
    mesh = _bar()
    with pytest.raises(ValueError, match="sum to zero"):
        HexConductionProblem(mesh, terminals=(CurrentTerminal(((0, 0, 0),), 1.0),), reference_node=(0, 0, 4))


def _oracle_potential(board, terminals):
    mesh = board.mesh
    fixed = np.zeros(mesh.node_shape, dtype=bool)
    fixed[board.reference_node] = True
    p = P.OracleProblem(
        thickness=np.asarray(mesh.slab_thickness_m), pitch_x=mesh.pitch_x_m, pitch_y=mesh.pitch_y_m,
        k_in=mesh.conductivity_w_per_m_k, k_z=mesh.through_plane_conductivity_w_per_m_k, active=mesh.active,
        fixed_mask=fixed, fixed_values=np.zeros(mesh.node_shape),
    )
    ops = P.operator_arrays(p)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], p.element_shape)
    rhs = np.zeros(p.size)
    for t in terminals:
        for n in t.nodes:
            rhs[np.ravel_multi_index(n, mesh.node_shape)] += t.current_a / len(t.nodes)
    rhs[~ops["free"]] = 0.0
    return A.solve_direct(a, rhs).reshape(mesh.node_shape)


def test_many_excitations_match_the_oracle_on_the_copper_cluster():
    print("dc many excitations", flush=True)
    board = workloads.dc_board(size=24, slabs=4, pairs=5)
    results = solve_hex_dc_many(board.mesh, board.excitations, board.reference_node,
                                config=MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30,
                                                  max_inner_iterations=20000))
    top = board.mesh.node_shape[0] - 1
    on_cluster = np.zeros(board.mesh.node_shape[1:], dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            on_cluster[dy:dy + board.cluster.shape[0], dx:dx + board.cluster.shape[1]] |= board.cluster
    for (potential, result), terminals in zip(results, board.excitations):
        assert result.converged
        exact = _oracle_potential(board, terminals)
        got = potential[top][on_cluster]
        want = exact[top][on_cluster]
        assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))


# Graded, unequal x and y pitches: a bar along x and one along y, each driven
# by voltage terminals on its end faces, carry sigma dV A / L.  An exchange of
# the in-plane axes anywhere on the device path changes the current.
_GX = np.array([0.10, 0.35, 0.20, 0.60]) * 1e-3
_GY = np.array([0.30, 0.70, 0.45, 0.15]) * 1e-3
_GT = (20e-6, 35e-6)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_graded_anisotropic_bar_carries_the_closed_form_current(axis):
    print(f"dc graded bar {axis}", flush=True)
    mesh = LayeredThermalMesh(_GT, _GX, _GY, SIGMA)
    nz, ny, nx = mesh.node_shape
    if axis == "x":
        low = tuple((z, y, 0) for z in range(nz) for y in range(ny))
        high = tuple((z, y, nx - 1) for z in range(nz) for y in range(ny))
        length, area = _GX.sum(), _GY.sum() * sum(_GT)
    else:
        low = tuple((z, 0, x) for z in range(nz) for x in range(nx))
        high = tuple((z, ny - 1, x) for z in range(nz) for x in range(nx))
        length, area = _GY.sum(), _GX.sum() * sum(_GT)
    problem = HexConductionProblem(mesh, voltage_terminals=(VoltageTerminal(low, 2.0e-3), VoltageTerminal(high, 0.0)))
    solution = solve_hex_dc(problem, config=MPIRConfig(relative_tolerance=1e-12, max_outer_iterations=40,
                                                       max_inner_iterations=20000))
    current = SIGMA * 2.0e-3 * area / length
    assert solution.solve.converged
    assert solution.voltage_terminal_current_a[0] == pytest.approx(current, rel=1e-9)
    assert solution.voltage_terminal_current_a[1] == pytest.approx(-current, rel=1e-9)
    # The potential is linear in the graded coordinate.
    coord = np.concatenate(([0.0], np.cumsum(_GX if axis == "x" else _GY)))
    shape = (1, 1, -1) if axis == "x" else (1, -1, 1)
    expected = 2.0e-3 * (1.0 - coord / coord[-1]).reshape(shape)
    np.testing.assert_allclose(solution.potential_v, np.broadcast_to(expected, mesh.node_shape), atol=1e-9 * 2.0e-3)
