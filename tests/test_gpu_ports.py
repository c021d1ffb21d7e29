"""The N-port basis of a layered conductor on the device, on the reference's
own checks (tests/test_dc_ports.py:72-234): the closed-form strip, the
current-driven strip, superposition against direct voltage-driven solves,
second moments against the time average of the instantaneous loss, warm
starts and rejected inputs; plus the conductance matrix against the
oracle's sparse-LU Schur complement and a closed-form three-port chain."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import (
    CurrentTerminal,
    LayeredPCBMesh,
    MPIRConfig,
    PCBConductionProblem,
    PortSet,
    ViaConnection,
    dc_port_basis,
    solve_pcb_dc,
)
from oracle import sheet as O
from pcbhip import _capi

pytestmark = pytest.mark.gpu

TIGHT = MPIRConfig(relative_tolerance=1.0e-12)
T = 35.0e-6


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _strip(columns=6, rows=2):
    mesh = LayeredPCBMesh(np.ones((1, rows, columns), dtype=bool), (T,), 1.0e-3, 1.0e-3)
    ports = PortSet(pads=(tuple((0, r, 0) for r in range(rows + 1)), tuple((0, r, columns) for r in range(rows + 1))),
                    names=("in", "out"), reference=1)
    return mesh, ports


def _board():
    """Two layers with a slot, graded pitches, two vias, three pads on both layers."""

    active = np.ones((2, 4, 7), dtype=bool)
    active[0, 1:3, 2:4] = False
    mesh = LayeredPCBMesh(active, (35.0e-6, 18.0e-6), (0.4e-3, 0.5e-3, 0.6e-3, 0.5e-3, 0.4e-3, 0.5e-3, 0.6e-3),
                          (0.5e-3, 0.4e-3, 0.5e-3, 0.6e-3), conductivity_s_per_m=(5.8e7, 5.2e7))
    ports = PortSet(pads=(((0, 0, 0), (0, 1, 0)), ((0, 4, 0), (0, 3, 0)), ((1, 2, 7), (1, 1, 7), (1, 3, 7))),
                    names=("A", "B", "load"), reference=2)
    vias = (ViaConnection((0, 2, 5), (1, 2, 5), 1.0e-3), ViaConnection((0, 1, 6), (1, 1, 6), 1.5e-3))
    return mesh, ports, vias


def test_strip_conductance_and_unit_fields_match_the_closed_form():
    print("ports strip", flush=True)
    mesh, ports = _strip()
    basis = dc_port_basis(mesh, ports, config=TIGHT)
    sigma = float(mesh.conductivity_s_per_m[0, 0, 0])
    r = 6.0e-3 / (sigma * 2.0e-3 * T)
    assert basis.converged and basis.workers == 1
    np.testing.assert_allclose(basis.conductance_s, np.array([[1.0, -1.0], [-1.0, 1.0]]) / r, rtol=1e-9)
    np.testing.assert_allclose(basis.reduced_resistance_ohm, [[r]], rtol=1e-9)
    phi = basis.potential_v((1.0, -1.0))
    for row in range(3):
        np.testing.assert_allclose(phi[0, row], r * np.linspace(1.0, 0.0, 7), rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(basis.port_voltage_v((2.0, -2.0)), (2.0 * r, 0.0), rtol=1e-9)
    correlation = np.array([[4.0, -4.0], [-4.0, 4.0]])
    element, via = basis.mean_loss_w(correlation)
    assert float(element.sum()) == pytest.approx(4.0 * r, rel=1e-9)
    assert via.shape == (0,)
    np.testing.assert_allclose(basis.rms_current_density_a_per_m2(correlation), 2.0 / (2.0e-3 * T), rtol=1e-9)


def test_pads_in_line_give_the_series_chain():
    # Full-width pads at x = 0, 4, 11: ports 0 and 2 couple only through port 1.
    print("ports in line", flush=True)
    mesh = LayeredPCBMesh(np.ones((1, 3, 11), dtype=bool), (T,), 1.0e-3, 2.0e-3)
    ports = PortSet(pads=tuple(tuple((0, y, x) for y in range(4)) for x in (0, 4, 11)), reference=0)
    basis = dc_port_basis(mesh, ports, config=TIGHT)
    sigma = float(mesh.conductivity_s_per_m[0, 0, 0])
    g1, g2 = sigma * T * 6.0e-3 / 4.0e-3, sigma * T * 6.0e-3 / 7.0e-3
    expected = np.array([[g1, -g1, 0.0], [-g1, g1 + g2, -g2], [0.0, -g2, g2]])
    np.testing.assert_allclose(basis.conductance_s, expected, rtol=1e-9, atol=1e-9 * g1)


def test_conductance_matches_the_oracle_schur_complement():
    print("ports vs oracle", flush=True)
    mesh, ports, vias = _board()
    basis = dc_port_basis(mesh, ports, vias=vias, config=TIGHT)
    shape = mesh.node_shape
    flat = lambda node: int(np.ravel_multi_index(node, shape))
    via_a = [flat(v.lower) for v in vias]
    via_b = [flat(v.upper) for v in vias]
    k = O.stiffness(np.stack(mesh.element_coefficients()), mesh.element_shape, via_a, via_b,
                    [v.conductance_s for v in vias])
    g, fields = O.port_conductance(k, O.active_nodes(mesh.element_active, via_a, via_b),
                                   [[flat(n) for n in pad] for pad in ports.pads])
    np.testing.assert_allclose(basis.conductance_s, g, rtol=1e-9, atol=1e-9 * np.abs(g).max())
    for column, port in enumerate(ports.driven):
        got = basis.unit_voltage_potential_v[column].reshape(-1)
        np.testing.assert_allclose(got, fields[port], rtol=0.0, atol=1e-10)


def test_basis_potentials_equal_direct_voltage_driven_solves():
    print("ports superposition", flush=True)
    mesh, ports, vias = _board()
    basis = dc_port_basis(mesh, ports, vias=vias, config=TIGHT)
    assert basis.converged and basis.conductance_s.shape == (3, 3)
    g = basis.conductance_s
    np.testing.assert_allclose(g, g.T, rtol=0.0, atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-9 * np.abs(g).max())
    assert np.all(np.linalg.eigvalsh(basis.reduced_conductance_s) > 0.0)

    voltages = np.array([3.0e-3, -1.0e-3, 0.0])
    direct = solve_pcb_dc(PCBConductionProblem(mesh, voltage_terminals=ports.voltage_terminals(voltages), vias=vias),
                          config=TIGHT)
    superposed = np.tensordot(voltages[list(ports.driven)], basis.unit_voltage_potential_v, axes=1)
    np.testing.assert_allclose(superposed, np.nan_to_num(direct.potential_v), rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(g @ voltages, direct.voltage_terminal_current_a, rtol=1e-8)

    currents = np.array([0.7, 0.5, -1.2])
    np.testing.assert_allclose(g @ basis.port_voltage_v(currents), currents, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(
        basis.potential_v(currents),
        np.tensordot(basis.port_voltage_v(currents)[list(ports.driven)], basis.unit_voltage_potential_v, axes=1),
        rtol=1e-9, atol=1e-15)


def test_each_unit_field_is_bitwise_its_own_solve():
    # The batch runs every unit solve as a slot with the arithmetic of a solve alone.
    print("ports batch invariance", flush=True)
    mesh, ports, vias = _board()
    basis = dc_port_basis(mesh, ports, vias=vias, config=TIGHT)
    for column, port in enumerate(ports.driven):
        volts = np.zeros(ports.count)
        volts[port] = 1.0
        alone = solve_pcb_dc(PCBConductionProblem(mesh, voltage_terminals=ports.voltage_terminals(volts), vias=vias),
                             config=TIGHT)
        assert np.array_equal(basis.solves[column].solution, alone.solve.solution)
        assert basis.solves[column].inner_iterations == alone.solve.inner_iterations


def test_constant_current_reproduces_the_current_driven_solve_on_a_strip():
    print("ports current-driven strip", flush=True)
    mesh, ports = _strip(rows=1)
    basis = dc_port_basis(mesh, ports, config=TIGHT)
    problem = PCBConductionProblem(mesh, terminals=(CurrentTerminal(ports.pads[0], 1.5, "in"),
                                                    CurrentTerminal(ports.pads[1], -1.5, "out")),
                                   reference_node=ports.pads[1][0])
    direct = solve_pcb_dc(problem, config=TIGHT)
    element, _ = basis.mean_loss_w(np.outer((1.5, -1.5), (1.5, -1.5)))
    np.testing.assert_allclose(element, direct.element_joule_loss_w, rtol=1e-9, atol=1e-16)
    np.testing.assert_allclose(basis.potential_v((1.5, -1.5)), np.nan_to_num(direct.potential_v), rtol=1e-9,
                               atol=1e-14)


def test_second_moments_reproduce_the_time_average_of_the_instantaneous_loss():
    print("ports second moments", flush=True)
    mesh, ports, vias = _board()
    basis = dc_port_basis(mesh, ports, vias=vias, config=TIGHT)
    t = np.linspace(0.0, 1.0, 240, endpoint=False)

    def triangle(phase):
        return 1.0 + 0.6 * (2.0 * np.abs(((t + phase) % 1.0) - 0.5) - 0.5)

    i_a, i_b = 2.0 * triangle(0.0), 2.0 * triangle(0.5)
    waveform = np.stack((i_a, i_b, -(i_a + i_b)))
    element_bf = np.zeros(mesh.element_active.shape)
    via_bf = np.zeros(len(vias))
    for sample in waveform.T:
        phi = basis.potential_v(sample)
        element_bf += basis.operator.element_joule_loss(phi)
        via_bf += basis.operator.via_joule_loss(phi)
    element_bf /= waveform.shape[1]
    via_bf /= waveform.shape[1]
    correlation = waveform @ waveform.T / waveform.shape[1]
    element, via = basis.mean_loss_w(correlation)
    np.testing.assert_allclose(element, element_bf, rtol=1e-10, atol=1e-18)
    np.testing.assert_allclose(via, via_bf, rtol=1e-10, atol=1e-18)

    mean = waveform.mean(axis=1)
    assert float(basis.mean_loss_w(np.outer(mean, mean))[0].sum()) < float(element.sum())
    only_a = np.zeros((3, 3))
    only_a[np.ix_((0, 2), (0, 2))] = correlation[0, 0] * np.array([[1, -1], [-1, 1]])
    only_b = np.zeros((3, 3))
    only_b[np.ix_((1, 2), (1, 2))] = correlation[1, 1] * np.array([[1, -1], [-1, 1]])
    separate = basis.mean_loss_w(only_a)[0] + basis.mean_loss_w(only_b)[0]
    assert float(separate.sum()) < float(element.sum())
    rms = basis.rms_current_density_a_per_m2(correlation)
    assert rms.shape == mesh.element_active.shape
    assert np.all(rms[~mesh.element_active] == 0.0) and np.all(rms >= 0.0)


def test_warm_start_from_a_previous_basis_saves_inner_iterations():
    print("ports warm start", flush=True)
    mesh, ports, vias = _board()
    cold = dc_port_basis(mesh, ports, vias=vias)
    heated = LayeredPCBMesh(mesh.element_active, mesh.layer_thickness_m, mesh.pitch_x_m, mesh.pitch_y_m,
                            conductivity_s_per_m=np.asarray(mesh.conductivity_s_per_m) / 1.1)
    warm = dc_port_basis(heated, ports, vias=vias, initial=cold)
    restart = dc_port_basis(heated, ports, vias=vias)
    assert warm.converged
    assert warm.inner_iterations < restart.inner_iterations
    np.testing.assert_allclose(warm.conductance_s, restart.conductance_s, rtol=1e-7)


def test_invalid_currents_and_correlations_are_rejected():
    print("ports rejected inputs", flush=True)
    mesh, ports, vias = _board()
    basis = dc_port_basis(mesh, ports, vias=vias)
    with pytest.raises(ValueError, match="sum to zero"):
        basis.potential_v((1.0, 1.0, 1.0))
    with pytest.raises(ValueError, match="one current per port"):
        basis.port_voltage_v((1.0, -1.0))
    with pytest.raises(ValueError, match="3×3"):
        basis.mean_loss_w(np.eye(2))
    with pytest.raises(ValueError, match="rows must sum to zero"):
        basis.mean_loss_w(np.eye(3))
    with pytest.raises(ValueError, match="positive semi-definite"):
        basis.mean_loss_w(-np.array([[2.0, -1.0, -1.0], [-1.0, 2.0, -1.0], [-1.0, -1.0, 2.0]]))
    with pytest.raises(ValueError, match="share the ports"):
        dc_port_basis(mesh, PortSet(pads=ports.pads, reference=0), vias=vias, initial=basis)
