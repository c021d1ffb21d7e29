"""Pins the oracle's layered-sheet operator (oracle/sheet.py) on the CPU:
closed-form strip and via-chain resistances, the reference's element-view
action against the assembled matrix, constants in the kernel and a
symmetric positive-definite free block."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import sheet as O

SIGMA, T = 5.8e7, 35e-6


def test_unit_sheet_matrices_reproduce_the_flux_of_a_linear_field():
    # A field v = g x on a unit-aspect element: U_x v carries -g/1 ... the
    # row sums vanish and the x-flux is exact.
    ux, uy = O.unit_matrices()
    assert np.allclose(ux.sum(axis=1), 0.0) and np.allclose(uy.sum(axis=1), 0.0)
    v = np.array([0.0, 1.0, 0.0, 1.0])  # corners 2 dy + dx: v = x
    np.testing.assert_allclose(ux @ v, [-0.5, 0.5, -0.5, 0.5])
    np.testing.assert_allclose(uy @ v, 0.0, atol=1e-15)


def test_strip_resistance_is_exact():
    # 1 x L strip of L cells: with the left edge at V and the right at 0 the
    # Q1 field is exactly linear, the current exactly sigma t W / L V.
    length, width_cells = 6, 3
    active = np.ones((1, width_cells, length), dtype=bool)
    coef = O.coefficients(active, [T], 1e-3, 2e-3, SIGMA)
    k = O.stiffness(coef, active.shape)
    shape = (1, width_cells + 1, length + 1)
    v = np.zeros(shape)
    v[0, :, 0] = 1.0
    free = np.ones(shape, dtype=bool)
    free[0, :, 0] = False
    free[0, :, -1] = False
    rhs = np.where(free.reshape(-1), -(k @ v.reshape(-1)), v.reshape(-1))
    sol = O.solve(k, free, rhs).reshape(shape)
    np.testing.assert_allclose(sol[0], np.tile(1.0 - np.arange(length + 1) / length, (width_cells + 1, 1)),
                               atol=1e-12)
    current = float(np.sum((k @ sol.reshape(-1)).reshape(shape)[0, :, 0]))
    assert current == pytest.approx(SIGMA * T * (width_cells * 2e-3) / (length * 1e-3), rel=1e-12)


def test_two_strips_in_series_through_two_vias():
    # Layer 0: strip of L1 cells; its right edge joined by one via per node
    # to the left edge of layer 1's strip of L2 cells.  R = R1 + R_via / 2 + R2.
    l1, l2, r_via = 4, 5, 1e-3
    active = np.zeros((2, 1, max(l1, l2)), dtype=bool)
    active[0, 0, :l1] = True
    active[1, 0, :l2] = True
    coef = O.coefficients(active, [T, T], 1e-3, 1e-3, SIGMA)
    shape = (2, 2, max(l1, l2) + 1)
    via_a = [np.ravel_multi_index((0, y, l1), shape) for y in (0, 1)]
    via_b = [np.ravel_multi_index((1, y, 0), shape) for y in (0, 1)]
    k = O.stiffness(coef, active.shape, via_a, via_b, [1.0 / r_via] * 2)
    act = O.active_nodes(active, via_a, via_b)
    free = act.copy()
    sink = [np.ravel_multi_index((1, y, l2), shape) for y in (0, 1)]
    free.flat[sink] = False
    rhs = np.zeros(k.shape[0])
    for y in (0, 1):
        rhs[np.ravel_multi_index((0, y, 0), shape)] = 0.5  # 1 A spread over the source edge
    rhs[~free.reshape(-1)] = 0.0
    sol = O.solve(k, free, rhs).reshape(shape)
    r_strip = 1.0 / (SIGMA * T)  # per square (1 x 1 mm cells)
    expected = l1 * r_strip + r_via / 2 + l2 * r_strip
    assert sol[0, 0, 0] == pytest.approx(expected, rel=1e-10)
    assert sol[0, 1, 0] == pytest.approx(expected, rel=1e-10)


@pytest.mark.parametrize("seed", [0, 1])
def test_element_view_action_matches_the_assembly(seed):
    rng = np.random.default_rng(seed)
    active = rng.random((3, 5, 7)) < 0.7
    coef = O.coefficients(active, [T, 2 * T, T], rng.uniform(0.5e-3, 1.5e-3, 7), rng.uniform(0.5e-3, 1.5e-3, 5),
                          SIGMA * rng.uniform(0.5, 1.5, active.shape))
    shape = (3, 6, 8)
    n = int(np.prod(shape))
    via_a = rng.choice(n // 3, 6, replace=False)
    via_b = via_a + n // 3
    via_g = rng.uniform(10.0, 1000.0, 6)
    k = O.stiffness(coef, active.shape, via_a, via_b, via_g)
    free = O.active_nodes(active, via_a, via_b) & (rng.random(shape) < 0.9)
    x = rng.standard_normal(n)
    want = O.system_matrix(k, free) @ x
    got = O.element_view_apply(coef, active.shape, x, free, via_a, via_b, via_g)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.max(np.abs(want)))
    # Constants in the kernel of K; the system is symmetric.
    np.testing.assert_allclose(k @ np.ones(n), 0.0, atol=1e-9 * SIGMA * T)
    a = O.system_matrix(k, free).toarray()
    np.testing.assert_allclose(a, a.T, rtol=0, atol=1e-9)


def test_connected_sheets_with_a_gauge_node_have_an_spd_system():
    active = np.ones((2, 4, 5), dtype=bool)
    coef = O.coefficients(active, [T, T], 1e-3, 1e-3, SIGMA)
    shape = (2, 5, 6)
    k = O.stiffness(coef, active.shape, [np.ravel_multi_index((0, 2, 3), shape)],
                    [np.ravel_multi_index((1, 2, 3), shape)], [50.0])
    free = np.ones(shape, dtype=bool)
    free[0, 0, 0] = False  # one gauge node; the via joins the layers
    a = O.system_matrix(k, free).toarray()
    assert np.min(np.linalg.eigvalsh(a)) > 0.0
    # Without the gauge the free block is singular (a floating conductor).
    assert abs(np.min(np.linalg.eigvalsh(k.toarray()))) < 1e-9 * SIGMA * T


def _in_line_ports(l1=4, l2=7, rows=3):
    """One strip cut by three full-width pads at x = 0, l1, l1 + l2."""

    active = np.ones((1, rows, l1 + l2), dtype=bool)
    shape = (1, rows + 1, l1 + l2 + 1)
    pads = [[np.ravel_multi_index((0, y, x), shape) for y in range(rows + 1)] for x in (0, l1, l1 + l2)]
    return active, shape, pads


def test_port_conductance_of_pads_in_line_is_the_series_chain():
    # A full-width middle pad is an equipotential cut: ports 0 and 2 couple
    # only through it, so G is the tridiagonal matrix of two strip conductances.
    l1, l2, rows = 4, 7, 3
    active, shape, pads = _in_line_ports(l1, l2, rows)
    coef = O.coefficients(active, [T], 1e-3, 2e-3, SIGMA)
    k = O.stiffness(coef, active.shape)
    g, fields = O.port_conductance(k, O.active_nodes(active), pads)
    g1 = SIGMA * T * rows * 2e-3 / (l1 * 1e-3)
    g2 = SIGMA * T * rows * 2e-3 / (l2 * 1e-3)
    expected = np.array([[g1, -g1, 0.0], [-g1, g1 + g2, -g2], [0.0, -g2, g2]])
    np.testing.assert_allclose(g, expected, rtol=1e-11, atol=1e-11 * g1)
    # Port 1 at 1 V: a tent, linear down to both outer pads.
    tent = fields[1].reshape(shape)[0, 0]
    np.testing.assert_allclose(tent[: l1 + 1], np.arange(l1 + 1) / l1, atol=1e-12)
    np.testing.assert_allclose(tent[l1:], 1.0 - np.arange(l2 + 1) / l2, atol=1e-12)


def test_port_conductance_obeys_reciprocity_kcl_and_the_energy_identity():
    rng = np.random.default_rng(3)
    active = rng.random((2, 6, 9)) < 0.9
    active[:, :, 0] = True
    active[:, :, -1] = True
    shape = (2, 7, 10)
    coef = O.coefficients(active, [35e-6, 18e-6], rng.uniform(0.4e-3, 0.6e-3, 9), rng.uniform(0.4e-3, 0.6e-3, 6),
                          5.8e7)
    via_a = [np.ravel_multi_index((0, y, 5), shape) for y in range(7)]
    via_b = [np.ravel_multi_index((1, y, 5), shape) for y in range(7)]
    via_g = rng.uniform(500.0, 1500.0, 7)
    k = O.stiffness(coef, active.shape, via_a, via_b, via_g)
    act = O.active_nodes(active, via_a, via_b)
    pads = [[np.ravel_multi_index((0, y, 0), shape) for y in (0, 1, 2)],
            [np.ravel_multi_index((0, y, 0), shape) for y in (4, 5, 6)],
            [np.ravel_multi_index((1, y, 9), shape) for y in range(7)]]
    g, fields = O.port_conductance(k, act, pads)
    scale = np.abs(g).max()
    np.testing.assert_allclose(g, g.T, atol=1e-10 * scale)
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-10 * scale)
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-10 * scale)
    assert np.all(np.linalg.eigvalsh(g[:2, :2]) > 0.0)
    # V^T G V is the power the pads drive in, which the elements and vias dissipate.
    v = np.array([2.0e-3, -1.0e-3, 0.5e-3])
    field = v @ fields
    element = O.element_loss(coef, active.shape, field).sum()
    drop = field[via_a] - field[via_b]
    assert v @ g @ v == pytest.approx(element + np.sum(via_g * drop * drop), rel=1e-10)
