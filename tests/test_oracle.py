"""Pins the oracle before it judges the device.

Each check is one the reference's own tests make for this path (cited) or a
fact the mathematics fixes: closed-form element matrices, Robin lumping
weights, symmetry and definiteness, the kernel of pure conduction, agreement
of two independent formulations (line gather vs element assembly), direct
solves and closed-form temperature fields.  No check retypes the formula it
tests.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import assembled as A
from oracle import problem as P
from oracle.native import OracleOperator
from tests.helpers import DOUBLE_EPS, SINGLE_EPS, random_body, stack, to_oracle


def _ops(problem):
    op = to_oracle(problem)
    return op, P.operator_arrays(op)


def test_unit_matrices_reproduce_the_exact_flux_of_a_linear_field():
    # For u = x on the unit cube, K_e u = integral of grad N_i . e_x = the
    # outward x-flux share of each corner: -1/4 on x = 0, +1/4 on x = 1; the
    # y and z tensors see no gradient.  Likewise for y and z.
    unit = P.unit_matrices()
    corners = np.array([[k >> 2, (k >> 1) & 1, k & 1] for k in range(8)], dtype=float)  # (z, y, x)
    for axis, tensor in zip((2, 1, 0), (0, 1, 2)):
        field = corners[:, axis]
        expected = np.where(field > 0, 0.25, -0.25)
        np.testing.assert_allclose(unit[tensor] @ field, expected, atol=1e-15)
        for other in {0, 1, 2} - {tensor}:
            np.testing.assert_allclose(unit[other] @ field, 0.0, atol=1e-15)
    for u in unit:
        np.testing.assert_array_equal(u, u.T)
        np.testing.assert_allclose(u.sum(axis=1), 0.0, atol=1e-15)
        assert np.linalg.eigvalsh(u)[0] > -1e-15


def test_robin_lumping_gives_full_half_and_quarter_face_areas():
    # tests/test_thermal_conduction.py:38-76: h A at interior top nodes, h A / 2
    # on edges, h A / 4 at corners; total h A_total.
    p = P.OracleProblem(
        thickness=np.array([35e-6, 0.4e-3]), pitch_x=np.full(2, 0.2e-3), pitch_y=np.full(2, 0.3e-3),
        k_in=np.ones((2, 2, 2)), k_z=np.ones((2, 2, 2)), active=np.ones((2, 2, 2), bool),
        fixed_mask=np.zeros((3, 3, 3), bool), fixed_values=np.zeros((3, 3, 3)),
        convection=[P.Convection("top", 12.0, 300.0)],
    )
    w, rw = P.convection_weights(p, p.convection[0])
    w = w.reshape(p.node_shape)
    area = 0.2e-3 * 0.3e-3
    assert w[-1, 1, 1] == pytest.approx(12.0 * area)
    assert w[-1, 0, 1] == pytest.approx(12.0 * area / 2)
    assert w[-1, 0, 0] == pytest.approx(12.0 * area / 4)
    assert np.all(w[:-1] == 0.0)
    assert w.sum() == pytest.approx(12.0 * area * 4)
    assert rw.sum() == pytest.approx(300.0 * 12.0 * area * 4)


@pytest.mark.parametrize("shape", [(1, 1), (1, 4), (3, 2), (12, 20)])
@pytest.mark.parametrize("fixed", [False, True])
def test_line_gather_matches_the_assembled_operator(shape, fixed):
    # Two independent formulations: the reference's node-owned line gather
    # (hex_q1.cpp:90-193) and element-by-element assembly.
    problem = stack(*shape, fixed=fixed)
    op, ops = _ops(problem)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape)
    gather = OracleOperator(ops)
    x = np.random.default_rng(3).standard_normal(op.size)
    y = a @ x
    scale = abs(a) @ np.abs(x)
    np.testing.assert_allclose(gather.apply_f64(x, threads=3), y, rtol=0, atol=64 * DOUBLE_EPS * scale.max())
    y32 = gather.apply_f32(x.astype(np.float32), threads=1)
    assert np.max(np.abs(y32 - a @ x.astype(np.float32).astype(float))) <= 16 * SINGLE_EPS * scale.max()
    fixed_nodes = ops["fixed"]
    np.testing.assert_array_equal(y32[fixed_nodes], x.astype(np.float32)[fixed_nodes])
    np.testing.assert_allclose(ops["diag"], a.diagonal(), rtol=1e-15)


def test_assembled_operator_is_spd_and_annihilates_constants():
    problem = random_body((2, 3, 4), seed=5, void=True)
    op, ops = _ops(problem)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape).toarray()
    np.testing.assert_allclose(a, a.T, atol=1e-18)
    assert np.linalg.eigvalsh(a)[0] > 0.0
    k = A.stiffness_matrix(ops["coef"], ops["unit"], op.element_shape)
    np.testing.assert_allclose(k @ np.ones(op.size), 0.0, atol=1e-12 * abs(k).max())


def test_oracle_mpir_matches_a_direct_solve_and_is_thread_independent():
    problem = random_body((3, 5, 7), seed=2)
    op, ops = _ops(problem)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], op.element_shape)
    ref = P.default_reference(ops, op)
    rhs = P.build_rhs(ops, ref, lambda v: k @ v)
    direct = A.solve_direct(a, rhs)
    gather = OracleOperator(ops)
    x1, r1 = gather.solve_mpir(rhs, rtol=1e-11, max_outer=16, max_inner=1000, threads=1)
    x8, _ = gather.solve_mpir(rhs, rtol=1e-11, max_outer=16, max_inner=1000, threads=8)
    assert r1["converged"]
    assert r1["low_operator_applications"] > r1["high_operator_applications"]
    np.testing.assert_allclose(x1, direct, rtol=0, atol=1e-9 * np.abs(direct).max())
    assert x1.tobytes() == x8.tobytes()
    assert [h[0] for h in r1["history"]] == list(range(1, r1["outer_iterations"] + 1))
    assert r1["history"][0][1] == pytest.approx(1.0)
    # The plain FP64 Jacobi PCG reaches the same field.
    x64, r64 = gather.solve_mpir(rhs, rtol=1e-12, inner_tol=1e-12, max_outer=4, max_inner=5000, inner_f64=True)
    assert r64["converged"]
    np.testing.assert_allclose(x64, direct, rtol=0, atol=1e-9 * np.abs(direct).max())


def _solve_oracle(problem, rtol=1e-12):
    op, ops = _ops(problem)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], op.element_shape)
    ref = P.default_reference(ops, op)
    rhs = P.build_rhs(ops, ref, lambda v: k @ v)
    t = A.solve_direct(a, rhs) + ref
    return op, ops, t, (lambda v: k @ v)


def test_fixed_faces_give_a_linear_profile():
    # tests/test_thermal_conduction.py:110-138.
    from thermal.matrix_free_mpir_fem import LayeredThermalMesh, ThermalConductionProblem

    mesh = LayeredThermalMesh((0.5e-3,) * 4, 1.0e-3, 1.0e-3, 1.0, element_shape=(2, 3))
    mask = np.zeros(mesh.node_shape, dtype=bool)
    mask[0] = mask[-1] = True
    values = np.zeros(mesh.node_shape)
    values[0], values[-1] = 300.0, 350.0
    op, ops, t, stiff = _solve_oracle(ThermalConductionProblem(mesh, fixed_temperature_mask=mask,
                                                               fixed_temperature_k=values))
    for plane, expected in zip(t.reshape(mesh.node_shape), 300.0 + 12.5 * np.arange(5)):
        np.testing.assert_allclose(plane, expected, rtol=1e-12)
    flux = P.element_heat_flux(op, t)
    np.testing.assert_allclose(flux[..., 2], -50.0 / 2.0e-3, rtol=1e-10)
    budget = P.heat_budget(ops, op, t, stiff)
    assert budget["fixed"] == pytest.approx(0.0, abs=1e-9)


def test_uniform_heating_with_one_cooled_face_matches_the_parabola():
    # tests/test_thermal_conduction.py:141-166: T_top = T_inf + q / h,
    # T_bottom = T_top + q t / (2 k).
    from thermal.matrix_free_mpir_fem import ConvectionBoundary, LayeredThermalMesh, ThermalConductionProblem

    mesh = LayeredThermalMesh((0.25e-3,) * 8, 1.0e-3, 1.0e-3, 0.3, element_shape=(4, 5))
    problem = ThermalConductionProblem(
        mesh, convection=(ConvectionBoundary("top", 100.0, 300.0),),
        element_heat_w=np.full(mesh.element_grid_shape, 1.0 / 160.0),
    )
    op, ops, t, stiff = _solve_oracle(problem)
    t = t.reshape(mesh.node_shape)
    flux = 1.0 / (4.0e-3 * 5.0e-3)
    top = 300.0 + flux / 100.0
    assert t[-1, 2, 2] == pytest.approx(top, rel=1e-12)
    assert t[0, 2, 2] == pytest.approx(top + flux * 2.0e-3 / (2 * 0.3), rel=1e-12)
    budget = P.heat_budget(ops, op, t, stiff)
    assert budget["convective"][0] == pytest.approx(1.0, rel=1e-12)
    assert budget["balance"] == pytest.approx(0.0, abs=1e-12)


def test_nodal_load_conserves_power():
    problem = stack(6, 7, fixed=True)
    op, ops = _ops(problem)
    total = float(problem.element_heat_w.sum()) + 0.05
    assert ops["load"].sum() == pytest.approx(total, rel=1e-14)


def test_oracle_budget_balances_on_a_random_body():
    op, ops, t, stiff = _solve_oracle(random_body((3, 6, 5), seed=9, void=True))
    budget = P.heat_budget(ops, op, t, stiff)
    assert abs(budget["balance"]) <= 1e-10 * budget["total"]
