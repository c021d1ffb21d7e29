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


# --- Anisotropic and graded pitch -------------------------------------------
# The checks above use equal x and y pitches, where a transposition of the
# in-plane axes is invisible.  These use graded, unequal pitches in x and y,
# distinct in-plane and through-plane conductivities, and closed forms that
# any swap of the x and y terms (coefficients, face areas, flux) breaks.

# As many cells in x as in y, so that an exchange of the two pitch arrays
# still broadcasts and must be caught by the closed forms, not by a shape.
_PITCH_X = np.array([0.10, 0.35, 0.20, 0.60]) * 1e-3
_PITCH_Y = np.array([0.30, 0.70, 0.45, 0.15]) * 1e-3
_THICK = np.array([0.20, 0.50]) * 1e-3
_K_IN, _K_Z = 2.0, 0.7


def _graded_box(fixed_axis, t_low=300.0, t_high=350.0, convection=()):
    """A box of graded cells with the two faces normal to ``fixed_axis``
    (2 = x, 1 = y, 0 = z, in node-array axes) held at ``t_low`` / ``t_high``."""

    shape = (len(_THICK), len(_PITCH_Y), len(_PITCH_X))
    nodes = tuple(n + 1 for n in shape)
    mask = np.zeros(nodes, dtype=bool)
    values = np.zeros(nodes)
    if fixed_axis is not None:
        low = [slice(None)] * 3
        high = [slice(None)] * 3
        low[fixed_axis], high[fixed_axis] = 0, -1
        mask[tuple(low)] = mask[tuple(high)] = True
        values[tuple(low)], values[tuple(high)] = t_low, t_high
    return P.OracleProblem(
        thickness=_THICK.copy(), pitch_x=_PITCH_X.copy(), pitch_y=_PITCH_Y.copy(),
        k_in=np.full(shape, _K_IN), k_z=np.full(shape, _K_Z), active=np.ones(shape, bool),
        fixed_mask=mask, fixed_values=values, convection=list(convection),
    )


def _node_coordinate(axis):
    edges = {2: _PITCH_X, 1: _PITCH_Y, 0: _THICK}[axis]
    return np.concatenate(([0.0], np.cumsum(edges)))


def _bar_errors(axis):
    """(max nodal deviation from the linear profile in kelvin, relative
    error of the face heat, relative error of the element flux) of the bar
    along ``axis`` solved by the oracle."""

    p = _graded_box(axis)
    ops = P.operator_arrays(p)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], p.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], p.element_shape)
    ref = P.default_reference(ops, p)
    t = A.solve_direct(a, P.build_rhs(ops, ref, lambda v: k @ v)) + ref
    coord = _node_coordinate(axis)
    length = coord[-1]
    shape = [1, 1, 1]
    shape[axis] = -1
    linear = 300.0 + 50.0 * (coord / length).reshape(shape)
    profile = np.max(np.abs(t.reshape(p.node_shape) - linear))
    # Heat leaving through the hot face = k dT A / L, A the cross-section.
    extent = {2: _PITCH_X.sum(), 1: _PITCH_Y.sum(), 0: _THICK.sum()}
    area = np.prod([v for ax, v in extent.items() if ax != axis])
    cond = _K_Z if axis == 0 else _K_IN
    exact = cond * 50.0 * area / length
    residual = (k @ t).reshape(p.node_shape)
    hot = [slice(None)] * 3
    hot[axis] = -1
    face_heat = float(np.sum(residual[tuple(hot)]))
    flux = P.element_heat_flux(p, t)
    want = np.zeros(3)
    want[{2: 0, 1: 1, 0: 2}[axis]] = -cond * 50.0 / length
    flux_err = float(np.max(np.abs(flux - want)) / (cond * 50.0 / length))
    return profile, abs(face_heat - exact) / exact, flux_err


@pytest.mark.parametrize("axis", [2, 1, 0], ids=["x", "y", "z"])
def test_graded_anisotropic_bar_has_the_linear_profile_and_exact_flux(axis):
    # tests/test_tensor_grid.py:181-199 checks the graded strip the same way:
    # the Q1 space holds the linear field exactly, whatever the grading.
    profile, heat, flux = _bar_errors(axis)
    assert profile <= 1e-10
    assert heat <= 1e-12
    assert flux <= 1e-12


def test_exposed_side_faces_lump_their_own_areas():
    # Sum of the lumped conductance of each side = h times that side's area:
    # y L_z for the x faces, x L_z for the y faces (boundaries.py:170-182).
    lx, ly, lz = _PITCH_X.sum(), _PITCH_Y.sum(), _THICK.sum()
    for direction, area in (("-x", ly * lz), ("+x", ly * lz), ("-y", lx * lz), ("+y", lx * lz)):
        p = _graded_box(None, convection=[P.Convection("exposed", 7.0, 290.0, (direction,))])
        w, rw = P.convection_weights(p, p.convection[0])
        assert w.sum() == pytest.approx(7.0 * area, rel=1e-13), direction
        assert rw.sum() == pytest.approx(7.0 * 290.0 * area, rel=1e-13), direction
        # Only the nodes of that face carry weight.
        axis, end = {"x": 2, "y": 1}[direction[1]], (0 if direction[0] == "-" else -1)
        grid = w.reshape(p.node_shape)
        face = [slice(None)] * 3
        face[axis] = end
        assert grid[tuple(face)].sum() == pytest.approx(w.sum(), rel=1e-14)


def test_a_transposed_pitch_fails_the_anisotropic_checks(monkeypatch):
    # Mutation check, on the CPU: the oracle with its x and y pitch arrays
    # exchanged in the coefficients, or its x and y face areas exchanged,
    # must fail the closed forms above.
    original = P.element_coefficients

    def swapped_coefficients(p):
        return original(P.OracleProblem(**{**p.__dict__, "pitch_x": p.pitch_y, "pitch_y": p.pitch_x}))

    monkeypatch.setattr(P, "element_coefficients", swapped_coefficients)
    profile, heat, _ = _bar_errors(2)
    assert profile > 1e-3 or heat > 1e-3
    monkeypatch.setattr(P, "element_coefficients", original)

    original_area = P.face_area

    def swapped_area(p, direction):
        other = {"x": "y", "y": "x", "z": "z"}[direction[1]]
        return original_area(p, direction[0] + other)

    monkeypatch.setattr(P, "face_area", swapped_area)
    p = _graded_box(None, convection=[P.Convection("exposed", 7.0, 290.0, ("-x",))])
    w, _ = P.convection_weights(p, p.convection[0])
    assert w.sum() != pytest.approx(7.0 * _PITCH_Y.sum() * _THICK.sum(), rel=1e-6)
