"""Device MPIR parity: the whole FP64-refined Jacobi-PCG solve on the GPU
against direct FP64 solves of the oracle's assembled matrix, closed-form
temperature fields (tests/test_thermal_conduction.py:79-194), the 64 x 64 x 8
board against the oracle's sparse direct solve, and batch independence."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import MPIRConfig, solve_mpir
from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi, workloads
from tests.helpers import random_body, stack, to_oracle
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    HeatSource,
    LayeredThermalMesh,
    MatrixFreeThermalOperator,
    ThermalConductionProblem,
    solve_thermal_conduction,
)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _oracle_solution(problem):
    """FP64 direct solve of the rise and its reference temperature."""

    op = to_oracle(problem)
    ops = P.operator_arrays(op)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], op.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], op.element_shape)
    ref = P.default_reference(ops, op)
    rhs = P.build_rhs(ops, ref, lambda v: k @ v)
    return A.solve_direct(a, rhs), rhs, ref


def _rise_error(rise, exact):
    return float(np.max(np.abs(rise - exact)) / np.max(np.abs(exact)))


@pytest.mark.parametrize("seed,void", [(2, False), (3, True)])
def test_mpir_matches_the_direct_solve(seed, void):
    print(f"mpir vs direct seed={seed} void={void}", flush=True)
    problem = random_body((3, 5, 70), seed=seed, void=void)
    exact, rhs, _ = _oracle_solution(problem)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    result = solve_mpir(operator, rhs, config=MPIRConfig(relative_tolerance=1e-11, max_outer_iterations=16,
                                                         max_inner_iterations=2000))
    assert result.converged
    assert result.low_runtime == "hip-fp32"
    assert result.low_operator_applications > result.high_operator_applications
    assert _rise_error(result.solution, exact) < 1e-9
    # One record per correction, numbered from 1, starting from the residual
    # of the zero guess and falling (tests/test_native_mpir.py:84-92).
    assert [s.outer_iteration for s in result.history] == list(range(1, result.outer_iterations + 1))
    assert result.history[0].high_relative_residual == pytest.approx(1.0)
    residuals = [s.high_relative_residual for s in result.history]
    assert all(b < a for a, b in zip(residuals, residuals[1:]))
    assert sum(s.inner_iterations for s in result.history) == result.inner_iterations
    assert result.relative_residual <= 1e-11
    # The device's final residual is the oracle's residual of the device answer.
    true_rel = np.linalg.norm(rhs - (A.system_matrix(*_arrays(problem)) @ result.solution)) / np.linalg.norm(rhs)
    assert true_rel == pytest.approx(result.relative_residual, rel=1e-3, abs=1e-14)
    # Bitwise repeatable.
    again = solve_mpir(operator, rhs, config=MPIRConfig(relative_tolerance=1e-11, max_outer_iterations=16,
                                                        max_inner_iterations=2000))
    assert again.solution.tobytes() == result.solution.tobytes()


def _arrays(problem):
    ops = P.operator_arrays(to_oracle(problem))
    return ops["coef"], ops["unit"], ops["robin_total"], ops["free"], problem.mesh.element_grid_shape


def test_fp64_inner_pcg_is_the_plain_fp64_solve():
    print("plain fp64 pcg", flush=True)
    problem = random_body((2, 7, 9), seed=11)
    exact, rhs, _ = _oracle_solution(problem)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    cfg = MPIRConfig(relative_tolerance=1e-12, inner_relative_tolerance=1e-12, max_outer_iterations=4,
                     max_inner_iterations=5000)
    x, result = _capi.solve_mpir(operator.device_operator, rhs, None, cfg.to_c(), inner_fp64=True)
    assert result["converged"]
    assert result["outer_iterations"] <= 2
    assert _rise_error(x, exact) < 1e-9


def test_solve_reaches_the_closed_form_parabola():
    print("parabola", flush=True)
    mesh = LayeredThermalMesh((0.25e-3,) * 8, 1.0e-3, 1.0e-3, 0.3, element_shape=(4, 5))
    problem = ThermalConductionProblem(
        mesh, convection=(ConvectionBoundary("top", 100.0, 300.0),),
        element_heat_w=np.full(mesh.element_grid_shape, 1.0 / 160.0),
    )
    solution = solve_thermal_conduction(problem, initial_temperature_k=300.0, preconditioner="jacobi",
                                        config=MPIRConfig(max_outer_iterations=16, max_inner_iterations=1000))
    flux = 1.0 / (4.0e-3 * 5.0e-3)
    top = 300.0 + flux / 100.0
    bottom = top + flux * 2.0e-3 / (2.0 * 0.3)
    assert solution.solve.converged
    assert solution.temperature_k[-1, 2, 2] == pytest.approx(top, rel=1e-9)
    assert solution.temperature_k[0, 2, 2] == pytest.approx(bottom, rel=1e-9)
    assert solution.total_heat_input_w == pytest.approx(1.0)
    assert solution.convective_heat_w[0] == pytest.approx(1.0, rel=1e-9)
    assert solution.heat_balance_error_w == pytest.approx(0.0, abs=1e-9)


def test_fixed_faces_give_a_linear_profile_and_exact_flux():
    print("linear profile", flush=True)
    mesh = LayeredThermalMesh((0.5e-3,) * 4, 1.0e-3, 1.0e-3, 1.0, element_shape=(2, 3))
    mask = np.zeros(mesh.node_shape, dtype=bool)
    mask[0] = mask[-1] = True
    values = np.zeros(mesh.node_shape)
    values[0], values[-1] = 300.0, 350.0
    solution = solve_thermal_conduction(ThermalConductionProblem(mesh, fixed_temperature_mask=mask,
                                                                 fixed_temperature_k=values),
                                        preconditioner="jacobi")
    assert solution.solve.converged
    for plane, expected in zip(solution.temperature_k, 300.0 + 12.5 * np.arange(5)):
        np.testing.assert_allclose(plane, expected, rtol=1e-8)
    np.testing.assert_allclose(solution.heat_flux_w_per_m2[..., 2], -50.0 / 2.0e-3, rtol=0,
                               atol=2 * 1e-8 * 350 / 0.5e-3)
    assert solution.fixed_temperature_heat_w == pytest.approx(0.0, abs=1e-9)
    assert solution.heat_balance_error_w == pytest.approx(0.0, abs=1e-9)


def test_nodal_source_heat_is_conserved():
    print("conservation", flush=True)
    mesh = LayeredThermalMesh((0.4e-3,) * 2, 1.0e-3, 1.0e-3, 5.0, element_shape=(3, 3))
    mask = np.zeros(mesh.node_shape, dtype=bool)
    mask[0, 0, :] = True
    problem = ThermalConductionProblem(
        mesh, convection=(ConvectionBoundary("bottom", 20.0, 290.0),), fixed_temperature_mask=mask,
        fixed_temperature_k=290.0, heat_sources=(HeatSource(((2, 1, 1), (2, 1, 2)), 0.3, "chip"),),
    )
    solution = solve_thermal_conduction(problem, preconditioner="jacobi")
    assert solution.solve.converged
    assert solution.total_heat_input_w == pytest.approx(0.3)
    assert solution.convective_heat_w[0] + solution.fixed_temperature_heat_w == pytest.approx(0.3, rel=1e-8)
    assert solution.min_temperature_k == pytest.approx(290.0)


def test_board_64x64x8_plain_fp64_and_mpir_match_the_oracle():
    print("board 64x64x8", flush=True)
    problem = workloads.board_64x64x8()
    exact, rhs, ref = _oracle_solution(problem)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    plain = MPIRConfig(relative_tolerance=1e-10, inner_relative_tolerance=1e-10, max_outer_iterations=4,
                       max_inner_iterations=20000)
    x64, r64 = _capi.solve_mpir(operator.device_operator, rhs, None, plain.to_c(), inner_fp64=True)
    assert r64["converged"]
    assert _rise_error(x64, exact) < 1e-9
    mixed = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=20000)
    result = solve_mpir(operator, rhs, config=mixed)
    assert result.converged
    assert _rise_error(result.solution, exact) < 1e-9


def test_batched_slots_equal_their_single_solves_bit_for_bit():
    print("batch independence", flush=True)
    problems = [random_body((2, 6, 9), seed=1), stack(12, 20, fixed=True), random_body((3, 40, 70), seed=4),
                workloads.board_64x64x8()]
    cfg = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=4000)
    ops = [MatrixFreeThermalOperator(p, preconditioner="jacobi") for p in problems]
    rhss = [o.build_rhs(o.default_reference_temperature()) for o in ops]
    singles = [solve_mpir(o, b, config=cfg).solution for o, b in zip(ops, rhss)]
    batch = _capi.Batch([o.device_operator for o in ops] + [ops[0].device_operator])
    for i, b in enumerate(rhss + [rhss[0]]):
        batch.set(i, b, None, cfg.to_c())
    batch.run()
    for i, single in enumerate(singles + [singles[0]]):
        x, res = batch.result(i)
        assert res["converged"]
        assert x.tobytes() == single.tobytes(), i


@pytest.mark.parametrize("shape", [(7, 45, 70), (4, 30, 66)], ids=lambda s: "x".join(map(str, s)))
def test_tile_shape_does_not_change_a_single_bit(shape):
    print(f"tile-shape independence shape={shape}", flush=True)
    # Inner products are summed per fixed reduction cell, so rows per wave,
    # z blocks (swept up or down; with 4 slabs the top plane is a block of its
    # own) and batch composition leave the bits alone.
    problem = random_body(shape, seed=12)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    rhs = operator.build_rhs(operator.default_reference_temperature())
    cfg = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=4000)
    answers = set()
    for rows, target in ((4, 1), (4, 4096), (8, 64), (8, 100000)):
        batch = _capi.Batch([operator.device_operator])
        batch.set_shape(rows, target)
        batch.set(0, rhs, None, cfg.to_c())
        batch.run()
        x, res = batch.result(0)
        assert res["converged"]
        answers.add(x.tobytes())
    assert len(answers) == 1


def test_finished_slots_leave_the_batch_without_changing_the_rest():
    print("compaction", flush=True)
    # Eight boards of very different difficulty: the easy ones finish early,
    # the batch drops their tiles, the hard ones continue unchanged.
    problems = [random_body((2, 4 + 3 * i, 5 + 7 * i), seed=20 + i) for i in range(8)]
    cfg = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=4000)
    ops = [MatrixFreeThermalOperator(p, preconditioner="jacobi") for p in problems]
    rhss = [o.build_rhs(o.default_reference_temperature()) for o in ops]
    batch = _capi.Batch([o.device_operator for o in ops])
    for i, b in enumerate(rhss):
        batch.set(i, b, None, cfg.to_c())
    batch.run()
    for i, (o, b) in enumerate(zip(ops, rhss)):
        single = solve_mpir(o, b, config=cfg).solution
        assert batch.result(i)[0].tobytes() == single.tobytes(), i
    batch.run()  # a second run after compaction starts from the full tables
    assert all(batch.result(i)[1]["converged"] for i in range(len(ops)))


def test_batch_schedule_does_not_change_a_single_bit():
    print("batch schedule independence", flush=True)
    problems = [random_body((2, 4 + 3 * i, 5 + 7 * i), seed=40 + i) for i in range(5)]
    cfg = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=4000)
    ops = [MatrixFreeThermalOperator(p, preconditioner="jacobi") for p in problems]
    rhss = [o.build_rhs(o.default_reference_temperature()) for o in ops]
    batch = _capi.Batch([o.device_operator for o in ops])
    for i, b in enumerate(rhss):
        batch.set(i, b, None, cfg.to_c())
    batch.run()
    first = [batch.result(i)[0].tobytes() for i in range(len(ops))]
    for steps, outer, compact in ((7, 3, 10), (1, 1, 100), (96, 32, 1)):
        batch.set_schedule(steps, outer, compact)
        batch.run()
        assert [batch.result(i)[0].tobytes() for i in range(len(ops))] == first
    for bad in ((0, 1, 50), (8, 9, 50), (8, 4, 0), (8, 4, 101)):
        with pytest.raises(ValueError):
            batch.set_schedule(*bad)


def test_a_capped_first_pass_re_references_and_converges():
    print("re-reference restart", flush=True)
    # tests/test_native_thermal_system.py:188-201.
    problem = stack(9, 11, fixed=True)
    solution = solve_thermal_conduction(problem, preconditioner="jacobi",
                                        config=MPIRConfig(relative_tolerance=1e-12, max_outer_iterations=2,
                                                          max_inner_iterations=3000))
    capped = solution.solve.outer_iterations
    assert capped >= 2
    exact, _, ref = _oracle_solution(problem)
    full = solve_thermal_conduction(problem, preconditioner="jacobi",
                                    config=MPIRConfig(relative_tolerance=1e-11, max_outer_iterations=30,
                                                      max_inner_iterations=3000))
    assert full.solve.converged
    np.testing.assert_allclose(full.temperature_k.reshape(-1) - ref, exact, rtol=0, atol=1e-9 * np.abs(exact).max())


def test_exact_initial_guess_converges_without_inner_work():
    print("exact warm start", flush=True)
    problem = random_body((2, 5, 6), seed=8)
    exact, rhs, _ = _oracle_solution(problem)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    result = solve_mpir(operator, rhs, config=MPIRConfig(relative_tolerance=1e-8), initial_guess=exact)
    assert result.converged and result.outer_iterations == 0 and result.inner_iterations == 0
    assert result.high_operator_applications == 1


def test_pass_timer_keeps_results_and_later_runs():
    print("pass timer", flush=True)
    problem = random_body((3, 30, 50), seed=21)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    rhs = operator.build_rhs(operator.default_reference_temperature())
    cfg = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=30, max_inner_iterations=4000)
    batch = _capi.Batch([operator.device_operator])
    batch.set(0, rhs, None, cfg.to_c())
    with pytest.raises(ValueError, match="run the batch"):
        batch.time_pass(5)
    batch.run()
    before, res_before = batch.result(0)
    us = batch.time_pass(20)
    assert 0.0 < us < 1e5
    after, res_after = batch.result(0)
    assert after.tobytes() == before.tobytes() and res_after == res_before
    batch.run()
    again, res_again = batch.result(0)
    assert again.tobytes() == before.tobytes() and res_again["inner_iterations"] == res_before["inner_iterations"]


def test_a_stack_deeper_than_the_plane_masks_matches_the_direct_solve():
    print("70-slab stack vs direct", flush=True)
    # 70 slabs: the FR-4 element layers 1..68 are uniform and the node planes
    # 1..69 are free with no Robin term, but the operator flags only those
    # below 64.  The others load like any layer, and the column's z blocks
    # sweep in both directions.
    rng = np.random.default_rng(41)
    masks = (workloads.copper_mask(rng, 6, 9, 0.5), workloads.copper_mask(rng, 6, 9, 0.5))
    problem = workloads.layered_board(70, 6, 9, masks, [workloads.Source(1, 2, 3, 0.5)])
    exact, rhs, _ = _oracle_solution(problem)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    a = A.system_matrix(*_arrays(problem))
    x = np.random.default_rng(42).standard_normal(problem.mesh.size)
    scale = float((abs(a) @ np.abs(x)).max())
    np.testing.assert_allclose(operator.device_operator.apply_f64(x), a @ x, rtol=0, atol=64 * 2.0**-52 * scale)
    x32 = x.astype(np.float32)
    y32 = operator.device_operator.apply_f32(x32)
    assert np.max(np.abs(y32 - a @ x32.astype(np.float64))) <= 32 * 2.0**-23 * scale
    result = solve_mpir(operator, rhs, config=MPIRConfig(relative_tolerance=1e-11, max_outer_iterations=30,
                                                         max_inner_iterations=20000))
    assert result.converged
    assert _rise_error(result.solution, exact) < 1e-9


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("preconditioner", ["jacobi", "two-level"])
def test_graded_anisotropic_bar_has_the_linear_profile_and_exact_heat(axis, preconditioner):
    # Graded, unequal x and y pitches with distinct in-plane and through-plane
    # conductivities; the faces normal to ``axis`` are held at 300 K and 350 K.
    # The Q1 space holds the linear field exactly and the heat through the
    # hot face is k dT A / L; an exchange of the in-plane axes breaks both.
    print(f"graded bar {axis} {preconditioner}", flush=True)
    gx = np.array([0.10, 0.35, 0.20, 0.60, 0.25, 0.40]) * 1e-3
    gy = np.array([0.30, 0.70, 0.45, 0.15, 0.50, 0.20]) * 1e-3
    thick = (0.2e-3, 0.5e-3)
    mesh = LayeredThermalMesh(thick, gx, gy, 2.0, through_plane_conductivity_w_per_m_k=0.7)
    mask = np.zeros(mesh.node_shape, dtype=bool)
    values = np.zeros(mesh.node_shape)
    if axis == "x":
        mask[..., 0] = mask[..., -1] = True
        values[..., 0], values[..., -1] = 300.0, 350.0
        coord, area = np.concatenate(([0.0], np.cumsum(gx))), gy.sum() * sum(thick)
        shape = (1, 1, -1)
    else:
        mask[:, 0, :] = mask[:, -1, :] = True
        values[:, 0, :], values[:, -1, :] = 300.0, 350.0
        coord, area = np.concatenate(([0.0], np.cumsum(gy))), gx.sum() * sum(thick)
        shape = (1, -1, 1)
    problem = ThermalConductionProblem(mesh, fixed_temperature_mask=mask, fixed_temperature_k=values)
    # Solved to 1e-12.  The residual is relative to |b|, which the fixed rows
    # (kelvin) dominate over the free rows (watts), so the free nodes keep an
    # error of about 1e-9 of the range; the bounds below are 1e-7, and an
    # exchange of the in-plane axes is an error of order 1e-1.
    config = MPIRConfig(relative_tolerance=1e-12, max_outer_iterations=40, max_inner_iterations=20000)
    solution = solve_thermal_conduction(problem, config=config, preconditioner=preconditioner, coarse_block_nodes=4)
    assert solution.solve.converged
    expected = 300.0 + 50.0 * (coord / coord[-1]).reshape(shape)
    np.testing.assert_allclose(solution.temperature_k, np.broadcast_to(expected, mesh.node_shape), rtol=0,
                               atol=1e-7 * 50.0)
    # The fixed heat sums both faces (zero); the x (or y) flux at every element
    # centre is -k dT / L.
    flux = solution.heat_flux_w_per_m2[..., 0 if axis == "x" else 1]
    np.testing.assert_allclose(flux, -2.0 * 50.0 / coord[-1], rtol=1e-6)
    assert abs(solution.fixed_temperature_heat_w) <= 1e-7 * 2.0 * 50.0 * area / coord[-1]
    other = solution.heat_flux_w_per_m2[..., 1 if axis == "x" else 0]
    assert np.max(np.abs(other)) <= 1e-6 * 2.0 * 50.0 / coord[-1]
