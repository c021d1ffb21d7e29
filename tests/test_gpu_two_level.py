"""The two-level preconditioner on the device, on the reference's own checks
(tests/test_thermal_conduction.py:259-327): the coarse matrix is the exact
Galerkin product and the device action tracks the FP64 action; two-level
PCG cuts the inner iterations at least five-fold and lands on the Jacobi
answer; the default configuration converges in under 600 inner iterations.
Beyond them: FP64-oracle parity, slot independence in mixed batches, the
FP64 inner solve, the layered-sheet operator and its N-port basis, and
served solves."""

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
    solve_mpir,
    solve_pcb_dc,
)
from oracle import assembled as A
from oracle import problem as P
from oracle import sheet as S
from oracle import two_level as T
from pcbhip import _capi, serving
from tests.helpers import copper_stack, random_body, stack, to_oracle
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator, ThermalConductionProblem, solve_thermal_conduction

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _oracle(problem):
    p = to_oracle(problem)
    ops = P.operator_arrays(p)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], p.element_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], p.element_shape)
    reference = P.default_reference(ops, p)
    rhs = P.build_rhs(ops, reference, lambda v: k @ v)
    return a, ops, rhs


def test_coarse_matrix_is_the_galerkin_product_and_the_device_action_tracks_fp64():
    print("two-level coarse matrix and device action", flush=True)
    base = copper_stack(7, 9)
    mask = np.zeros(base.mesh.node_shape, dtype=bool)
    mask[0, :2, :2] = True  # one patch made only of fixed nodes
    problem = ThermalConductionProblem(base.mesh, convection=base.convection, fixed_temperature_mask=mask,
                                       fixed_temperature_k=298.15, element_heat_w=base.element_heat_w)
    operator = MatrixFreeThermalOperator(problem, coarse_block_nodes=2)
    correction = operator.coarse_correction
    assert correction is not None and operator.preconditioner == "two-level"
    assert correction.coarse_shape == (4, 4, 5)
    a, ops, _ = _oracle(problem)
    np.testing.assert_allclose(correction.coarse_matrix, T.galerkin(a, problem.mesh.node_shape, ops["free"], 2),
                               rtol=1e-12, atol=1e-15)
    vector = np.random.default_rng(1).standard_normal(operator.size)
    high = T.preconditioner(a, ops["diag"], problem.mesh.node_shape, ops["free"], 2)(vector)
    np.testing.assert_allclose(correction.apply_high(vector), high, rtol=0.0, atol=1e-10 * np.max(np.abs(high)))
    low = operator.precondition_low(vector)
    assert low.dtype == np.float32
    np.testing.assert_allclose(low, high, rtol=0.0, atol=1e-5 * np.max(np.abs(high)))
    # Jacobi alone: r / diag.
    jacobi = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    assert jacobi.coarse_correction is None
    np.testing.assert_allclose(jacobi.precondition_low(vector), vector / ops["diag"], rtol=2e-7, atol=0.0)


@pytest.mark.parametrize("block", [1, 3, 7])
def test_device_galerkin_products_match_the_oracle(block):
    print(f"device galerkin block={block}", flush=True)
    problem = stack(13, 70, fixed=True, slabs=(35e-6, 0.4e-3, 0.4e-3, 0.4e-3, 35e-6))
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    a, ops, _ = _oracle(problem)
    shape = problem.mesh.node_shape
    p = T.prolongation(shape, ops["free"], block)
    vectors = np.random.default_rng(block).standard_normal((3, p.shape[1]))
    got = operator.device_operator.galerkin_f64(block, vectors)
    expected = np.stack([p.T @ (a @ (p @ v)) for v in vectors])
    np.testing.assert_allclose(got, expected, rtol=0.0, atol=1e-12 * np.abs(expected).max())
    with pytest.raises(ValueError, match="vectors must be"):
        operator.device_operator.galerkin_f64(block, vectors[:, :-1])


def test_device_factorisation_inverts_the_coarse_matrix_and_refuses_an_indefinite_one():
    print("device coarse factorisation", flush=True)
    problem = stack(9, 10)
    operator = MatrixFreeThermalOperator(problem, coarse_block_nodes=2)
    correction = operator.coarse_correction
    n = correction.coarse_size
    np.testing.assert_allclose(correction.coarse_inverse @ correction.coarse_matrix, np.eye(n), atol=1e-9)
    np.testing.assert_allclose(correction.coarse_inverse, np.linalg.inv(correction.coarse_matrix), rtol=1e-9,
                               atol=1e-12 * np.abs(correction.coarse_inverse).max())
    # The host-inverse entry point installs the same preconditioner.
    twin = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    twin.device_operator.set_coarse(2, correction.coarse_inverse)
    r = np.random.default_rng(4).standard_normal(operator.size)
    assert twin.precondition_low(r).tobytes() == operator.precondition_low(r).tobytes()
    twin.device_operator.set_coarse(0, None)
    np.testing.assert_array_equal(twin.precondition_low(r), np.float32(r) * (1.0 / twin._diagonal_high).astype(np.float32))
    with pytest.raises(ValueError, match="positive definite"):
        operator.device_operator.set_coarse_matrix(2, -correction.coarse_matrix)
    with pytest.raises(ValueError, match="does not match"):
        operator.device_operator.set_coarse_matrix(3, correction.coarse_matrix)


def test_two_level_preconditioner_cuts_inner_iterations_on_a_copper_stack():
    print("two-level vs jacobi copper stack 24x24", flush=True)
    problem = copper_stack(24, 24)
    config = MPIRConfig(max_inner_iterations=3000)
    two_level = solve_thermal_conduction(problem, config=config, initial_temperature_k=298.15)
    jacobi = solve_thermal_conduction(problem, config=config, initial_temperature_k=298.15, preconditioner="jacobi")
    assert two_level.solve.converged and jacobi.solve.converged
    assert two_level.solve.inner_iterations * 5 < jacobi.solve.inner_iterations
    np.testing.assert_allclose(two_level.temperature_k, jacobi.temperature_k, rtol=1e-9, atol=1e-7)
    assert two_level.heat_balance_error_w == pytest.approx(0.0, abs=1e-9)


def test_default_config_converges_on_a_copper_stack():
    print("two-level default config 40x40", flush=True)
    solution = solve_thermal_conduction(copper_stack(40, 40), initial_temperature_k=298.15)
    assert solution.solve.converged
    assert solution.solve.inner_iterations < 600


@pytest.mark.parametrize("seed,void", [(2, False), (3, True)])
def test_two_level_mpir_matches_the_direct_solve(seed, void):
    print(f"two-level mpir vs direct seed={seed} void={void}", flush=True)
    problem = random_body((3, 5, 70), seed=seed, void=void)
    a, _, rhs = _oracle(problem)
    exact = A.solve_direct(a, rhs)
    operator = MatrixFreeThermalOperator(problem, coarse_block_nodes=4)
    # Conductivities span five decades: the error is the conditioning times
    # the residual, so the residual target is 1e-12.
    result = solve_mpir(operator, rhs, config=MPIRConfig(relative_tolerance=1e-12, max_outer_iterations=16,
                                                         max_inner_iterations=2000))
    assert result.converged
    assert float(np.max(np.abs(result.solution - exact)) / np.max(np.abs(exact))) < 1e-9


def test_fp64_inner_two_level_solve():
    print("two-level fp64 inner", flush=True)
    problem = stack(11, 13, fixed=True)
    a, _, rhs = _oracle(problem)
    exact = A.solve_direct(a, rhs)
    operator = MatrixFreeThermalOperator(problem, coarse_block_nodes=3)
    config = _capi.config_c(1e-12, 0.0, 1e-10, 4, 5000)
    x, r = _capi.solve_mpir(operator.device_operator, rhs, None, config, inner_fp64=True)
    assert r["converged"]
    assert float(np.max(np.abs(x - exact)) / np.max(np.abs(exact))) < 1e-10


def test_two_level_slots_equal_their_solves_alone_in_a_mixed_batch():
    print("two-level mixed batch", flush=True)
    problems = [stack(9, 12, fixed=True), random_body((2, 7, 20), seed=5), stack(17, 6)]
    two = [MatrixFreeThermalOperator(p, coarse_block_nodes=3) for p in problems]
    plain = [MatrixFreeThermalOperator(p, preconditioner="jacobi") for p in problems]
    config = _capi.config_c(1e-10, 0.0, 2e-3, 16, 4000)
    rhs = [_oracle(p)[2] for p in problems]
    alone = [_capi.solve_mpir(op.device_operator, b, None, config) for op, b in zip(two, rhs)]
    ops = [two[0].device_operator, plain[1].device_operator, two[1].device_operator, plain[0].device_operator,
           two[2].device_operator, two[0].device_operator]
    order = [0, 1, 1, 0, 2, 0]
    batch = _capi.Batch(ops)
    for slot, k in enumerate(order):
        batch.set(slot, rhs[k], None, config)
    batch.run()
    for slot, op in enumerate(ops):
        x, r = batch.result(slot)
        assert r["converged"]
        if op.coarse_size:
            ref_x, ref_r = alone[order[slot]]
            assert x.tobytes() == ref_x.tobytes()
            assert r["inner_iterations"] == ref_r["inner_iterations"]
    batch.close()


def _sheet_board():
    active = np.ones((3, 10, 14), dtype=bool)
    active[1, 3:7, 4:9] = False
    mesh = LayeredPCBMesh(active, (35e-6, 18e-6, 35e-6), 0.5e-3, 0.5e-3)
    vias = (ViaConnection((0, 2, 2), (1, 2, 2), 2e-3), ViaConnection((1, 8, 12), (2, 8, 12), 2e-3),
            ViaConnection((0, 9, 13), (2, 9, 13), 1e-3))
    problem = PCBConductionProblem(
        mesh, terminals=(CurrentTerminal(((0, 0, 0), (0, 1, 0)), 2.0, "in"),
                         CurrentTerminal(((2, 10, 14), (2, 9, 14)), -2.0, "out")),
        reference_node=(2, 10, 14), vias=vias)
    return problem


def test_sheet_operator_two_level_solve_matches_the_oracle():
    print("two-level sheet dc", flush=True)
    problem = _sheet_board()
    tight = MPIRConfig(relative_tolerance=1e-12)
    two = solve_pcb_dc(problem, config=tight, coarse_block_nodes=3)
    jacobi = solve_pcb_dc(problem, config=MPIRConfig(relative_tolerance=1e-12, max_inner_iterations=20000),
                          preconditioner="jacobi")
    assert two.solve.converged and jacobi.solve.converged
    assert two.solve.inner_iterations < jacobi.solve.inner_iterations
    mesh = problem.mesh
    shape = mesh.node_shape
    flat = lambda node: int(np.ravel_multi_index(node, shape))
    via_a = [flat(v.lower) for v in problem.vias]
    via_b = [flat(v.upper) for v in problem.vias]
    k = S.stiffness(np.stack(mesh.element_coefficients()), mesh.element_shape, via_a, via_b,
                    [v.conductance_s for v in problem.vias])
    active = S.active_nodes(mesh.element_active, via_a, via_b).reshape(-1)
    free = active.copy()
    free[flat(problem.reference_node)] = False
    rhs = np.zeros(mesh.size)
    for t in problem.terminals:
        for node in t.nodes:
            rhs[flat(node)] += t.current_a / len(t.nodes)
    rhs[~free] = 0.0
    exact = S.solve(k, free, rhs)
    got = np.nan_to_num(two.potential_v.reshape(-1))
    np.testing.assert_allclose(got[active], exact[active], rtol=0.0, atol=1e-9 * np.max(np.abs(exact)))
    np.testing.assert_allclose(two.joule_loss_w, jacobi.joule_loss_w, rtol=1e-9)


def test_port_basis_is_the_same_with_either_preconditioner():
    print("two-level port basis", flush=True)
    problem = _sheet_board()
    ports = PortSet(pads=(((0, 0, 0), (0, 1, 0)), ((2, 10, 14), (2, 9, 14)), ((1, 0, 14), (1, 1, 14))),
                    reference=1)
    two = dc_port_basis(problem.mesh, ports, vias=problem.vias, config=MPIRConfig(relative_tolerance=1e-12),
                        coarse_block_nodes=3)
    jacobi = dc_port_basis(problem.mesh, ports, vias=problem.vias, preconditioner="jacobi",
                           config=MPIRConfig(relative_tolerance=1e-12, max_inner_iterations=20000))
    assert two.converged and jacobi.converged
    assert two.operator.coarse_correction is not None
    np.testing.assert_allclose(two.conductance_s, jacobi.conductance_s, rtol=1e-9,
                               atol=1e-9 * np.abs(jacobi.conductance_s).max())


def test_served_two_level_solves_equal_their_solves_alone():
    print("two-level served", flush=True)
    problems = [stack(6, 8), stack(13, 9, fixed=True), random_body((2, 9, 30), seed=7)]
    two = [MatrixFreeThermalOperator(p, coarse_block_nodes=3) for p in problems]
    plain = MatrixFreeThermalOperator(problems[1], preconditioner="jacobi")
    config = _capi.config_c(1e-10, 0.0, 2e-3, 16, 4000)
    rhs = [_oracle(p)[2] for p in problems]
    alone = [_capi.solve_mpir(op.device_operator, b, None, config) for op, b in zip(two, rhs)]
    plain_alone = _capi.solve_mpir(plain.device_operator, rhs[1], None, config)
    # One server, two-level and Jacobi slots in flight together.
    with serving.ServedSolves(max_slots=8, steps_per_replay=4) as executor, serving.route_solves(executor):
        import threading

        got = {}

        def run(key, op, b):
            got[key] = serving.solve_mpir_routed(op, b, None, config)

        threads = [threading.Thread(target=run, args=(k, op.device_operator, b)) for k, (op, b) in
                   enumerate(zip(two, rhs))]
        threads.append(threading.Thread(target=run, args=("plain", plain.device_operator, rhs[1])))
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    for k in range(len(problems)):
        x, r = got[k]
        assert r["converged"]
        assert x.tobytes() == alone[k][0].tobytes()
        assert r["inner_iterations"] == alone[k][1]["inner_iterations"]
    assert got["plain"][0].tobytes() == plain_alone[0].tobytes()
