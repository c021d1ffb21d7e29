"""The two-level preconditioner's host side on the CPU: the block-size
choice on the reference's known answers (tests/test_thermal_conduction.py:
330-337), the coloured Galerkin assembly against the oracle's explicit
``P^T A P`` on the reference's own fixture (tests/test_thermal_conduction.py:
259-303) and on deeper stacks and via-joined sheets, and the oracle itself
against a brute-force double sum."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import AggregationCoarseCorrection, choose_block_size
from oracle import assembled as A
from oracle import problem as P
from oracle import sheet as S
from oracle import two_level as T
from tests.helpers import copper_stack, stack, to_oracle
from thermal.matrix_free_mpir_fem import ThermalConductionProblem
from thermal.matrix_free_mpir_fem import choose_block_size as thermal_choose_block_size


def _system(problem):
    p = to_oracle(problem)
    ops = P.operator_arrays(p)
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], p.element_shape)
    return a, ops


def _correction(a, ops, node_shape, block, **kw):
    return AggregationCoarseCorrection(node_shape=node_shape, free_nodes=ops["free"], diagonal_high=ops["diag"],
                                       apply_high=lambda v: a @ v, block=block, **kw)


def test_block_size_selection_respects_the_coarse_size_cap():
    for choose in (choose_block_size, thermal_choose_block_size):
        assert choose((4, 41, 41)) == 4
        assert choose((4, 201, 201), max_coarse_size=2048) == 10
        assert choose((2, 3, 3)) == 4
    # The cap can always be met by one patch per layer; widths stop there.
    assert choose_block_size((2, 5, 9), max_coarse_size=1) == 9


def test_oracle_galerkin_matches_a_brute_force_double_sum():
    problem = stack(3, 4, fixed=True)
    a, ops = _system(problem)
    shape = problem.mesh.node_shape
    for block in (1, 2, 3):
        np.testing.assert_allclose(T.galerkin(a, shape, ops["free"], block),
                                   T.galerkin_brute(a.toarray(), shape, ops["free"], block), rtol=1e-13, atol=1e-18)


def test_coarse_matrix_is_the_exact_galerkin_product():
    # The reference's fixture: a 7 x 9 copper stack, one 2 x 2 patch of
    # fixed nodes, patches of 2 nodes.
    base = copper_stack(7, 9)
    mask = np.zeros(base.mesh.node_shape, dtype=bool)
    mask[0, :2, :2] = True
    problem = ThermalConductionProblem(base.mesh, convection=base.convection, fixed_temperature_mask=mask,
                                       fixed_temperature_k=298.15, element_heat_w=base.element_heat_w)
    a, ops = _system(problem)
    correction = _correction(a, ops, problem.mesh.node_shape, 2)
    assert correction.coarse_shape == (4, 4, 5)
    expected = T.galerkin(a, problem.mesh.node_shape, ops["free"], 2)
    empty = np.asarray(T.prolongation(problem.mesh.node_shape, ops["free"], 2).sum(axis=0)).reshape(-1) == 0
    assert empty.sum() == 1
    np.testing.assert_allclose(correction.coarse_matrix, expected, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(correction.coarse_inverse @ correction.coarse_matrix, np.eye(80), atol=1e-9)
    # The preconditioner is SPD and is the oracle's action.
    n = problem.mesh.size
    dense = np.column_stack([correction.apply_high(np.eye(n)[:, i]) for i in range(n)])
    np.testing.assert_allclose(dense, dense.T, atol=1e-12 * np.abs(dense).max())
    assert np.linalg.eigvalsh(dense)[0] > 0.0
    oracle_m = T.preconditioner(a, ops["diag"], problem.mesh.node_shape, ops["free"], 2)
    r = np.random.default_rng(1).standard_normal(n)
    np.testing.assert_allclose(correction.apply_high(r), oracle_m(r), rtol=1e-10, atol=1e-12 * np.abs(oracle_m(r)).max())


@pytest.mark.parametrize("slabs,block", [((35e-6, 0.5e-3, 0.5e-3, 0.5e-3, 35e-6), 3), ((35e-6,) * 7, 4)])
def test_coloured_assembly_on_deeper_stacks(slabs, block):
    # Six and eight node planes: the layer colours repeat, so planes three
    # apart share a colour.
    problem = stack(9, 11, fixed=True, slabs=slabs)
    a, ops = _system(problem)
    correction = _correction(a, ops, problem.mesh.node_shape, block)
    np.testing.assert_allclose(correction.coarse_matrix, T.galerkin(a, problem.mesh.node_shape, ops["free"], block),
                               rtol=1e-12, atol=1e-15 * np.abs(correction.coarse_matrix).max())


def _sheet(layers, rows, cols, vias, seed=0):
    rng = np.random.default_rng(seed)
    shape = (layers, rows, cols)
    coef = S.coefficients(np.ones(shape, bool), np.full(layers, 35e-6), np.full(cols, 0.5e-3), np.full(rows, 0.5e-3),
                          np.full(shape, 5.8e7))
    node_shape = (layers, rows + 1, cols + 1)
    via_a = [np.ravel_multi_index(a_, node_shape) for a_, _ in vias]
    via_b = [np.ravel_multi_index(b_, node_shape) for _, b_ in vias]
    via_g = list(rng.uniform(1.0, 10.0, len(vias)))
    k = S.stiffness(coef, shape, via_a, via_b, via_g)
    free = np.ones(node_shape, bool)
    free[0, 0, 0] = False
    a = S.system_matrix(k, free.reshape(-1))
    diag = a.diagonal()
    return a, free.reshape(-1), diag, node_shape


def test_sheet_layers_are_colours_of_their_own():
    # Vias from layer 0 to layer 3 (same colour under a period of 3) and
    # between neighbours.
    vias = [((0, 2, 3), (3, 2, 3)), ((1, 5, 5), (2, 5, 5)), ((0, 7, 1), (1, 7, 1)), ((2, 0, 8), (3, 0, 8))]
    a, free, diag, node_shape = _sheet(4, 8, 9, vias)
    correction = AggregationCoarseCorrection(node_shape=node_shape, free_nodes=free, diagonal_high=diag,
                                             apply_high=lambda v: a @ v, block=3, layer_colours=node_shape[0])
    np.testing.assert_allclose(correction.coarse_matrix, T.galerkin(a, node_shape, free, 3), rtol=1e-12,
                               atol=1e-15 * np.abs(correction.coarse_matrix).max())
    # With a period of three the 0-3 via would land on the diagonal: the
    # assembly sees it and refuses rather than returning a wrong matrix.
    with pytest.raises(ValueError, match="does not reproduce"):
        AggregationCoarseCorrection(node_shape=node_shape, free_nodes=free, diagonal_high=diag,
                                    apply_high=lambda v: a @ v, block=3, layer_colours=3)


def test_a_via_longer_than_a_patch_is_refused():
    a, free, diag, node_shape = _sheet(2, 12, 12, [((0, 1, 1), (1, 11, 11)), ((0, 6, 6), (1, 6, 6))])
    with pytest.raises(ValueError, match="does not reproduce"):
        AggregationCoarseCorrection(node_shape=node_shape, free_nodes=free, diagonal_high=diag,
                                    apply_high=lambda v: a @ v, block=3, layer_colours=2)


def test_an_indefinite_operator_is_refused():
    problem = stack(4, 4)
    a, ops = _system(problem)
    with pytest.raises(ValueError, match="positive definite"):
        _correction(-a, ops, problem.mesh.node_shape, 2)


def test_oracle_two_level_pcg_cuts_iterations_on_a_copper_stack():
    # The reference's claim (tests/test_thermal_conduction.py:306-321) on the
    # FP64 oracle: on its copper stack the coarse space cuts the PCG
    # iterations of one inner solve (to 2e-3) at least five-fold.
    problem = copper_stack(24, 24)
    a, ops = _system(problem)
    shape = problem.mesh.node_shape
    b = ops["load"] * ops["free"]
    block = choose_block_size(shape)
    _, jacobi = T.pcg(a, b, lambda r: r / ops["diag"], 2e-3, 20000)
    x, two = T.pcg(a, b, T.preconditioner(a, ops["diag"], shape, ops["free"], block), 2e-3, 20000)
    assert two * 5 < jacobi
    assert np.linalg.norm(a @ x - b) <= 2e-3 * np.linalg.norm(b)


def _port(problem):
    from oracle.native import OracleOperator

    a, ops = _system(problem)
    return a, ops, OracleOperator(ops)


def test_oracle_block_rule_is_the_reference_rule():
    for shape, width in (((4, 41, 41), 4), ((4, 201, 201), 10), ((2, 3, 3), 4), ((5, 51, 51), 4), ((5, 101, 101), 6),
                         ((5, 201, 201), 11)):
        assert T.block_size(shape) == width == choose_block_size(shape)
    assert T.block_size((2, 5, 9), max_coarse_size=1) == 9


@pytest.mark.parametrize("block", [1, 2, 3, 5])
def test_c_port_coloured_assembly_is_the_galerkin_product(block):
    # The 27-colour probing of the native path (hex_q1.cpp:703-835) against
    # the sparse P^T A P, which the brute-force double sum pins; scattered
    # fixed nodes leave some patches empty at width 1.
    problem = stack(7, 9, fixed=True)
    a, ops, port = _port(problem)
    expected = T.galerkin(a, problem.mesh.node_shape, ops["free"], block)
    coarse = port.two_level(block, threads=2)
    np.testing.assert_allclose(coarse["matrix"], expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())
    inverse = T.spd_inverse(coarse["matrix"])
    np.testing.assert_allclose(inverse @ expected, np.eye(len(expected)), atol=1e-9)
    assert coarse["inverse32"].tobytes() == inverse.astype(np.float32).tobytes()


@pytest.mark.parametrize("block", [1, 2, 3])
def test_c_port_two_level_first_step_is_along_the_preconditioned_residual(block):
    # One FP32 PCG step from zero is x1 = alpha z with z = M^-1 r: the
    # restriction, the coarse inverse and the prolongation of the C port
    # against the FP64 action through the explicit sparse P.  The direction
    # is pinned node by node; alpha = r.z / z.Az only loosely, because the
    # FP32 apply cancels heavily on the smooth coarse part of z (about 1e-3
    # of the curvature).  The expectation uses the same FP32-rounded coarse
    # inverse, applied in FP64.
    problem = stack(6, 8, fixed=True)
    a, ops, port = _port(problem)
    shape = problem.mesh.node_shape
    rhs = (ops["load"] + np.random.default_rng(1).uniform(0.0, 1e-3, port.size)) * ops["free"]
    coarse = port.two_level(block)
    prolong = T.prolongation(shape, ops["free"], block)
    z = rhs / ops["diag"] + prolong @ (coarse["inverse32"].astype(np.float64) @ (prolong.T @ rhs))
    expected = (rhs @ z) / (z @ (a @ z)) * z
    x1, its, _, _ = port.pcg_f32(rhs, 1e-30, 1, threads=2, coarse=coarse)
    assert its == 1
    moved = expected != 0.0
    assert not x1[~moved].any()
    ratio = x1[moved] / expected[moved]
    assert np.ptp(ratio) <= 1e-5 * abs(ratio.mean())
    assert abs(ratio.mean() - 1.0) <= 1e-2
    # A plain Jacobi step points elsewhere.
    jacobi = port.pcg_f32(rhs, 1e-30, 1, threads=2)[0][moved] / expected[moved]
    assert np.ptp(jacobi) > 0.1


def test_c_port_two_level_solve_is_independent_of_the_thread_count():
    problem = copper_stack(16, 20)
    _, ops, port = _port(problem)
    coarse = port.two_level()
    rhs = ops["load"] * ops["free"]
    one = port.solve_mpir(rhs, max_outer=16, threads=1, coarse=coarse)
    three = port.solve_mpir(rhs, max_outer=16, threads=3, coarse=coarse)
    assert one[1]["converged"] and one[1]["inner_iterations"] == three[1]["inner_iterations"]
    assert one[0].tobytes() == three[0].tobytes()


def test_c_port_two_level_solve_reproduces_the_published_fixture_counts():
    # The reference's 4-slab fixture at 50 x 50 elements (13,005 nodes;
    # experiments/thermal_native_benchmark.py:43-63) with its MPIR limits
    # (16 outer steps, inner 2e-3 / 200) takes 7 outer and 584 inner
    # iterations with the native two-level solve
    # (docs/THERMAL_MPIR_FEM.md:472-474).
    from pcbhip import workloads

    problem = workloads.four_slab_fixture(50)
    p = to_oracle(problem)
    ops = P.operator_arrays(p)
    from oracle.native import OracleOperator

    port = OracleOperator(ops)
    stiffness = OracleOperator(ops, stiffness=True)
    rhs = P.build_rhs(ops, P.default_reference(ops, p), stiffness.apply_f64)
    coarse = port.two_level()
    assert coarse["block"] == 4 and coarse["matrix"].shape == (845, 845)
    _, result = port.solve_mpir(rhs, rtol=1e-10, max_outer=16, max_inner=200, coarse=coarse)
    assert result["converged"]
    assert (result["outer_iterations"], result["inner_iterations"]) == (7, 584)
