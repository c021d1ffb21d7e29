"""The multi-right-hand-side block apply: slots that share an operator
applied in groups, with the matrix-core (split-bf16) element action or the
FP32 vector action.  Each right-hand side keeps its own scalars and FP64
refinement, so every one must reach the FP64 oracle's solution, and its bits
must not depend on the group it was applied in."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem.hex_dc import MatrixFreeHexDCOperator
from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi, workloads
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu

CONFIG = _capi.config_c(1e-10, 0.0, 2e-3, 30, 20000)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


@pytest.fixture(scope="module")
def dc():
    board = workloads.dc_board(seed=9, size=30, rows=22, slabs=4, pairs=6)
    op = MatrixFreeHexDCOperator(board.mesh, reference_node=board.reference_node)
    rhs = [op.build_rhs(t) for t in board.excitations]
    return board, op, rhs


def _run(op, rhss, group, engine, rows=4):
    batch = _capi.Batch([op.device_operator] * len(rhss))
    batch.set_shape(rows, 1024)
    batch.set_block(group, engine)
    for i, b in enumerate(rhss):
        batch.set(i, b, None, CONFIG)
    batch.run()
    return [batch.result(i) for i in range(len(rhss))]


def _dc_oracle(board, rhs):
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
    return A.solve_direct(a, rhs).reshape(mesh.node_shape)


@pytest.mark.parametrize("engine", [2, 3])
def test_block_mfma_solves_reach_the_oracle_on_the_copper_cluster(dc, engine):
    print(f"block mfma: DC excitations vs oracle, engine {engine}", flush=True)
    board, op, rhs = dc
    results = _run(op, rhs, 4, engine)
    top = board.mesh.node_shape[0] - 1
    on = np.zeros(board.mesh.node_shape[1:], dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            on[dy:dy + board.cluster.shape[0], dx:dx + board.cluster.shape[1]] |= board.cluster
    for (x, res), b in zip(results, rhs):
        assert res["converged"] and res["relative_residual"] <= 1e-10
        exact = _dc_oracle(board, b)
        got = x.reshape(board.mesh.node_shape)[top][on]
        want = exact[top][on]
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"  error vs oracle {err:.2e}, inner {res['inner_iterations']}", flush=True)
        assert err <= 1e-9


def test_block_bits_do_not_depend_on_the_group(dc):
    print("block mfma: group independence", flush=True)
    _, op, rhs = dc
    reference = _run(op, rhs, 6, 3)
    for group, rows in ((1, 4), (2, 8), (4, 4)):
        other = _run(op, rhs, group, 3, rows)
        for (x0, r0), (x1, r1) in zip(reference, other):
            np.testing.assert_array_equal(x1, x0)
            assert r1 == r0
    alone = _run(op, rhs[2:3], 1, 3)
    np.testing.assert_array_equal(alone[0][0], reference[2][0])
    assert alone[0][1] == reference[2][1]


def test_vector_block_apply_is_the_plain_slot_path_bit_for_bit(dc):
    print("block vector: equals plain slots", flush=True)
    _, op, rhs = dc
    plain = _run(op, rhs, 0, 0)
    block = _run(op, rhs, 3, 0)
    for (x0, r0), (x1, r1) in zip(plain, block):
        np.testing.assert_array_equal(x1, x0)
        assert r1 == r0


def test_block_mfma_thermal_right_hand_sides_reach_the_direct_solve():
    print("block mfma: thermal right-hand sides", flush=True)
    rng = np.random.default_rng(21)
    masks = (workloads.copper_mask(rng, 36, 44, 0.5), workloads.copper_mask(rng, 36, 44, 0.4))
    problem = workloads.layered_board(5, 36, 44, masks, workloads.random_sources(rng, 36, 44, 2, 0.1, 2.0))
    op = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    base = op.build_rhs(op.default_reference_temperature()) - op.nodal_load()
    free = op.free_nodes.reshape(-1)
    rhss = []
    for _ in range(5):
        heat = np.zeros(problem.mesh.element_grid_shape)
        for s in workloads.random_sources(rng, 36, 44, 3, 0.1, 3.0, side_lo=1, side_hi=5):
            heat[-1, s.row:s.row + s.side, s.col:s.col + s.side] += s.power_w / (s.side * s.side)
        rhss.append(np.where(free, base + workloads.element_heat_load(problem.mesh, heat), base))
    ops = P.operator_arrays(P.from_problem(problem))
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], problem.mesh.element_grid_shape)
    for (x, res), b in zip(_run(op, rhss, 5, 2), rhss):
        exact = A.solve_direct(a, b)
        err = np.max(np.abs(x - exact)) / np.max(np.abs(exact))
        assert res["converged"] and err < 1e-9, err
