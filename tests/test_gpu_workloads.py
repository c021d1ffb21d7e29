"""The benchmark configurations at their full sizes on the device.

* C2: the seeded 512 x 512 x 16 board, one MPIR solve (FP32 inner, FP64
  outer) to 1e-10, against the oracle's own FP64 solve of the same system.
* C3: 64 terminal-pair excitations of the 256 x 256 x 12 DC board in one
  batch: per-column FP64 residuals through the oracle's apply, KCL at the
  gauge node, the energy identity, and bitwise equality with solves alone.
* C4: the 200 uneven boards in one batch: every slot converges, sampled
  boards' residuals through the oracle's apply, sampled slots bitwise equal
  to their solves alone.
* C5: a few seconds of Poisson arrivals of the serving pool through the
  continuous-batching server: every served result bitwise equal to its solve
  alone.

The oracle is the C restatement of the reference's native FP64 path
(``oracle/q1hex_oracle.c``) on the host; it never runs on the device.
"""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import MPIRConfig
from electrical.matrix_free_mpir_fem.hex_dc import MatrixFreeHexDCOperator
from oracle import problem as P
from oracle.native import OracleOperator
from pcbhip import _capi, serving, workloads
from tests.helpers import to_oracle
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu

MPIR = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _oracle_ops(problem):
    ops = P.operator_arrays(to_oracle(problem))
    return ops, OracleOperator(ops), OracleOperator(ops, stiffness=True)


def _relative_residual(oracle, rhs, x):
    return float(np.linalg.norm(rhs - oracle.apply_f64(x)) / np.linalg.norm(rhs))


def test_c2_board_512x512x16_matches_the_fp64_oracle():
    print("C2 512x512x16: building", flush=True)
    problem = workloads.board_512x512x16()
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    reference = operator.default_reference_temperature()
    rhs = operator.build_rhs(reference)
    ops, oracle, stiffness = _oracle_ops(problem)
    # The right-hand side the oracle builds from the same problem.
    rhs_oracle = P.build_rhs(ops, P.default_reference(ops, to_oracle(problem)), stiffness.apply_f64)
    assert np.max(np.abs(rhs - rhs_oracle)) <= 1e-12 * np.max(np.abs(rhs_oracle))
    print("C2: device solve", flush=True)
    x, result = _capi.solve_mpir(operator.device_operator, rhs, None, MPIR.to_c())
    assert result["converged"] and result["relative_residual"] <= 1e-10
    # The FP64 residual through the oracle's own apply.
    assert _relative_residual(oracle, rhs, x) <= 1.5e-10
    # The oracle's FP64 solve (FP64 inner PCG) to 1e-12, started from the
    # device's answer (only to shorten it: the result is set by its own
    # residual); the device answer is within 1e-9 of it.
    print("C2: oracle FP64 solve", flush=True)
    x_oracle, res = oracle.solve_mpir(rhs, x, rtol=1e-12, inner_tol=1e-3, max_outer=40, max_inner=100000,
                                      inner_f64=True)
    assert res["converged"]
    error = float(np.max(np.abs(x - x_oracle)) / np.max(np.abs(x_oracle)))
    print(f"C2: fp64 oracle error {error:.2e}, {result['inner_iterations']} inner", flush=True)
    assert error <= 1e-9
    # Heat balance of the device answer through the oracle's stiffness.
    budget = P.heat_budget(ops, to_oracle(problem), x + reference, stiffness.apply_f64)
    assert abs(budget["balance"]) <= 1e-9 * budget["total"]


def test_c3_dc_board_64_excitations_in_one_batch():
    print("C3 256x256x12 x 64: building", flush=True)
    board = workloads.dc_board()
    op = MatrixFreeHexDCOperator(board.mesh, reference_node=board.reference_node)
    dev = op.device_operator
    gauge = int(np.ravel_multi_index(board.reference_node, board.mesh.node_shape))
    excitations = board.excitations
    assert len(excitations) == 64
    rhss = []
    for terminals in excitations:
        rhs = op.build_rhs(terminals)
        # The injections themselves: +-1/16 A on each terminal node, zero on
        # the gauge row.
        want = np.zeros(op.size)
        for t in terminals:
            for n in t.nodes:
                want[np.ravel_multi_index(n, board.mesh.node_shape)] += t.current_a / len(t.nodes)
        want[gauge] = 0.0
        np.testing.assert_array_equal(rhs, want)
        rhss.append(rhs)
    config = MPIR.to_c()
    batch = _capi.Batch([dev] * len(rhss))
    for i, rhs in enumerate(rhss):
        batch.set(i, rhs, None, config)
    print("C3: batch of 64", flush=True)
    batch.run()
    mesh = board.mesh
    fixed = np.zeros(mesh.node_shape, dtype=bool)
    fixed[board.reference_node] = True
    p = P.OracleProblem(
        thickness=np.asarray(mesh.slab_thickness_m), pitch_x=mesh.pitch_x_m, pitch_y=mesh.pitch_y_m,
        k_in=mesh.conductivity_w_per_m_k, k_z=mesh.through_plane_conductivity_w_per_m_k, active=mesh.active,
        fixed_mask=fixed, fixed_values=np.zeros(mesh.node_shape),
    )
    ops = P.operator_arrays(p)
    oracle, stiffness = OracleOperator(ops), OracleOperator(ops, stiffness=True)
    solutions = []
    for i, (rhs, terminals) in enumerate(zip(rhss, excitations)):
        x, result = batch.result(i)
        solutions.append(x)
        assert result["converged"], i
        residual = rhs - oracle.apply_f64(x)
        assert np.linalg.norm(residual) <= 1.5e-10 * np.linalg.norm(rhs), i
        kx = stiffness.apply_f64(x)
        # KCL: the gauge node carries no net current; what it carries is the
        # sum of the free rows' residuals, a few 1e-8 A at most here, against
        # the 1 A of the excitation.
        assert abs(kx[gauge]) <= 1e-6, (i, kx[gauge])
        # Energy: the injected power x.b equals the Joule loss x.Kx of the
        # element quadratic forms, up to x.(b - Kx) on the free rows.
        injected = float(np.dot(rhs, x))
        loss = float(np.sum(op.element_joule_loss(x)))
        bound = np.linalg.norm(x) * np.linalg.norm(residual) + 1e-12 * abs(loss)
        assert abs(injected - loss) <= bound, (i, injected, loss, bound)
        assert loss > 0.0
    # Columns solved alone equal their batched bits.
    for i in (0, 37):
        x_alone, _ = _capi.solve_mpir(dev, rhss[i], None, config)
        assert x_alone.tobytes() == solutions[i].tobytes(), i


def test_c4_two_hundred_uneven_boards_in_one_batch():
    print("C4 200 uneven boards: building", flush=True)
    boards = workloads.uneven_boards(200)
    ops = [MatrixFreeThermalOperator(b, preconditioner="jacobi") for b in boards]
    rhss = [o.build_rhs(o.default_reference_temperature()) for o in ops]
    config = MPIR.to_c()
    batch = _capi.Batch([o.device_operator for o in ops])
    for i, rhs in enumerate(rhss):
        batch.set(i, rhs, None, config)
    print("C4: batch", flush=True)
    batch.run()
    results = [batch.result(i) for i in range(len(ops))]
    assert all(r["converged"] and r["relative_residual"] <= 1e-10 for _, r in results)
    for i in range(0, 200, 8):
        _, oracle, _ = _oracle_ops(boards[i])
        assert _relative_residual(oracle, rhss[i], results[i][0]) <= 1.5e-10, i
    for i in (0, 99, 199):
        x_alone, _ = _capi.solve_mpir(ops[i].device_operator, rhss[i], None, config)
        assert x_alone.tobytes() == results[i][0].tobytes(), i


def test_c5_served_poisson_stream_equals_the_solves_alone():
    print("C5 serving pool: building", flush=True)
    pool = workloads.serving_pool(thermal_boards=8, dc_boards=4, max_rhs=16)
    rng = np.random.default_rng(5)
    config = MPIR.to_c()
    arrivals = serving.poisson_arrivals(rng, 20.0, 2.0)
    requests = []
    for k, t in enumerate(arrivals):
        # 70 % thermal, 30 % DC, 1..16 right-hand sides each.
        kind = "thermal" if rng.random() < 0.7 else "dc"
        boards = [b for b in pool if b.kind == kind]
        board = boards[int(rng.integers(len(boards)))]
        m = int(rng.integers(1, 17))
        requests.append(serving.ServeRequest(k, board.operator, board.rhs[:m], config, kind=kind,
                                             arrival_s=float(t)))
    server = serving.SolveServer(0, max_slots=512, memory_budget_bytes=200e9)
    print(f"C5: serving {len(requests)} requests", flush=True)
    serving.run_arrivals(server, requests, time_limit_s=120.0)
    assert server.idle and len(server.completed) == len(requests)
    done = {r.request_id: r for r in server.completed}
    checked = 0
    for req in requests:
        res = done[req.request_id]
        assert res.error is None and res.converged
        for x, rhs in zip(res.solutions, req.rhs):
            x_alone, _ = _capi.solve_mpir(req.operator, rhs, None, config)
            assert x_alone.tobytes() == x.tobytes(), req.request_id
            checked += 1
    print(f"C5: {checked} served right-hand sides equal their solves alone", flush=True)
    server.close()
