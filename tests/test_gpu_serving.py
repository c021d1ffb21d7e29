"""Continuous batching on the device: requests submitted while others
iterate, through a server with fewer slots than right-hand sides (so slots
and their memory blocks are reused) and with tile tables that grow and
shrink, must give exactly the solve each right-hand side gets alone."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem.hex_dc import MatrixFreeHexDCOperator
from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi, workloads
from pcbhip.serving import ServeRequest, SolveServer
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

from tests.helpers import to_oracle

pytestmark = pytest.mark.gpu

CONFIG = _capi.config_c(1e-10, 0.0, 2e-3, 30, 20000)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _thermal(rng, slabs, rows, cols, m):
    masks = (workloads.copper_mask(rng, rows, cols, 0.4), workloads.copper_mask(rng, rows, cols, 0.5))
    problem = workloads.layered_board(slabs, rows, cols, masks, workloads.random_sources(rng, rows, cols, 2, 0.1, 3.0))
    op = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    base = op.build_rhs(op.default_reference_temperature()) - op.nodal_load()
    free = op.free_nodes.reshape(-1)
    rhs = []
    for _ in range(m):
        heat = np.zeros(problem.mesh.element_grid_shape)
        for s in workloads.random_sources(rng, rows, cols, 3, 0.1, 3.0, side_lo=1, side_hi=4):
            heat[-1, s.row:s.row + s.side, s.col:s.col + s.side] += s.power_w / (s.side * s.side)
        rhs.append(np.where(free, base + workloads.element_heat_load(problem.mesh, heat), base))
    return problem, op, rhs


def _alone(op: _capi.Operator, rhs):
    batch = _capi.Batch([op])
    batch.set(0, rhs, None, CONFIG)
    batch.run()
    return batch.result(0)


def _requests():
    rng = np.random.default_rng(11)
    boards = [_thermal(rng, 4, 20, 30, 2), _thermal(rng, 6, 41, 17, 3), _thermal(rng, 3, 64, 70, 1)]
    dc = workloads.dc_board(seed=7, size=28, rows=20, slabs=4, pairs=3)
    dc_op = MatrixFreeHexDCOperator(dc.mesh, reference_node=dc.reference_node)
    entries = [("thermal", op.device_operator, rhs, problem) for problem, op, rhs in boards]
    entries.append(("dc", dc_op.device_operator, [dc_op.build_rhs(t) for t in dc.excitations], None))
    return entries, dc_op, [op for _, op, _ in boards]


def test_served_results_equal_the_solves_alone():
    print("serving: mixed requests vs alone", flush=True)
    entries, _keep_dc, _keep = _requests()
    alone = [[_alone(op, b) for b in rhs] for _, op, rhs, _ in entries]
    # 5 slots for 9 right-hand sides: requests queue, slots and blocks are reused.
    server = SolveServer(max_slots=5, steps_per_replay=4)
    requests = [ServeRequest(i, op, rhs, CONFIG, kind=kind) for i, (kind, op, rhs, _) in enumerate(entries)]
    server.submit(requests[0])
    server.submit(requests[3])
    for _ in range(3):
        server.pump()
    server.submit(requests[1])
    server.pump()
    server.submit(requests[2])
    results = {r.request_id: r for r in server.completed}
    results.update({r.request_id: r for r in server.drain()})
    assert sorted(results) == [0, 1, 2, 3]
    for i, (kind, op, rhs, problem) in enumerate(entries):
        served = results[i]
        assert served.converged
        for k in range(len(rhs)):
            x_alone, r_alone = alone[i][k]
            np.testing.assert_array_equal(served.solutions[k], x_alone)
            assert served.results[k] == r_alone
    info = server.backend.info()
    assert info["bytes_in_use"] == 0 and info["running"] == 0 and info["finished"] == 0
    assert info["bytes_pooled"] > 0
    server.close()


def test_served_thermal_result_matches_the_oracle():
    print("serving: oracle", flush=True)
    rng = np.random.default_rng(12)
    problem, op, rhs = _thermal(rng, 4, 24, 40, 1)
    server = SolveServer(max_slots=2, steps_per_replay=8)
    server.submit(ServeRequest(0, op.device_operator, rhs, CONFIG))
    (result,) = server.drain()
    oracle = to_oracle(problem)
    ops = P.operator_arrays(oracle)
    matrix = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], oracle.element_shape)
    exact = A.solve_direct(matrix, rhs[0])
    err = np.max(np.abs(result.solutions[0] - exact)) / np.max(np.abs(exact))
    print(f"served vs oracle: {err:.3e}", flush=True)
    assert result.results[0]["relative_residual"] <= 1e-10
    assert err < 1e-9
    server.close()


def test_tile_tables_grow_and_shrink_without_changing_results():
    print("serving: table growth", flush=True)
    rng = np.random.default_rng(13)
    problem, op, rhs = _thermal(rng, 8, 64, 64, 1)
    x_alone, r_alone = _alone(op.device_operator, rhs[0])
    raw = _capi.Server(0, max_slots=48, steps_per_replay=8)
    first = [raw.submit(op.device_operator, rhs[0], None, CONFIG, tag=k) for k in range(2)]
    raw.step()
    more = [raw.submit(op.device_operator, rhs[0], None, CONFIG, tag=10 + k) for k in range(40)]
    seen = {}
    while len(seen) < 42:
        running, finished = raw.step()
        for slot, tag in raw.finished():
            seen[tag] = raw.collect(slot)
    assert set(seen) == {0, 1} | {10 + k for k in range(40)}
    for tag, (x, res) in seen.items():
        np.testing.assert_array_equal(x, x_alone)
        assert res == r_alone
    info = raw.info()
    assert info["graph_builds"] >= 2  # grown for 42 boards (and maybe shrunk again)
    assert info["free_slots"] == 48 and info["bytes_in_use"] == 0
    raw.trim()
    assert raw.info()["bytes_pooled"] == 0
    assert len(first) == 2 and len(more) == 40
    raw.close()


def test_a_full_server_refuses_and_unfinished_slots_cannot_be_collected():
    print("serving: errors", flush=True)
    rng = np.random.default_rng(14)
    _, op, rhs = _thermal(rng, 3, 16, 16, 1)
    raw = _capi.Server(0, max_slots=1, steps_per_replay=4)
    slot = raw.submit(op.device_operator, rhs[0], None, CONFIG)
    with pytest.raises(_capi.ServerFull, match="every slot"):
        raw.submit(op.device_operator, rhs[0], None, CONFIG)
    with pytest.raises(ValueError, match="no finished result"):
        raw.collect(slot)
    with pytest.raises(ValueError, match="finite"):
        bad = rhs[0].copy()
        bad[3] = np.nan
        _capi.Server(0, max_slots=1).submit(op.device_operator, bad, None, CONFIG)
    while not raw.finished():
        raw.step()
    x, res = raw.collect(slot)
    assert res["converged"]
    raw.close()
