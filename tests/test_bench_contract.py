"""The benchmark's host-side contract on the CPU: the committed profile
summaries hold an HBM traffic figure for every tile shape the autotuner may
pick for the apply on the 512 x 512 x 16 board and for the default shape of
the fused PCG pass, and a multi-GPU request is refused with a message (one
GPU per run)."""

from __future__ import annotations

import importlib.util
import subprocess
import sys

from pcbhip import autotune
from tests.conftest import ROOT


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


# On this board the stencils skip the Robin words of the node planes that are
# all free with no Robin term and the coefficient rows of the uniform FR-4
# layers, so the measured bytes sit well below the algorithmic 108.5 MB
# (apply) and 229.3 MB (pass); the bounds catch a missing or garbled line.


def test_profile_summary_has_traffic_for_every_apply_candidate():
    bench = _bench()
    shape = (16, 512, 512)
    for rows, zb in autotune.apply_candidates(shape):
        traffic = bench.profiled_traffic(shape, rows, zb)
        assert isinstance(traffic, int) and 30e6 < traffic < 200e6, (rows, zb, traffic)
    assert bench.profiled_traffic(shape, 4, 3) is None


def test_more_than_one_gpu_is_refused():
    proc = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)
    assert proc.returncode != 0
    assert "one GPU" in (proc.stdout + proc.stderr)


def test_profile_summary_has_traffic_for_the_default_pcg_pass():
    bench = _bench()
    traffic = bench.profiled_pass_traffic((16, 512, 512), 4, 2048)
    assert isinstance(traffic, int) and 100e6 < traffic < 400e6, traffic
    assert bench.profiled_pass_traffic((16, 512, 512), 8, 1024) is None


def test_general_board_profile_has_the_pass_traffic_of_its_best_shape():
    bench = _bench()
    traffic = bench.profiled_pass_traffic((16, 1024, 1024), 4, 4096, "general")
    # Above the algorithmic 915.7 MB (halo re-reads), below twice it.
    assert isinstance(traffic, int) and 915.7e6 < traffic < 1.9e9, traffic


def test_host_figure_operators_are_the_oracle_operators_of_the_front_ends():
    # The bench's host legs build the oracle port's operator from a front-end
    # operator: a thermal one from its problem, a DC one from its mesh and
    # fixed nodes.  Both must be the assembled system of the same problem.
    import numpy as np

    from oracle import assembled as A
    from oracle import problem as P
    from pcbhip import workloads
    from thermal.matrix_free_mpir_fem.mesh import flat_index

    bench = _bench()

    class Front:
        pass

    thermal = Front()
    thermal.problem = workloads.board_64x64x8()
    op = bench._oracle_operator(thermal)
    ops = P.operator_arrays(P.from_problem(thermal.problem))
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], ops["element_shape"])
    x = np.random.default_rng(0).standard_normal(op.size)
    np.testing.assert_allclose(op.apply_f64(x), a @ x, rtol=0, atol=1e-12 * np.abs(a @ x).max())

    board = workloads.dc_board(size=24, slabs=3, pairs=2)
    dc = Front()
    dc.mesh, dc.size = board.mesh, board.mesh.size
    fixed = ~board.mesh.active_nodes
    fixed.flat[flat_index(board.reference_node, board.mesh.node_shape)] = True
    dc.free_nodes = ~fixed
    op = bench._oracle_operator(dc)
    coef = np.stack(board.mesh.element_coefficients())
    a = A.system_matrix(coef, P.unit_matrices(), np.zeros(dc.size), dc.free_nodes.reshape(-1),
                        board.mesh.element_grid_shape)
    x = np.random.default_rng(1).standard_normal(op.size)
    np.testing.assert_allclose(op.apply_f64(x), a @ x, rtol=0, atol=1e-12 * np.abs(a @ x).max())

    rhs = np.zeros(op.size)
    rhs[np.flatnonzero(dc.free_nodes)[:3]] = [1.0, -0.5, -0.5]
    cost = bench.cpu_node_iteration_cost([(dc, rhs)], 2)
    assert cost["solves"] == 1 and cost["node_iterations"] >= op.size and cost["seconds_per_node_iteration"] > 0
