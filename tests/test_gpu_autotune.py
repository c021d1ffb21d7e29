"""The tile-shape autotuner on the device: the tuned batch and server give the
same bits as the default shape (a shape changes only the time), the timing
sweep covers every candidate, and a second tuning of the same board set comes
from the cache without timing."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import MPIRConfig
from pcbhip import _capi, autotune, serving, workloads
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu

CONFIG = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000).to_c()


@pytest.fixture(scope="module")
def boards():
    _capi.require_device(0)
    problems = workloads.uneven_boards(6, seed=11)
    ops = [MatrixFreeThermalOperator(p, preconditioner="jacobi") for p in problems]
    return ops, [o.build_rhs(o.default_reference_temperature()) for o in ops]


def test_tuned_batch_gives_the_default_shape_bits_and_is_cached(boards, tmp_path):
    ops, rhss = boards
    print("autotune: default-shape batch", flush=True)
    plain = _capi.Batch([o.device_operator for o in ops])
    for i, b in enumerate(rhss):
        plain.set(i, b, None, CONFIG)
    plain.run()
    expected = [plain.result(i)[0] for i in range(len(ops))]
    plain.close()

    cache = autotune.ShapeCache(tmp_path / "shapes.json")
    key = autotune.board_set_key("batch", [o.mesh.element_grid_shape for o in ops])
    for attempt in range(2):
        print(f"autotune: tuned batch, attempt {attempt}", flush=True)
        batch = _capi.Batch([o.device_operator for o in ops])
        slots = [(b, None, CONFIG) for b in rhss]
        for i, b in enumerate(rhss):
            batch.set(i, b, None, CONFIG)
        chosen = autotune.tune_batch(batch, key, slots=slots, cache=cache)
        assert chosen["cached"] == (attempt == 1)
        if attempt == 0:
            assert len(chosen["sweep"]) == len(autotune.pass_candidates())
            assert all(row["us"] > 0 for row in chosen["sweep"])
        batch.run()
        for i in range(len(ops)):
            x, result = batch.result(i)
            assert result["converged"]
            assert x.tobytes() == expected[i].tobytes(), (attempt, i)
        batch.close()


def test_tuned_server_gives_the_solves_alone(boards, tmp_path, monkeypatch):
    ops, rhss = boards
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "serve.json"))
    print("autotune: tuned server", flush=True)
    server = serving.SolveServer(max_slots=32, tune_with=[o.device_operator for o in ops])
    assert server.shape["tuned"] and (server.shape["rows"], server.shape["target"]) in autotune.pass_candidates()
    requests = [serving.ServeRequest(i, o.device_operator, [rhss[i]], CONFIG, kind="thermal", arrival_s=0.0)
                for i, o in enumerate(ops)]
    serving.run_arrivals(server, requests, time_limit_s=120.0)
    done = {r.request_id: r for r in server.completed}
    server.close()
    assert sorted(done) == list(range(len(ops)))
    for i, o in enumerate(ops):
        x_alone, _ = _capi.solve_mpir(o.device_operator, rhss[i], None, CONFIG)
        assert np.asarray(done[i].solutions[0]).tobytes() == x_alone.tobytes(), i
