"""A failing slot ends alone: an operator the inner PCG cannot handle, in a
batch and in a server, reports "not SPD" for its own right-hand side only,
through the kernels' own curvature check (every kernel exits normally), and
every other slot keeps exactly the bits of its solve alone.

The reference stops its inner PCG on non-positive curvature
(src/thermal/matrix_free_mpir_fem/native/hex_q1.cpp:460-463;
cpp/cuda/inner_pcg.cu:491-494) and raises RuntimeError.
"""

from __future__ import annotations

import numpy as np
import pytest

from pcbhip import _capi, workloads
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu

CONFIG = _capi.config_c(1e-10, 0.0, 2e-3, 30, 20000)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _overflowing():
    """Coefficients of 1e300: finite and positive in FP64, so the operator is
    accepted and its FP64 outer residual is finite, but infinite in the FP32
    inner precision, so the first inner pass meets a non-finite curvature
    p.Ap, which the curvature check reports as not SPD."""

    shape = (2, 6, 8)
    s, r, c = shape
    coef = np.full((3,) + shape, 1e300)
    nodes = (s + 1, r + 1, c + 1)
    free = np.ones(nodes, dtype=bool)
    free[0] = False
    op = _capi.Operator(0, shape, coef, np.zeros(int(np.prod(nodes))), free)
    rhs = np.where(free, 1.0, 0.0).reshape(-1)
    return op, rhs


def _good():
    rng = np.random.default_rng(21)
    out = []
    for slabs, rows, cols in ((4, 20, 30), (3, 33, 17), (5, 12, 70)):
        masks = (workloads.copper_mask(rng, rows, cols, 0.4), workloads.copper_mask(rng, rows, cols, 0.5))
        problem = workloads.layered_board(slabs, rows, cols, masks,
                                          workloads.random_sources(rng, rows, cols, 2, 0.1, 3.0))
        op = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
        out.append((op.device_operator, op.build_rhs(op.default_reference_temperature())))
    return out


def test_failing_slot_fails_alone_in_a_batch():
    print("failure isolation: batch", flush=True)
    bad_op, bad_rhs = _overflowing()
    good = _good()
    alone = [_capi.solve_mpir(op, rhs, None, CONFIG)[0] for op, rhs in good]
    with pytest.raises(RuntimeError, match="positive-definite"):
        _capi.solve_mpir(bad_op, bad_rhs, None, CONFIG)
    entries = [good[0], (bad_op, bad_rhs), good[1], good[2]]
    batch = _capi.Batch([op for op, _ in entries])
    for i, (_, rhs) in enumerate(entries):
        batch.set(i, rhs, None, CONFIG)
    batch.run()
    with pytest.raises(RuntimeError, match="positive-definite"):
        batch.result(1)
    for slot, want in zip((0, 2, 3), alone):
        x, result = batch.result(slot)
        assert result["converged"]
        assert x.tobytes() == want.tobytes(), slot


def test_failing_request_fails_alone_in_a_server():
    print("failure isolation: server", flush=True)
    bad_op, bad_rhs = _overflowing()
    good = _good()
    alone = [_capi.solve_mpir(op, rhs, None, CONFIG)[0] for op, rhs in good]
    server = _capi.Server(0, max_slots=8)
    slots = {}
    for k, (op, rhs) in enumerate([good[0], (bad_op, bad_rhs), good[1], good[2]]):
        slots[server.submit(op, rhs, None, CONFIG, tag=k)] = k
    results, failures = {}, []
    for _ in range(100000):
        server.step()
        for slot, tag in server.finished():
            try:
                x, _ = server.collect(slot)
                results[tag] = x
            except RuntimeError as exc:
                assert "positive-definite" in str(exc)
                failures.append(tag)
        if len(results) + len(failures) == 4:
            break
    assert failures == [1]
    for tag, want in zip((0, 2, 3), alone):
        assert results[tag].tobytes() == want.tobytes(), tag
    server.close()
