"""Host logic of the serving loop on the CPU: FIFO admission by slots,
device-memory budget and nodes in flight, request completion and latency
bookkeeping, driven through a stand-in for the native server that finishes
each slot a fixed number of steps after it was submitted."""

from __future__ import annotations

import numpy as np
import pytest

from pcbhip import _capi
from pcbhip.serving import ServeRequest, SolveServer, latency_summary, poisson_arrivals, run_arrivals


class FakeOperator:
    def __init__(self, size: int, op_bytes: int = 1000):
        self.size = size
        self._bytes = op_bytes

    def device_bytes(self) -> int:
        return self._bytes


class FakeBackend:
    def __init__(self, max_slots: int, steps: int = 3):
        self.free = list(range(max_slots))
        self.steps = steps
        self.running: dict[int, list] = {}
        self.done: dict[int, tuple] = {}
        self.log: list[tuple[str, int]] = []

    def slot_bytes(self, op, max_outer):
        return 10 * op.size + max_outer

    def submit(self, op, rhs, x0, config, tag=0):
        if not self.free:
            raise _capi.ServerFull("every slot of the server is in use")
        slot = self.free.pop(0)
        self.running[slot] = [tag, self.steps, np.asarray(rhs, dtype=float)]
        self.log.append(("submit", tag))
        return slot

    def step(self):
        for slot in list(self.running):
            entry = self.running[slot]
            entry[1] -= 1
            if entry[1] == 0:
                self.done[slot] = (entry[0], entry[2])
                del self.running[slot]
        return len(self.running), len(self.done)

    def finished(self):
        return [(slot, tag) for slot, (tag, _) in sorted(self.done.items())]

    def collect(self, slot, solution=True):
        tag, rhs = self.done.pop(slot)
        self.free.append(slot)
        return (2.0 * rhs if solution else None), {"converged": True, "tag": tag}


CONFIG = _capi.config_c(1e-10, 0.0, 2e-3, 8, 200)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def _request(i, op, m, arrival=0.0):
    return ServeRequest(i, op, [np.full(op.size, float(i + k)) for k in range(m)], CONFIG, arrival_s=arrival)


def test_requests_complete_with_every_right_hand_side_in_order():
    backend = FakeBackend(max_slots=8, steps=2)
    server = SolveServer(max_slots=8, backend=backend, clock=Clock())
    op = FakeOperator(5)
    for i in range(4):
        server.submit(_request(i, op, m=3))
    results = server.drain()
    assert sorted(r.request_id for r in results) == [0, 1, 2, 3]
    for r in results:
        assert len(r.solutions) == 3
        for k, x in enumerate(r.solutions):
            np.testing.assert_array_equal(x, 2.0 * (r.request_id + k))
        assert r.converged
        assert r.finished_s >= r.admitted_s >= r.arrival_s
    # Slots never oversubscribed: 8 slots hold two 3-RHS requests at a time.
    assert server.peak_running_slots <= 8
    assert server.peak_running_requests == 2
    assert server.bytes_admitted == 0 and server.running_nodes == 0 and server.free_slots == 8


def test_memory_budget_holds_and_admission_is_fifo():
    backend = FakeBackend(max_slots=64, steps=2)
    op_small, op_big = FakeOperator(10, 100), FakeOperator(100, 100)
    budget = 100 + 2 * (10 * 100 + 8)  # the big operator plus two of its slots
    server = SolveServer(max_slots=64, memory_budget_bytes=budget, backend=backend, clock=Clock())
    server.submit(_request(0, op_big, m=2))
    server.submit(_request(1, op_big, m=1))   # waits: no memory left
    server.submit(_request(2, op_small, m=1))  # fits in memory, but FIFO keeps it behind request 1
    server.pump()
    assert list(server.inflight) == [0]
    assert [r.request_id for r in server.queue] == [1, 2]
    server.drain()
    assert server.peak_bytes <= budget
    order = [tag for what, tag in backend.log if what == "submit"]
    assert order == [0, 0, 1, 2]


def test_shared_operator_is_counted_once():
    backend = FakeBackend(max_slots=16, steps=5)
    op = FakeOperator(10, op_bytes=5000)
    server = SolveServer(max_slots=16, backend=backend, clock=Clock())
    server.submit(_request(0, op, m=1))
    server.submit(_request(1, op, m=2))
    server.pump()
    assert len(server.inflight) == 2
    assert server.bytes_admitted == 5000 + 3 * (10 * 10 + 8)
    server.drain()
    assert server.bytes_admitted == 0


def test_a_request_beyond_the_budget_or_the_slots_is_refused():
    server = SolveServer(max_slots=4, memory_budget_bytes=1000, backend=FakeBackend(4), clock=Clock())
    with pytest.raises(ValueError, match="memory budget"):
        server.submit(_request(0, FakeOperator(200, 10), m=1))
    with pytest.raises(ValueError, match="exceeds the 4 slots"):
        server.submit(_request(1, FakeOperator(1, 10), m=5))
    with pytest.raises(ValueError, match="at least one"):
        server.submit(ServeRequest(2, FakeOperator(1, 10), [], CONFIG))


def test_nodes_in_flight_are_capped_but_never_block_an_idle_server():
    backend = FakeBackend(max_slots=64, steps=3)
    server = SolveServer(max_slots=64, max_running_nodes=25, backend=backend, clock=Clock())
    op = FakeOperator(10)
    server.submit(_request(0, op, m=2))  # 20 nodes
    server.submit(_request(1, op, m=1))  # would make 30 > 25
    server.submit(_request(2, FakeOperator(40), m=1))  # alone exceeds the cap: admitted when idle
    server.pump()
    assert list(server.inflight) == [0]
    results = server.drain()
    assert sorted(r.request_id for r in results) == [0, 1, 2]
    assert server.peak_running_requests == 1


def test_latency_counts_from_arrival():
    clock = Clock()
    backend = FakeBackend(max_slots=4, steps=4)
    server = SolveServer(max_slots=4, backend=backend, clock=clock)
    op = FakeOperator(3)
    requests = [_request(i, op, m=2, arrival=0.0) for i in range(3)]
    elapsed = run_arrivals(server, requests)
    assert elapsed > 0
    assert len(server.completed) == 3
    lat = [r.latency_s for r in server.completed]
    # The third request waits for a pair of slots: its latency is the longest.
    assert max(lat) == server.completed[-1].latency_s
    summary = latency_summary(server.completed)
    assert summary["p50_ms"] <= summary["p99_ms"] <= summary["max_ms"]


def test_poisson_arrivals_have_the_requested_rate():
    rng = np.random.default_rng(5)
    times = poisson_arrivals(rng, 200.0, 30.0)
    assert np.all(np.diff(times) > 0) and times[-1] < 30.0
    assert abs(times.size / 30.0 - 200.0) < 200.0 * 0.05


class FailingBackend(FakeBackend):
    """Fails the collection of right-hand sides whose first entry is negative."""

    def collect(self, slot, solution=True):
        tag, rhs = self.done[slot]
        if rhs[0] < 0:
            self.done.pop(slot)
            self.free.append(slot)
            raise RuntimeError("inner PCG requires a finite symmetric positive-definite operator")
        return super().collect(slot, solution)


def test_served_solves_from_many_threads_return_each_result_to_its_caller():
    from concurrent.futures import ThreadPoolExecutor

    from pcbhip.serving import ServedSolves, route_solves, solve_mpir_routed

    op = FakeOperator(4)
    with ServedSolves(max_slots=4, backend=FailingBackend(max_slots=4, steps=3)) as served:
        def chain(k):
            # A chain of dependent solves, each started from the last answer.
            x = np.full(op.size, float(k + 1))
            with route_solves(served):
                for _ in range(3):
                    x, res = solve_mpir_routed(op, x, None, CONFIG)
                    assert res["converged"]
            return x

        with ThreadPoolExecutor(max_workers=6) as pool:
            out = list(pool.map(chain, range(6)))
        for k, x in enumerate(out):
            np.testing.assert_array_equal(x, 8.0 * (k + 1))
        # A failed solve raises in its own caller only; the server keeps serving.
        with route_solves(served), pytest.raises(RuntimeError, match="positive-definite"):
            solve_mpir_routed(op, np.full(op.size, -1.0), None, CONFIG)
        x, _ = served.solve(op, np.ones(op.size), None, CONFIG)
        np.testing.assert_array_equal(x, 2.0)
    with pytest.raises(RuntimeError, match="closed"):
        served.solve(op, np.ones(op.size), None, CONFIG)
