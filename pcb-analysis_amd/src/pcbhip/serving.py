"""Continuous-batching serving of MPIR conduction solves on one HIP device.

A request is one resident operator (a thermal board or a DC conductor) and
``m`` right-hand sides.  ``SolveServer`` keeps a FIFO queue of requests and a
native server (``pcbhip_server``) of slots.  Each ``pump``:

* admits queued requests, oldest first, while the slots, the device-memory
  budget (resident operators plus the vectors of every admitted slot) and the
  optional cap on nodes in flight allow; a request is admitted whole;
* runs one replay boundary of the native server, which retires the oldest
  replay in flight, puts the admitted slots into the tile tables and launches
  the next captured graph of MPIR steps while the previous one still runs;
* collects the slots that ended and completes a request when all of its
  right-hand sides have, stamping its latency from arrival.

Every slot's arithmetic is that of a solve alone (fixed-order reductions per
slot, a state machine per slot), so a result does not depend on what else
was in flight.  The reference has no serving layer: its callers solve one
right-hand side at a time (src/multiphysics/staggered_coupling/
electro_thermal.py:403-423, cpp/src/fem/dc_ports.cpp:26-49).
"""

from __future__ import annotations

import threading
import time
from collections import deque
from contextlib import contextmanager
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _capi


@dataclass
class ServeRequest:
    """``m`` right-hand sides of one resident operator."""

    request_id: int
    operator: _capi.Operator
    rhs: Sequence[np.ndarray]
    config: _capi.MpirConfigC
    kind: str = "thermal"
    arrival_s: float = 0.0
    initial_guess: Sequence[np.ndarray] | None = None

    @property
    def nodes(self) -> int:
        return int(self.operator.size) * len(self.rhs)


@dataclass
class ServeResult:
    request_id: int
    kind: str
    solutions: list
    results: list
    arrival_s: float
    admitted_s: float
    finished_s: float
    error: BaseException | None = None  # the first failed right-hand side's error (e.g. not SPD)

    @property
    def latency_s(self) -> float:
        return self.finished_s - self.arrival_s

    @property
    def converged(self) -> bool:
        return self.error is None and all(r["converged"] for r in self.results)


@dataclass
class _InFlight:
    request: ServeRequest
    admitted_s: float
    slot_bytes: int
    solutions: list = field(default_factory=list)
    results: list = field(default_factory=list)
    left: int = 0
    error: BaseException | None = None


class SolveServer:
    """FIFO admission by slots, memory and nodes in flight over a native
    continuous-batching server.

    ``memory_budget_bytes`` bounds the device bytes of the operators of
    admitted requests (each counted once while any admitted request uses it)
    plus every admitted slot's vectors.  ``max_running_nodes`` (optional)
    bounds the nodes times right-hand sides in flight: past the point where
    the device is full, more work in flight only stretches every request's
    latency.  ``backend`` is the native server (a ``_capi.Server`` unless a
    test passes a stand-in with the same methods).  ``tune_with``, a list of
    device operators typical of the traffic, picks the tile shape by timing
    the fused pass over them (``autotune.tune_server``, cached per set of
    board shapes) instead of ``rows_inner`` / ``target_groups``."""

    def __init__(self, device: int = 0, *, max_slots: int = 512, memory_budget_bytes: float = 200e9,
                 steps_per_replay: int = 16, max_running_nodes: int | None = None, keep_solutions: bool = True,
                 rows_inner: int = 4, target_groups: int = 1024, tune_with=None, backend=None,
                 clock=time.perf_counter):
        self.backend = backend if backend is not None else _capi.Server(device, max_slots, False, steps_per_replay)
        self.shape = {"rows": int(rows_inner), "target": int(target_groups), "tuned": False}
        if tune_with:
            from . import autotune

            chosen = autotune.tune_server(self.backend, tune_with)
            self.shape = {"rows": chosen["rows"], "target": chosen["target"], "tuned": True,
                          "cached": chosen["cached"]}
        elif backend is None:
            self.backend.set_shape(rows_inner, target_groups)
        self.max_slots = int(max_slots)
        self.memory_budget_bytes = float(memory_budget_bytes)
        self.max_running_nodes = None if max_running_nodes is None else int(max_running_nodes)
        self.keep_solutions = bool(keep_solutions)
        self.clock = clock
        self.queue: deque[ServeRequest] = deque()
        self.inflight: dict[int, _InFlight] = {}
        self._slot_of: dict[int, tuple[int, int]] = {}  # native slot -> (request id, rhs index)
        self._op_refs: dict[int, list] = {}  # id(operator) -> [refcount, bytes]
        self.free_slots = self.max_slots
        self.bytes_admitted = 0
        self.running_nodes = 0
        self.peak_bytes = 0
        self.peak_running_requests = 0
        self.peak_running_slots = 0
        self.completed: list[ServeResult] = []
        self.pumps = 0
        self.pump_s = 0.0  # wall time inside pump(), the native step included
        self.step_s = 0.0  # wall time inside the native step (it blocks on the oldest replay in flight)

    # -- admission ---------------------------------------------------------------

    def footprint(self, request: ServeRequest) -> tuple[int, int]:
        """(operator bytes, bytes of all the request's slots)."""

        op_bytes = int(request.operator.device_bytes())
        slot = int(self.backend.slot_bytes(request.operator, int(request.config.max_outer_iterations)))
        return op_bytes, slot * len(request.rhs)

    def submit(self, request: ServeRequest) -> None:
        if not request.rhs:
            raise ValueError("a request needs at least one right-hand side")
        if len(request.rhs) > self.max_slots:
            raise ValueError(f"a request of {len(request.rhs)} right-hand sides exceeds the {self.max_slots} slots")
        op_bytes, slot_bytes = self.footprint(request)
        if op_bytes + slot_bytes > self.memory_budget_bytes:
            raise ValueError("a request's device footprint exceeds the memory budget")
        self.queue.append(request)

    def _fits(self, request: ServeRequest) -> bool:
        m = len(request.rhs)
        if m > self.free_slots:
            return False
        op_bytes, slot_bytes = self.footprint(request)
        extra = slot_bytes + (0 if id(request.operator) in self._op_refs else op_bytes)
        if self.bytes_admitted + extra > self.memory_budget_bytes:
            return False
        if self.max_running_nodes is not None and self.inflight and \
                self.running_nodes + request.nodes > self.max_running_nodes:
            return False
        return True

    def _admit(self, request: ServeRequest, now: float) -> None:
        op_bytes, slot_bytes = self.footprint(request)
        key = id(request.operator)
        if key in self._op_refs:
            self._op_refs[key][0] += 1
        else:
            self._op_refs[key] = [1, op_bytes]
            self.bytes_admitted += op_bytes
        self.bytes_admitted += slot_bytes
        m = len(request.rhs)
        entry = _InFlight(request, now, slot_bytes, [None] * m, [None] * m, m)
        self.inflight[request.request_id] = entry
        for k, rhs in enumerate(request.rhs):
            x0 = None if request.initial_guess is None else request.initial_guess[k]
            slot = self.backend.submit(request.operator, rhs, x0, request.config, tag=request.request_id)
            self._slot_of[slot] = (request.request_id, k)
        self.free_slots -= m
        self.running_nodes += request.nodes
        self.peak_bytes = max(self.peak_bytes, self.bytes_admitted)
        self.peak_running_requests = max(self.peak_running_requests, len(self.inflight))
        self.peak_running_slots = max(self.peak_running_slots, self.max_slots - self.free_slots)

    def _release(self, entry: _InFlight) -> None:
        request = entry.request
        key = id(request.operator)
        ref = self._op_refs[key]
        ref[0] -= 1
        if ref[0] == 0:
            self.bytes_admitted -= ref[1]
            del self._op_refs[key]
        self.bytes_admitted -= entry.slot_bytes
        self.free_slots += len(request.rhs)
        self.running_nodes -= request.nodes

    # -- the loop ----------------------------------------------------------------

    @property
    def idle(self) -> bool:
        return not self.queue and not self.inflight

    def pump(self) -> list[ServeResult]:
        """Admit, run one replay boundary, collect; returns the requests that
        completed in this call."""

        start = time.perf_counter()
        try:
            return self._pump()
        finally:
            self.pumps += 1
            self.pump_s += time.perf_counter() - start

    def _pump(self) -> list[ServeResult]:
        now = self.clock()
        while self.queue and self._fits(self.queue[0]):
            self._admit(self.queue.popleft(), now)
        if not self.inflight:
            return []
        t0 = time.perf_counter()
        self.backend.step()
        self.step_s += time.perf_counter() - t0
        done: list[ServeResult] = []
        for slot, _tag in self.backend.finished():
            request_id, k = self._slot_of.pop(slot)
            entry = self.inflight[request_id]
            try:
                x, res = self.backend.collect(slot, solution=self.keep_solutions)
            except RuntimeError as exc:  # this right-hand side failed; its slot is freed
                x, res = None, None
                entry.error = entry.error or exc
            entry.solutions[k] = x
            entry.results[k] = res
            entry.left -= 1
            if entry.left == 0:
                finished = self.clock()
                del self.inflight[request_id]
                self._release(entry)
                result = ServeResult(request_id, entry.request.kind, entry.solutions, entry.results,
                                     entry.request.arrival_s, entry.admitted_s, finished, entry.error)
                done.append(result)
        self.completed.extend(done)
        return done

    def drain(self) -> list[ServeResult]:
        out: list[ServeResult] = []
        while not self.idle:
            out.extend(self.pump())
        return out

    def close(self) -> None:
        close = getattr(self.backend, "close", None)
        if close is not None:
            close()


class ServedSolves:
    """Blocking MPIR solves from many host threads, continuously batched.

    ``solve`` submits one right-hand side to a shared :class:`SolveServer`
    and waits for it; a pump thread runs the server's replay boundaries
    while any solve is queued or in flight.  Host threads that each run a
    chain of dependent solves (an electro-thermal loop alternating DC and
    thermal solves, say) thus keep their solves in flight together.  A
    solve's result is bitwise the result of the single-solve path."""

    def __init__(self, device: int = 0, **server_options) -> None:
        self.server = SolveServer(device, keep_solutions=True, **server_options)
        self._lock = threading.Lock()
        self._wake = threading.Condition(self._lock)
        self._waiting: dict[int, list] = {}
        self._next_id = 0
        self._closed = False
        self._failure: BaseException | None = None
        self._pump = threading.Thread(target=self._run, name="served-solves", daemon=True)
        self._pump.start()

    def solve(self, operator: _capi.Operator, rhs, x0, config: _capi.MpirConfigC, kind: str = "solve"):
        """(solution, result dict) of one right-hand side, as ``_capi.solve_mpir`` returns them."""

        done = threading.Event()
        slot: list = [done, None]
        with self._lock:
            if self._failure is not None:
                raise RuntimeError("the served-solve pump failed") from self._failure
            if self._closed:
                raise RuntimeError("served solves are closed")
            request_id = self._next_id
            self._next_id += 1
            self._waiting[request_id] = slot
            self.server.submit(ServeRequest(request_id, operator, [np.asarray(rhs, dtype=np.float64)], config,
                                            kind=kind, arrival_s=self.server.clock(),
                                            initial_guess=None if x0 is None else [x0]))
            self._wake.notify()
        done.wait()
        result = slot[1]
        if isinstance(result, BaseException):
            raise RuntimeError("the served-solve pump failed") from result
        if result.error is not None:
            raise result.error
        return result.solutions[0], result.results[0]

    def _run(self) -> None:
        with self._lock:
            while True:
                while self.server.idle and not self._closed:
                    self._wake.wait()
                if self.server.idle:
                    return
                try:
                    done = self.server.pump()
                except BaseException as exc:  # fail every waiter rather than hang them
                    self._failure = exc
                    for slot in self._waiting.values():
                        slot[1] = exc
                        slot[0].set()
                    self._waiting.clear()
                    return
                for result in done:
                    slot = self._waiting.pop(result.request_id)
                    slot[1] = result
                    slot[0].set()

    def close(self) -> None:
        with self._lock:
            self._closed = True
            self._wake.notify()
        self._pump.join()
        self.server.close()

    def __enter__(self) -> "ServedSolves":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


_route = threading.local()


@contextmanager
def route_solves(executor: ServedSolves | None):
    """Sends this thread's device MPIR solves (``solve_mpir_routed``) to
    ``executor`` for the duration of the block."""

    previous = getattr(_route, "executor", None)
    _route.executor = executor
    try:
        yield executor
    finally:
        _route.executor = previous


def solve_mpir_routed(operator: _capi.Operator, rhs, x0, config: _capi.MpirConfigC):
    """One device MPIR solve: through this thread's :class:`ServedSolves` when
    one is routed, else through the single-solve path ``_capi.solve_mpir``."""

    executor = getattr(_route, "executor", None)
    if executor is None:
        return _capi.solve_mpir(operator, rhs, x0, config)
    return executor.solve(operator, rhs, x0, config)


def run_arrivals(server: SolveServer, requests: Sequence[ServeRequest], *, log=None,
                 time_limit_s: float | None = None) -> float:
    """Feeds ``requests`` (``arrival_s`` relative to the start, ascending)
    to ``server`` at their arrival times and pumps until all complete, or
    until ``time_limit_s`` has passed (requests still queued or in flight
    then stay uncompleted).  Arrival times are rewritten to the server clock.
    Returns the wall time."""

    start = server.clock()
    pending = deque(requests)
    last_log = start
    while pending or not server.idle:
        now = server.clock()
        if time_limit_s is not None and now - start > time_limit_s:
            break
        while pending and pending[0].arrival_s <= now - start:
            r = pending.popleft()
            r.arrival_s = start + r.arrival_s
            server.submit(r)
        if server.idle:
            if pending:
                time.sleep(max(0.0, min(0.01, pending[0].arrival_s - (server.clock() - start))))
            continue
        server.pump()
        if log is not None and now - last_log > 5.0:
            last_log = now
            log(f"t={now - start:.1f}s queued={len(server.queue)} in flight={len(server.inflight)} "
                f"completed={len(server.completed)}")
    return server.clock() - start


def latency_summary(results: Sequence[ServeResult]) -> dict:
    lat = np.sort(np.asarray([r.latency_s for r in results]))
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None, "max_ms": None}
    return {
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "max_ms": float(lat[-1] * 1e3),
    }


def poisson_arrivals(rng: np.random.Generator, rate_per_s: float, duration_s: float) -> np.ndarray:
    """Arrival times of a Poisson process of ``rate_per_s`` over ``duration_s``."""

    gaps = rng.exponential(1.0 / rate_per_s, size=int(rate_per_s * duration_s * 1.5) + 16)
    times = np.cumsum(gaps)
    return times[times < duration_s]
