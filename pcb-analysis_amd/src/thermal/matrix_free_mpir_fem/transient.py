"""Transient conduction by backward Euler on the steady operator.

The reference interface is ``TimeSchedule``, ``TransientStep``,
``TransientThermalSolution`` and ``solve_thermal_transient`` of
src/thermal/matrix_free_mpir_fem/transient.py:45-258 (C++ march
cpp/src/thermal/thermal_solve.cpp:165-324): the same arguments, records,
stopping rule and messages.

With the element heat capacity lumped onto the nodes (``C``), one implicit
step solves ``(K + R + C/dt) T_{n+1} = q + R T_amb + (C/dt) T_n``: the steady
operator with ``C/dt`` added where the Robin conductance sits, so every step
is a steady MPIR solve on the HIP device, warm-started from ``T_n`` (with
radiating faces, a Newton loop of such solves).  ``BackwardEulerMarch`` keeps
the device operator of the last step length and reuses it while the step
repeats (a uniform schedule builds it once); the radiating case rebuilds per
Newton iterate because its Robin terms move.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from electrical.matrix_free_mpir_fem.solver import MPIRConfig

from .mesh import Preconditioner
from .operator import MatrixFreeThermalOperator
from .problem import ThermalConductionProblem
from .solve import LinearConduction, ThermalConductionSolution, radiation_fixed_point

# Steps generated per block while a geometric schedule is laid out.
_SCHEDULE_BLOCK = 256


@dataclass(frozen=True)
class TimeSchedule:
    """Time points ``0 = t_0 < t_1 < ... < t_N`` of a transient run, in seconds."""

    times_s: tuple[float, ...]

    def __post_init__(self) -> None:
        times = np.asarray(self.times_s, dtype=np.float64).reshape(-1)
        if times.size < 2 or times[0] != 0.0:
            raise ValueError("times_s must start at 0 and hold at least one step")
        if not (np.isfinite(times).all() and (np.diff(times) > 0.0).all()):
            raise ValueError("times_s must be finite and strictly increasing")
        object.__setattr__(self, "times_s", tuple(times.tolist()))

    @property
    def steps_s(self) -> np.ndarray:
        return np.diff(np.asarray(self.times_s))

    @property
    def end_s(self) -> float:
        return self.times_s[-1]

    @classmethod
    def uniform(cls, step_s: float, end_s: float) -> "TimeSchedule":
        """Equal steps of ``step_s`` up to the first multiple at or past ``end_s``
        (a point within 1e-9 of a step below ``end_s`` counts as reaching it)."""

        if not (step_s > 0.0 and end_s > 0.0):
            raise ValueError("step_s and end_s must be positive")
        count = int(np.ceil(end_s / step_s - 1.0e-9))
        return cls(tuple(np.linspace(0.0, count * step_s, count + 1).tolist()))

    @classmethod
    def geometric(cls, first_step_s: float, end_s: float, *, growth: float = 1.5,
                  max_step_s: float | None = None) -> "TimeSchedule":
        """Steps growing by ``growth`` from ``first_step_s`` (later steps at most
        ``max_step_s``), the last one trimmed to end at ``end_s``.

        The steps are running products and the times running sums, laid out a
        block at a time with ``ufunc.accumulate``, which rounds exactly as a
        step-by-step loop does.  A step past the cap stays at the cap, so
        capping the running product equals capping each step in turn.
        """

        if not (first_step_s > 0.0 and end_s > 0.0):
            raise ValueError("first_step_s and end_s must be positive")
        if growth < 1.0:
            raise ValueError("growth must be at least one")
        end = float(end_s)
        reached = end - 1.0e-12 * end
        cap = np.inf if max_step_s is None else float(max_step_s)
        factors = np.full(_SCHEDULE_BLOCK, float(growth))
        blocks = [np.zeros(1)]
        product, first = float(first_step_s), True
        while blocks[-1][-1] < reached:
            factors[0] = product
            with np.errstate(over="ignore"):  # products past the end of the run may overflow
                uncapped = np.multiply.accumulate(factors)
            steps = np.minimum(uncapped, cap)
            if first:
                steps[0], first = uncapped[0], False
            product = float(uncapped[-1]) * float(growth)
            times = np.add.accumulate(np.concatenate(([blocks[-1][-1]], steps)))[1:]
            done = np.flatnonzero(times >= reached)
            if done.size:
                times = times[: done[0] + 1]
                times[-1] = min(times[-1], end)
            blocks.append(times)
        return cls(tuple(np.concatenate(blocks).tolist()))


@dataclass(frozen=True)
class TransientStep:
    """One backward-Euler step: its time, the largest rise and change, the
    budget terms of the step's solve and its iteration counts."""

    index: int
    time_s: float
    step_s: float
    max_temperature_k: float
    max_change_k: float
    stored_heat_w: float
    convective_heat_w: float
    radiative_heat_w: float
    heat_balance_error_w: float
    inner_iterations: int
    radiation_iterations: int
    converged: bool


@dataclass(frozen=True)
class TransientThermalSolution:
    """Temperature history of a backward-Euler run and the last step's budget.
    ``temperature_k`` is ``(steps + 1, *node_shape)`` with ``store="all"``,
    else only the final field ``(1, *node_shape)``."""

    times_s: np.ndarray
    temperature_k: np.ndarray
    history: tuple[TransientStep, ...]
    final: ThermalConductionSolution
    reached_steady: bool

    @property
    def max_temperature_k(self) -> np.ndarray:
        return np.asarray([step.max_temperature_k for step in self.history])

    @property
    def final_temperature_k(self) -> np.ndarray:
        return self.temperature_k[-1]


def initial_field(problem: ThermalConductionProblem, initial_temperature_k) -> np.ndarray:
    """The field at ``t = 0``: the guess, or the first convective (else
    radiative) ambient, else the mean fixed temperature; fixed nodes at their
    values."""

    mesh = problem.mesh
    if initial_temperature_k is None:
        films = problem.convection or problem.radiation
        level = (films[0].mean_ambient_k() if films
                 else np.mean(problem.fixed_temperature_k[problem.fixed_temperature_mask]))
        start = np.full(mesh.node_shape, float(level))
    else:
        guess = np.asarray(initial_temperature_k, dtype=np.float64)
        start = np.array(np.broadcast_to(guess, mesh.node_shape) if guess.ndim == 0 else guess.reshape(mesh.node_shape))
        if not np.isfinite(start[mesh.active_nodes]).all():
            raise ValueError("initial_temperature_k must be finite on active nodes")
    return np.where(problem.fixed_temperature_mask, problem.fixed_temperature_k, start)


@dataclass
class FieldRecord:
    """The stored times and fields: every step, or only the latest."""

    keep_all: bool
    times: list = field(default_factory=list)
    fields: list = field(default_factory=list)

    def add(self, time_s: float, values: np.ndarray) -> None:
        if not self.keep_all:
            self.times.clear()
            self.fields.clear()
        self.times.append(float(time_s))
        self.fields.append(values)


class BackwardEulerMarch:
    """Implicit steps of one problem.  ``options`` are the fields of
    ``LinearConduction`` other than the capacity and previous field."""

    def __init__(self, problem: ThermalConductionProblem, options: dict, radiation_max_iterations: int,
                 radiation_tolerance_k: float):
        self.problem = problem
        self.options = options
        self.radiation = (radiation_max_iterations, radiation_tolerance_k)
        self.capacity = problem.mesh.nodal_heat_capacity_j_per_k().reshape(-1)
        self.active = problem.mesh.active_nodes
        self._operator: MatrixFreeThermalOperator | None = None
        self._operator_step: float | None = None

    def operator_for(self, linear: LinearConduction, step_s: float) -> MatrixFreeThermalOperator:
        if self._operator is None or self._operator_step != step_s:
            self._operator = None  # free the old device operator before building the new one
            self._operator = linear.operator(self.problem)
            self._operator_step = step_s
        return self._operator

    def step(self, current: np.ndarray, step_s: float) -> ThermalConductionSolution:
        linear = LinearConduction(capacity_per_s=self.capacity / step_s, previous_temperature_k=current, **self.options)
        if self.problem.radiation:
            return radiation_fixed_point(self.problem, linear, current, *self.radiation)
        return linear.run(self.problem, current, self.operator_for(linear, step_s))

    def change(self, before: np.ndarray, after: np.ndarray) -> float:
        """Largest change on active nodes."""

        return float(np.max(np.abs(after - before)[self.active])) if self.active.any() else 0.0


def solve_thermal_transient(
    problem: ThermalConductionProblem,
    schedule: TimeSchedule,
    *,
    initial_temperature_k: np.ndarray | float | None = None,
    store: str = "all",
    until_steady: bool = False,
    steady_tolerance_k_per_s: float = 1.0e-3,
    config: MPIRConfig | None = None,
    runtime=None,
    backend: str | None = None,
    device_id: int = 0,
    preconditioner: Preconditioner = "two-level",
    coarse_block_nodes: int | None = None,
    native: bool | None = None,
    radiation_max_iterations: int = 25,
    radiation_tolerance_k: float = 1.0e-4,
) -> TransientThermalSolution:
    """March ``problem`` through ``schedule`` by backward Euler.  The mesh
    must carry a volumetric heat capacity.  ``store="all"`` keeps every
    field, ``"final"`` only the last; ``until_steady`` stops once
    ``max |T_{n+1} - T_n| / dt`` falls to ``steady_tolerance_k_per_s``."""

    if store not in ("all", "final"):
        raise ValueError("store must be 'all' or 'final'")
    if not steady_tolerance_k_per_s > 0.0:
        raise ValueError("steady_tolerance_k_per_s must be positive")
    march = BackwardEulerMarch(
        problem,
        dict(config=config, runtime=runtime, backend=backend, device_id=device_id, preconditioner=preconditioner,
             coarse_block_nodes=coarse_block_nodes, native=native),
        radiation_max_iterations, radiation_tolerance_k,
    )
    current = initial_field(problem, initial_temperature_k)
    record = FieldRecord(store == "all")
    record.add(0.0, np.where(march.active, current, np.nan))
    history: list[TransientStep] = []
    solution, steady = None, False
    for index, (time_s, step_s) in enumerate(zip(schedule.times_s[1:], schedule.steps_s.tolist()), start=1):
        solution = march.step(current, step_s)
        after = np.where(np.isfinite(solution.temperature_k), solution.temperature_k, current)
        change = march.change(current, after)
        current = after
        shown = np.where(march.active, current, np.nan)
        record.add(time_s, shown)
        history.append(TransientStep(
            index=index, time_s=float(time_s), step_s=step_s, max_temperature_k=float(np.nanmax(shown)),
            max_change_k=change, stored_heat_w=solution.stored_heat_w,
            convective_heat_w=float(np.sum(solution.convective_heat_w)),
            radiative_heat_w=float(np.sum(solution.radiative_heat_w)),
            heat_balance_error_w=solution.heat_balance_error_w, inner_iterations=solution.solve.inner_iterations,
            radiation_iterations=solution.radiation_iterations,
            converged=solution.solve.converged and solution.radiation_converged,
        ))
        steady = until_steady and change / step_s <= steady_tolerance_k_per_s
        if steady:
            break
    return TransientThermalSolution(
        times_s=np.asarray(record.times),
        temperature_k=np.stack(record.fields),
        history=tuple(history),
        final=solution,
        reached_steady=bool(steady),
    )


__all__ = ["TimeSchedule", "TransientStep", "TransientThermalSolution", "solve_thermal_transient"]
