"""Transient conduction by backward Euler on the steady operator (reference:
src/thermal/matrix_free_mpir_fem/transient.py:1-258; C++ march
cpp/src/thermal/thermal_solve.cpp:165-324).

With the element heat capacity lumped onto the nodes (``C``), one implicit
step solves ``(K + R + C/dt) T_{n+1} = q + R T_amb + (C/dt) T_n``: the steady
operator with ``C/dt`` added where the Robin conductance sits, so every step
is a steady MPIR solve on the HIP device, warm-started from ``T_n`` (with
radiating faces, a Newton loop of such solves).  Without radiation the
device operator is prepared once per distinct step length and reused while
the step repeats (a uniform schedule builds it once).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from electrical.matrix_free_mpir_fem.solver import MPIRConfig

from .mesh import Preconditioner
from .operator import MatrixFreeThermalOperator
from .problem import ThermalConductionProblem
from .solve import ThermalConductionSolution, _solve, _solve_linear


@dataclass(frozen=True)
class TimeSchedule:
    """Time points ``0 = t_0 < t_1 < ... < t_N`` of a transient run, in seconds
    (transient.py:45-100)."""

    times_s: tuple[float, ...]

    def __post_init__(self) -> None:
        times = np.asarray(self.times_s, dtype=np.float64).reshape(-1)
        if times.size < 2 or times[0] != 0.0:
            raise ValueError("times_s must start at 0 and hold at least one step")
        if not np.all(np.isfinite(times)) or np.any(np.diff(times) <= 0.0):
            raise ValueError("times_s must be finite and strictly increasing")
        object.__setattr__(self, "times_s", tuple(float(t) for t in times))

    @property
    def steps_s(self) -> np.ndarray:
        return np.diff(np.asarray(self.times_s))

    @property
    def end_s(self) -> float:
        return self.times_s[-1]

    @classmethod
    def uniform(cls, step_s: float, end_s: float) -> "TimeSchedule":
        """Equal steps of ``step_s`` up to the first multiple at or past ``end_s``."""

        if not (step_s > 0.0 and end_s > 0.0):
            raise ValueError("step_s and end_s must be positive")
        count = int(np.ceil(end_s / step_s - 1.0e-9))
        return cls(tuple(float(v) for v in np.linspace(0.0, count * step_s, count + 1)))

    @classmethod
    def geometric(cls, first_step_s: float, end_s: float, *, growth: float = 1.5,
                  max_step_s: float | None = None) -> "TimeSchedule":
        """Steps growing by ``growth`` from ``first_step_s`` (at most
        ``max_step_s``), the last one trimmed to end at ``end_s``."""

        if not (first_step_s > 0.0 and end_s > 0.0):
            raise ValueError("first_step_s and end_s must be positive")
        if growth < 1.0:
            raise ValueError("growth must be at least one")
        times = [0.0]
        step = float(first_step_s)
        while times[-1] < end_s * (1.0 - 1.0e-12):
            times.append(min(times[-1] + step, float(end_s)))
            step = step * growth if max_step_s is None else min(step * growth, float(max_step_s))
        return cls(tuple(times))


@dataclass(frozen=True)
class TransientStep:
    """Record of one step (transient.py:103-117)."""

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
    """Temperature history of a backward-Euler run and the last step's budget
    (transient.py:120-137).  ``temperature_k`` is ``(steps + 1, *node_shape)``
    with ``store="all"``, else only the final field ``(1, *node_shape)``."""

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


def _start_temperature(problem: ThermalConductionProblem, initial_temperature_k) -> np.ndarray:
    """The initial field: the first convective (else radiative) ambient, else
    the mean fixed temperature (transient.py:166-180); fixed nodes at their
    values."""

    mesh = problem.mesh
    if initial_temperature_k is None:
        if problem.convection:
            start = float(problem.convection[0].mean_ambient_k())
        elif problem.radiation:
            start = float(problem.radiation[0].mean_ambient_k())
        else:
            mask = problem.fixed_temperature_mask
            start = float(np.mean(np.asarray(problem.fixed_temperature_k)[mask]))
        current = np.full(mesh.node_shape, start)
    else:
        guess = np.asarray(initial_temperature_k, dtype=np.float64)
        current = np.full(mesh.node_shape, float(guess)) if guess.ndim == 0 else guess.reshape(mesh.node_shape).copy()
        if not np.all(np.isfinite(current[mesh.active_nodes])):
            raise ValueError("initial_temperature_k must be finite on active nodes")
    fixed_values = np.broadcast_to(np.asarray(problem.fixed_temperature_k, dtype=np.float64), mesh.node_shape)
    return np.where(problem.fixed_temperature_mask, fixed_values, current)


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
    """March ``problem`` through ``schedule`` by backward Euler
    (transient.py:140-255).  ``store="all"`` keeps every field, ``"final"``
    only the last; ``until_steady`` stops once ``max |T_{n+1} - T_n| / dt``
    falls to ``steady_tolerance_k_per_s``."""

    mesh = problem.mesh
    if store not in ("all", "final"):
        raise ValueError("store must be 'all' or 'final'")
    if not steady_tolerance_k_per_s > 0.0:
        raise ValueError("steady_tolerance_k_per_s must be positive")
    capacity = mesh.nodal_heat_capacity_j_per_k().reshape(-1)
    current = _start_temperature(problem, initial_temperature_k)
    active = mesh.active_nodes
    fields = [np.where(active, current, np.nan)]
    times = [0.0]
    history: list[TransientStep] = []
    final: ThermalConductionSolution | None = None
    reached_steady = False
    cached_step = None
    operator = None
    for index, (time_s, step_s) in enumerate(zip(schedule.times_s[1:], schedule.steps_s), start=1):
        step_s = float(step_s)
        options = dict(config=config, runtime=runtime, backend=backend, device_id=device_id,
                       initial_temperature_k=current, preconditioner=preconditioner,
                       coarse_block_nodes=coarse_block_nodes, reference_temperature_k=None, native=native,
                       capacity_per_s=capacity / step_s, previous_temperature_k=current)
        if problem.radiation:
            # The Robin terms change with every Newton iterate: no reuse.
            final = _solve(problem, radiation_max_iterations=radiation_max_iterations,
                           radiation_tolerance_k=radiation_tolerance_k, **options)
        else:
            if operator is None or step_s != cached_step:
                operator = MatrixFreeThermalOperator(
                    problem, runtime=runtime, backend=backend, device_id=device_id, preconditioner=preconditioner,
                    coarse_block_nodes=coarse_block_nodes, native=native, capacity_per_s=capacity / step_s,
                )
                cached_step = step_s
            final = _solve_linear(problem, operator=operator, **options)
        proposed = np.where(np.isfinite(final.temperature_k), final.temperature_k, current)
        change = float(np.max(np.abs(proposed - current)[active])) if np.any(active) else 0.0
        current = proposed
        times.append(float(time_s))
        field = np.where(active, current, np.nan)
        if store == "all":
            fields.append(field)
        else:
            fields = [field]
        history.append(TransientStep(
            index=index,
            time_s=float(time_s),
            step_s=step_s,
            max_temperature_k=float(np.nanmax(field)),
            max_change_k=change,
            stored_heat_w=final.stored_heat_w,
            convective_heat_w=float(np.sum(final.convective_heat_w)),
            radiative_heat_w=float(np.sum(final.radiative_heat_w)),
            heat_balance_error_w=final.heat_balance_error_w,
            inner_iterations=final.solve.inner_iterations,
            radiation_iterations=final.radiation_iterations,
            converged=final.solve.converged and final.radiation_converged,
        ))
        if until_steady and change / step_s <= steady_tolerance_k_per_s:
            reached_steady = True
            break
    assert final is not None
    return TransientThermalSolution(
        times_s=np.asarray(times) if store == "all" else np.asarray(times[-1:]),
        temperature_k=np.stack(fields),
        history=tuple(history),
        final=final,
        reached_steady=reached_steady,
    )


__all__ = ["TimeSchedule", "TransientStep", "TransientThermalSolution", "solve_thermal_transient"]
