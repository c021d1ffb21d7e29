"""Steady conduction solve with the heat budget (reference:
src/thermal/matrix_free_mpir_fem/solve.py:19-280; C++ twin
cpp/src/thermal/thermal_solve.cpp:49-163).

The solve works on the rise above a reference temperature (the first
convective ambient, else the mean fixed temperature), runs MPIR on the HIP
device, re-references once to the mean rise if the first pass did not
converge, and reports the heat budget; radiating faces are Newton-linearised
around the last iterate until the field stops moving.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np

from electrical.matrix_free_mpir_fem.solver import MPIRConfig, MPIRResult, default_config, solve_mpir

from .mesh import Preconditioner
from .operator import MatrixFreeThermalOperator
from .problem import ThermalConductionProblem


@dataclass(frozen=True)
class ThermalConductionSolution:
    """Nodal temperatures and the heat budget of one steady solve (solve.py:19-36)."""

    temperature_k: np.ndarray
    heat_flux_w_per_m2: np.ndarray
    max_temperature_k: float
    min_temperature_k: float
    total_heat_input_w: float
    convective_heat_w: np.ndarray
    fixed_temperature_heat_w: float
    heat_balance_error_w: float
    solve: MPIRResult
    radiative_heat_w: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    radiation_iterations: int = 0
    radiation_converged: bool = True
    radiation_change_k: float = 0.0
    stored_heat_w: float = 0.0


def solve_thermal_conduction(
    problem: ThermalConductionProblem,
    *,
    config: MPIRConfig | None = None,
    runtime=None,
    backend: str | None = None,
    device_id: int = 0,
    initial_temperature_k: np.ndarray | float | None = None,
    preconditioner: Preconditioner = "two-level",
    coarse_block_nodes: int | None = None,
    reference_temperature_k: float | None = None,
    native: bool | None = None,
    radiation_max_iterations: int = 25,
    radiation_tolerance_k: float = 1.0e-4,
) -> ThermalConductionSolution:
    """Solve steady heat conduction in a layered PCB with matrix-free MPIR on the HIP device."""

    return _solve(
        problem,
        config=config,
        runtime=runtime,
        backend=backend,
        device_id=device_id,
        initial_temperature_k=initial_temperature_k,
        preconditioner=preconditioner,
        coarse_block_nodes=coarse_block_nodes,
        reference_temperature_k=reference_temperature_k,
        native=native,
        radiation_max_iterations=radiation_max_iterations,
        radiation_tolerance_k=radiation_tolerance_k,
    )


def _solve(
    problem: ThermalConductionProblem,
    *,
    config: MPIRConfig | None = None,
    runtime=None,
    backend: str | None = None,
    device_id: int = 0,
    initial_temperature_k=None,
    preconditioner: Preconditioner = "two-level",
    coarse_block_nodes: int | None = None,
    reference_temperature_k: float | None = None,
    native: bool | None = None,
    radiation_max_iterations: int = 25,
    radiation_tolerance_k: float = 1.0e-4,
    capacity_per_s: np.ndarray | None = None,
    previous_temperature_k: np.ndarray | None = None,
) -> ThermalConductionSolution:
    """A steady solve, or one backward-Euler step when ``capacity_per_s`` is
    given; radiating faces by Newton iteration on their linearisation
    (solve.py:80-168): each iteration adds the linearised faces as convection
    faces and solves the linear problem warm-started from the last iterate,
    until no node moves more than ``radiation_tolerance_k``."""

    linear_options = dict(
        config=config, runtime=runtime, backend=backend, device_id=device_id, preconditioner=preconditioner,
        coarse_block_nodes=coarse_block_nodes, reference_temperature_k=reference_temperature_k, native=native,
        capacity_per_s=capacity_per_s, previous_temperature_k=previous_temperature_k,
    )
    if not problem.radiation:
        return _solve_linear(problem, initial_temperature_k=initial_temperature_k, **linear_options)
    if radiation_max_iterations < 1:
        raise ValueError("radiation_max_iterations must be positive")
    if not radiation_tolerance_k > 0.0:
        raise ValueError("radiation_tolerance_k must be positive")
    mesh = problem.mesh
    fallback = problem.convection[0].mean_ambient_k() if problem.convection else problem.radiation[0].mean_ambient_k()
    if initial_temperature_k is None:
        current = np.full(mesh.node_shape, float(fallback))
    else:
        guess = np.asarray(initial_temperature_k, dtype=np.float64)
        current = np.full(mesh.node_shape, float(guess)) if guess.ndim == 0 else guess.reshape(mesh.node_shape).copy()
        current = np.where(np.isfinite(current), current, problem.radiation[0].mean_ambient_k())
    n_convection = len(problem.convection)
    solution = None
    change = float("inf")
    converged = False
    iteration = 0
    for iteration in range(1, radiation_max_iterations + 1):
        faces = tuple(boundary.linearize(mesh, current) for boundary in problem.radiation)
        linear = dataclasses.replace(problem, convection=tuple(problem.convection) + faces, radiation=())
        solution = _solve_linear(linear, initial_temperature_k=current, **linear_options)
        proposed = np.where(np.isfinite(solution.temperature_k), solution.temperature_k, current)
        change = float(np.max(np.abs(proposed - current)))
        current = proposed
        if change <= radiation_tolerance_k:
            converged = True
            break
    assert solution is not None
    heat = solution.convective_heat_w
    return dataclasses.replace(
        solution,
        convective_heat_w=heat[:n_convection],
        radiative_heat_w=heat[n_convection:],
        radiation_iterations=iteration,
        radiation_converged=converged and solution.solve.converged,
        radiation_change_k=change,
    )


def merged_result(first: MPIRResult, resumed: MPIRResult) -> MPIRResult:
    """Counts of a re-referenced solve summed over both passes (solve.py:240-253)."""

    return MPIRResult(
        solution=resumed.solution,
        converged=resumed.converged,
        outer_iterations=first.outer_iterations + resumed.outer_iterations,
        inner_iterations=first.inner_iterations + resumed.inner_iterations,
        relative_residual=resumed.relative_residual,
        high_operator_applications=first.high_operator_applications + resumed.high_operator_applications,
        low_operator_applications=first.low_operator_applications + resumed.low_operator_applications,
        low_runtime=resumed.low_runtime,
        history=first.history + resumed.history,
    )


def initial_rise(operator: MatrixFreeThermalOperator, initial_temperature_k, reference: float) -> np.ndarray | None:
    """The warm start as a rise above ``reference``, fixed rows at their values (solve.py:206-222)."""

    if initial_temperature_k is None:
        return None
    guess = np.asarray(initial_temperature_k, dtype=np.float64)
    start = np.full(operator.size, float(guess)) if guess.ndim == 0 else guess.reshape(-1).copy()
    if start.size != operator.size:
        raise ValueError("initial_temperature_k must hold one value per node")
    return (
        np.where(
            operator.fixed_nodes.reshape(-1),
            operator.fixed_temperature_k.reshape(-1) + reference * ~operator.active_nodes.reshape(-1),
            start,
        )
        - reference
    )


def _solve_linear(
    problem: ThermalConductionProblem,
    *,
    config: MPIRConfig | None,
    runtime,
    backend: str | None,
    device_id: int,
    initial_temperature_k,
    preconditioner: Preconditioner,
    coarse_block_nodes: int | None,
    reference_temperature_k: float | None,
    native: bool | None,
    capacity_per_s: np.ndarray | None = None,
    previous_temperature_k: np.ndarray | None = None,
    operator: MatrixFreeThermalOperator | None = None,
) -> ThermalConductionSolution:
    """One linear solve and its heat budget.  ``operator`` (optional) is a
    prepared operator of ``problem`` with this ``capacity_per_s``, reused by a
    transient march while its step repeats."""

    if operator is None:
        operator = MatrixFreeThermalOperator(
            problem,
            runtime=runtime,
            backend=backend,
            device_id=device_id,
            preconditioner=preconditioner,
            coarse_block_nodes=coarse_block_nodes,
            native=native,
            capacity_per_s=capacity_per_s,
        )
    previous = None
    if previous_temperature_k is not None:
        previous = np.asarray(previous_temperature_k, dtype=np.float64).reshape(-1)
        previous = np.where(np.isfinite(previous), previous, 0.0)
    reference = (
        operator.default_reference_temperature() if reference_temperature_k is None else float(reference_temperature_k)
    )
    if not np.isfinite(reference):
        raise ValueError("reference_temperature_k must be finite")
    rhs = operator.build_rhs(reference, previous)
    initial = initial_rise(operator, initial_temperature_k, reference)
    if config is None:
        # A stiff copper/laminate stack lets the FP32 inner solve buy only
        # about one decade per outer step (solve.py:223-226): 16 outer steps,
        # and with Jacobi the Jacobi inner budget.
        config = default_config(operator.preconditioner, 16)
    result = solve_mpir(operator, rhs, config=config, initial_guess=initial)
    free = operator.free_nodes.reshape(-1)
    if not result.converged and np.any(free):
        # Re-reference to the mean free rise and continue (solve.py:228-253).
        shift = float(np.mean(result.solution[free]))
        if abs(shift) > 0.0:
            reference += shift
            rhs = operator.build_rhs(reference, previous)
            resumed = solve_mpir(operator, rhs, config=config, initial_guess=result.solution - shift)
            result = merged_result(result, resumed)
    temperature = result.solution.reshape(problem.mesh.node_shape) + reference

    load = operator.nodal_load()
    residual = operator.unconstrained_residual(temperature, previous)
    fixed = operator.fixed_nodes.reshape(-1)
    convective = operator.convective_heat(temperature)
    fixed_heat = -float(np.sum(residual[fixed]))
    stored = float(np.sum(operator.stored_heat_w(temperature, previous)))
    total_input = float(np.sum(load))
    heat_flux = operator.element_heat_flux(temperature)
    if not problem.mesh.is_full:
        heat_flux = np.where(problem.mesh.active[..., None], heat_flux, 0.0)  # type: ignore[index]
        temperature = np.where(operator.active_nodes, temperature, np.nan)
    return ThermalConductionSolution(
        temperature_k=temperature,
        heat_flux_w_per_m2=heat_flux,
        max_temperature_k=float(np.nanmax(temperature)),
        min_temperature_k=float(np.nanmin(temperature)),
        total_heat_input_w=total_input,
        convective_heat_w=convective,
        fixed_temperature_heat_w=fixed_heat,
        heat_balance_error_w=total_input - float(np.sum(convective)) - fixed_heat - stored,
        solve=result,
        stored_heat_w=stored,
    )
