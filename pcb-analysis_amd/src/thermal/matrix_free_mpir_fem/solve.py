"""Steady conduction solve and its heat budget.

The reference interface is ``solve_thermal_conduction`` and
``ThermalConductionSolution`` of src/thermal/matrix_free_mpir_fem/solve.py:19-280
(C++ twin cpp/src/thermal/thermal_solve.cpp:49-163): the same arguments,
result fields and behaviour -- the solve works on the rise above a
reference temperature (the first convective ambient, else the mean fixed
temperature), re-references once to the mean rise if the first pass did not
converge, and reports the heat budget; radiating faces are Newton-linearised
around the last iterate until the field stops moving.

Structure here: ``LinearConduction`` holds the options of one linear solve
and runs it on a prepared device operator (``run``) in three stages --
right-hand side and warm start, the MPIR solve with its restart, the budget;
``radiation_fixed_point`` drives it through the Newton iterates.  Transient
steps reuse both with a capacity term.
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
    """Nodal temperatures and the heat budget of one steady solve."""

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

    return _solve(problem, config=config, runtime=runtime, backend=backend, device_id=device_id,
                  initial_temperature_k=initial_temperature_k, preconditioner=preconditioner,
                  coarse_block_nodes=coarse_block_nodes, reference_temperature_k=reference_temperature_k,
                  native=native, radiation_max_iterations=radiation_max_iterations,
                  radiation_tolerance_k=radiation_tolerance_k)


def combined_result(parts: list[MPIRResult]) -> MPIRResult:
    """One result for a solve that ran in several passes (a re-referenced
    restart): the last pass's answer, the counts and history of all."""

    last = parts[-1]
    return dataclasses.replace(
        last,
        outer_iterations=sum(p.outer_iterations for p in parts),
        inner_iterations=sum(p.inner_iterations for p in parts),
        high_operator_applications=sum(p.high_operator_applications for p in parts),
        low_operator_applications=sum(p.low_operator_applications for p in parts),
        history=tuple(step for p in parts for step in p.history),
    )


@dataclass
class LinearConduction:
    """Options of a linear conduction solve (a steady solve, or a
    backward-Euler step with ``capacity_per_s`` and the previous field)."""

    config: MPIRConfig | None = None
    runtime: object = None
    backend: str | None = None
    device_id: int = 0
    preconditioner: Preconditioner = "two-level"
    coarse_block_nodes: int | None = None
    reference_temperature_k: float | None = None
    native: bool | None = None
    capacity_per_s: np.ndarray | None = None
    previous_temperature_k: np.ndarray | None = None

    def operator(self, problem: ThermalConductionProblem) -> MatrixFreeThermalOperator:
        return MatrixFreeThermalOperator(
            problem, runtime=self.runtime, backend=self.backend, device_id=self.device_id,
            preconditioner=self.preconditioner, coarse_block_nodes=self.coarse_block_nodes, native=self.native,
            capacity_per_s=self.capacity_per_s,
        )

    def previous(self) -> np.ndarray | None:
        if self.previous_temperature_k is None:
            return None
        field = np.asarray(self.previous_temperature_k, dtype=np.float64).reshape(-1)
        return np.nan_to_num(field, nan=0.0, posinf=0.0, neginf=0.0)

    @staticmethod
    def warm_rise(operator: MatrixFreeThermalOperator, initial_temperature_k, reference: float) -> np.ndarray | None:
        """A warm start as a rise above ``reference``: the guess on free rows,
        the fixed values on fixed rows (a void node's row holds the reference,
        so its rise is zero)."""

        if initial_temperature_k is None:
            return None
        guess = np.asarray(initial_temperature_k, dtype=np.float64)
        start = np.broadcast_to(guess, (operator.size,)) if guess.ndim == 0 else guess.reshape(-1)
        if start.size != operator.size:
            raise ValueError("initial_temperature_k must hold one value per node")
        fixed = operator.fixed_nodes.reshape(-1)
        void = ~operator.active_nodes.reshape(-1)
        held = operator.fixed_temperature_k.reshape(-1) + np.where(void, reference, 0.0)
        return np.where(fixed, held, start) - reference

    def mpir(self, operator: MatrixFreeThermalOperator, reference: float, initial) -> tuple[MPIRResult, float]:
        """The MPIR solve of the rise, and once more from the mean free rise
        as the new reference when the first pass stops unconverged; returns
        the result and the reference of its answer."""

        previous = self.previous()
        # A stiff copper/laminate stack buys about one decade per outer step
        # of the FP32 inner solve (the reference's solve.py:223-226): 16 outer
        # steps, with the preconditioner's inner budget.
        config = self.config or default_config(operator.preconditioner, 16)
        passes = [solve_mpir(operator, operator.build_rhs(reference, previous), config=config,
                             initial_guess=initial)]
        free = operator.free_nodes.reshape(-1)
        if not passes[0].converged and free.any():
            shift = float(np.mean(passes[0].solution[free]))
            if shift != 0.0:
                reference += shift
                passes.append(solve_mpir(operator, operator.build_rhs(reference, previous), config=config,
                                         initial_guess=passes[0].solution - shift))
        return (passes[0] if len(passes) == 1 else combined_result(passes)), reference

    def budget(self, problem: ThermalConductionProblem, operator: MatrixFreeThermalOperator, result: MPIRResult,
               reference: float) -> ThermalConductionSolution:
        """Temperatures, element heat flux and the heat budget: input equals
        convected plus fixed-node plus stored heat."""

        mesh = problem.mesh
        previous = self.previous()
        temperature = result.solution.reshape(mesh.node_shape) + reference
        residual = operator.unconstrained_residual(temperature, previous)
        fixed_heat = -float(residual[operator.fixed_nodes.reshape(-1)].sum())
        convective = operator.convective_heat(temperature)
        stored = float(np.sum(operator.stored_heat_w(temperature, previous)))
        heat_input = float(operator.nodal_load().sum())
        flux = operator.element_heat_flux(temperature)
        if not mesh.is_full:
            flux = np.where(np.asarray(mesh.active)[..., None], flux, 0.0)
            temperature = np.where(operator.active_nodes, temperature, np.nan)
        return ThermalConductionSolution(
            temperature_k=temperature,
            heat_flux_w_per_m2=flux,
            max_temperature_k=float(np.nanmax(temperature)),
            min_temperature_k=float(np.nanmin(temperature)),
            total_heat_input_w=heat_input,
            convective_heat_w=convective,
            fixed_temperature_heat_w=fixed_heat,
            heat_balance_error_w=heat_input - float(np.sum(convective)) - fixed_heat - stored,
            solve=result,
            stored_heat_w=stored,
        )

    def run(self, problem: ThermalConductionProblem, initial_temperature_k=None,
            operator: MatrixFreeThermalOperator | None = None) -> ThermalConductionSolution:
        """One linear solve; ``operator`` is a prepared operator of ``problem``
        with this capacity, which a transient march reuses while its step
        repeats."""

        operator = operator or self.operator(problem)
        reference = (operator.default_reference_temperature() if self.reference_temperature_k is None
                     else float(self.reference_temperature_k))
        if not np.isfinite(reference):
            raise ValueError("reference_temperature_k must be finite")
        result, reference = self.mpir(operator, reference, self.warm_rise(operator, initial_temperature_k, reference))
        return self.budget(problem, operator, result, reference)


def radiation_fixed_point(problem: ThermalConductionProblem, linear: LinearConduction, initial_temperature_k,
                          max_iterations: int, tolerance_k: float) -> ThermalConductionSolution:
    """Newton iteration on radiating faces: each iterate solves the problem
    with the radiating faces replaced by their tangent convection faces at
    the last field, until no node moves more than ``tolerance_k``."""

    if max_iterations < 1:
        raise ValueError("radiation_max_iterations must be positive")
    if not tolerance_k > 0.0:
        raise ValueError("radiation_tolerance_k must be positive")
    mesh = problem.mesh
    level = (problem.convection or problem.radiation)[0].mean_ambient_k()
    if initial_temperature_k is None:
        field = np.full(mesh.node_shape, float(level))
    else:
        guess = np.asarray(initial_temperature_k, dtype=np.float64)
        field = np.array(np.broadcast_to(guess, mesh.node_shape) if guess.ndim == 0 else guess.reshape(mesh.node_shape))
        field[~np.isfinite(field)] = problem.radiation[0].mean_ambient_k()
    n_convection = len(problem.convection)
    solution, change, iterations = None, float("inf"), 0
    while iterations < max_iterations and not change <= tolerance_k:
        iterations += 1
        tangents = tuple(face.linearize(mesh, field) for face in problem.radiation)
        linear_problem = dataclasses.replace(problem, convection=tuple(problem.convection) + tangents, radiation=())
        solution = linear.run(linear_problem, field)
        updated = np.where(np.isfinite(solution.temperature_k), solution.temperature_k, field)
        change = float(np.max(np.abs(updated - field)))
        field = updated
    heat = solution.convective_heat_w
    return dataclasses.replace(
        solution,
        convective_heat_w=heat[:n_convection],
        radiative_heat_w=heat[n_convection:],
        radiation_iterations=iterations,
        radiation_converged=bool(change <= tolerance_k and solution.solve.converged),
        radiation_change_k=change,
    )


def _solve(problem: ThermalConductionProblem, *, initial_temperature_k=None, radiation_max_iterations: int = 25,
           radiation_tolerance_k: float = 1.0e-4, **options) -> ThermalConductionSolution:
    """A steady solve, or one backward-Euler step when ``capacity_per_s`` is
    among ``options`` (the fields of ``LinearConduction``)."""

    linear = LinearConduction(**options)
    if not problem.radiation:
        return linear.run(problem, initial_temperature_k)
    return radiation_fixed_point(problem, linear, initial_temperature_k, radiation_max_iterations,
                                 radiation_tolerance_k)
