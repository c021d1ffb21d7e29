"""Mixed-precision iterative refinement: configuration, results and the solve
entry point.

The contract is the reference's (src/electrical/matrix_free_mpir_fem/
solver.py:16-116, 580-698): an FP64 outer loop ``r = b - A x``, stop when
``||r|| <= atol + rtol ||b||``, otherwise an inner Jacobi-PCG correction in
FP32 to ``inner_relative_tolerance`` and ``x += correction``.  Here the whole
loop runs on the HIP device (``pcbhip_solve_mpir`` / ``pcbhip_batch_run``):
a system hands its solve to the device through ``native_solve_mpir``, and a
system without that hook is refused -- there is no host fallback.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Literal, Protocol

import numpy as np


class MatrixFreeMPIRSystem(Protocol):
    """Split-precision operator contract (solver.py:16-38); the HIP systems
    additionally provide ``native_solve_mpir``."""

    size: int
    runtime: Any
    high_dtype: Any
    inner_solver: Literal["pcg", "gmres"]

    def apply_high(self, vector: np.ndarray) -> np.ndarray: ...

    def apply_low(self, vector: Any) -> Any: ...

    def diagonal_low(self) -> Any: ...


@dataclass(frozen=True)
class MPIRConfig:
    """Accuracy and work limits for the nested solve (solver.py:52-79)."""

    relative_tolerance: float = 1.0e-10
    absolute_tolerance: float = 0.0
    inner_relative_tolerance: float = 2.0e-3
    max_outer_iterations: int = 8
    max_inner_iterations: int = 200
    gmres_restart: int = 24

    def __post_init__(self) -> None:
        for name in ("relative_tolerance", "absolute_tolerance", "inner_relative_tolerance"):
            value = float(getattr(self, name))
            if not np.isfinite(value) or value < 0.0:
                raise ValueError(f"{name} must be finite and non-negative")
        if self.relative_tolerance == 0.0 and self.absolute_tolerance == 0.0:
            raise ValueError("at least one outer tolerance must be positive")
        if not 0.0 < self.inner_relative_tolerance < 1.0:
            raise ValueError("inner_relative_tolerance must be between zero and one")
        if self.max_outer_iterations < 1 or self.max_inner_iterations < 1:
            raise ValueError("iteration limits must be positive")
        if self.gmres_restart < 2:
            raise ValueError("gmres_restart must be at least two")

    def to_c(self):
        from pcbhip import _capi

        return _capi.config_c(self.relative_tolerance, self.absolute_tolerance, self.inner_relative_tolerance,
                              self.max_outer_iterations, self.max_inner_iterations)


# The reference's default limits (8 outer, 200 inner iterations) are sized for
# its default two-level preconditioner (two_level.py:54-229).  Jacobi needs
# longer inner solves to reach the same 2e-3 inner tolerance, so the front
# ends give a Jacobi solve this inner budget when the caller passes no
# configuration.
JACOBI_MAX_INNER_ITERATIONS = 20000


def jacobi_default_config(max_outer_iterations: int = 8) -> MPIRConfig:
    """The default limits of a front end for the Jacobi-preconditioned device solve."""

    return MPIRConfig(max_outer_iterations=max_outer_iterations, max_inner_iterations=JACOBI_MAX_INNER_ITERATIONS)


def default_config(preconditioner: str, max_outer_iterations: int = 8) -> MPIRConfig:
    """A front end's limits when the caller passes none: the reference's own
    with the two-level preconditioner, the Jacobi budget with Jacobi."""

    if preconditioner == "two-level":
        return MPIRConfig(max_outer_iterations=max_outer_iterations)
    return jacobi_default_config(max_outer_iterations)


@dataclass(frozen=True)
class MPIRStep:
    outer_iteration: int
    high_relative_residual: float
    inner_iterations: int
    inner_relative_residual: float


@dataclass(frozen=True)
class MPIRResult:
    solution: np.ndarray
    converged: bool
    outer_iterations: int
    inner_iterations: int
    relative_residual: float
    high_operator_applications: int
    low_operator_applications: int
    low_runtime: str
    history: tuple[MPIRStep, ...]


def mpir_result_from_core(result: dict, low_runtime: str) -> MPIRResult:
    """An :class:`MPIRResult` from a core result dictionary (solver.py:103-116)."""

    return MPIRResult(
        solution=result["solution"],
        converged=bool(result["converged"]),
        outer_iterations=int(result["outer_iterations"]),
        inner_iterations=int(result["inner_iterations"]),
        relative_residual=float(result["relative_residual"]),
        high_operator_applications=int(result["high_operator_applications"]),
        low_operator_applications=int(result["low_operator_applications"]),
        low_runtime=low_runtime,
        history=tuple(MPIRStep(*step) for step in result["history"]),
    )


def solve_mpir(
    system: MatrixFreeMPIRSystem,
    rhs: np.ndarray,
    *,
    config: MPIRConfig | None = None,
    initial_guess: np.ndarray | None = None,
) -> MPIRResult:
    """Solve ``A x = rhs`` with FP64 refinement around an FP32 inner PCG, on the device."""

    config = config or MPIRConfig()
    high_dtype = getattr(system, "high_dtype", np.float64)
    rhs_high = np.asarray(rhs, dtype=high_dtype).reshape(-1)
    if rhs_high.size != system.size:
        raise ValueError(f"rhs has size {rhs_high.size}, expected {system.size}")
    if not np.all(np.isfinite(rhs_high)):
        raise ValueError("rhs must contain only finite values")
    if initial_guess is not None:
        guess = np.asarray(initial_guess, dtype=high_dtype).reshape(-1)
        if guess.size != system.size:
            raise ValueError(f"initial_guess has size {guess.size}, expected {system.size}")
    if getattr(system, "inner_solver", "pcg") == "gmres":
        # The reference dispatches complex (Maxwell) systems to restarted
        # GMRES (solver.py:665-675); those operators are not on this path.
        raise NotImplementedError("the restarted GMRES inner solve of complex systems is not on the HIP path")
    native = getattr(system, "native_solve_mpir", None)
    if native is None:
        raise TypeError("solve_mpir runs on the HIP device and needs a system with native_solve_mpir")
    out = native(rhs_high, config, initial_guess)
    if out is None:
        raise RuntimeError("the system declined the device solve")
    sol, conv, outer, inner, rel, n_high, n_low, steps = out
    return MPIRResult(
        solution=np.asarray(sol, dtype=high_dtype),
        converged=bool(conv),
        outer_iterations=int(outer),
        inner_iterations=int(inner),
        relative_residual=float(rel),
        high_operator_applications=int(n_high),
        low_operator_applications=int(n_low),
        low_runtime=system.runtime.name,
        history=tuple(MPIRStep(*step) for step in steps),
    )


def solve_mpir_device(
    device_system: Any,
    rhs: np.ndarray,
    *,
    config: MPIRConfig | None = None,
    initial_guess: np.ndarray | None = None,
) -> MPIRResult:
    """The device-system contract of the reference (solver.py:119-148):
    ``device_system.solve(rhs, x0, rtol, atol, inner_rtol, max_outer,
    max_inner, threads) -> dict``."""

    config = config or MPIRConfig()
    rhs_high = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
    if not np.all(np.isfinite(rhs_high)):
        raise ValueError("rhs must contain only finite values")
    result = device_system.solve(
        rhs_high,
        None if initial_guess is None else np.ascontiguousarray(initial_guess, dtype=np.float64).reshape(-1),
        float(config.relative_tolerance),
        float(config.absolute_tolerance),
        float(config.inner_relative_tolerance),
        int(config.max_outer_iterations),
        int(config.max_inner_iterations),
        1,
    )
    return mpir_result_from_core(result, "hip-fp32")
