"""Mixed-precision iterative refinement: configuration, results and the solve
entry point.

The contract is the reference's (src/electrical/matrix_free_mpir_fem/
solver.py:16-116, 580-698): an FP64 outer loop ``r = b - A x``, stop when
``||r|| <= atol + rtol ||b||``, otherwise an FP32 inner correction to
``inner_relative_tolerance`` and ``x += correction``.

Two ways through it, in the reference's order:
- a system with ``native_solve_mpir`` (every operator of this package) runs
  the whole loop on the HIP device (``pcbhip_solve_mpir`` /
  ``pcbhip_batch_run``), Jacobi or two-level PCG fused into device kernels;
- any other system is refined here, its inner correction solved by PCG or by
  restarted right-preconditioned GMRES (``inner_solver``) on the FP32 vector
  runtime of the HIP device: the system's ``apply_low`` and preconditioner on
  device vectors, and the runtime's device dot products, updates and
  Gram-Schmidt products.  Only the outer FP64 vectors and GMRES's small
  Hessenberg least-squares problem live on the host, as in the reference.
  A system whose runtime is not the HIP one is refused: there is no host
  fallback.
"""

from __future__ import annotations

import math
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
    """Solve ``A x = rhs`` with FP64 refinement around an FP32 inner solve, on the device."""

    config = config or MPIRConfig()
    high_dtype = getattr(system, "high_dtype", np.float64)
    if np.dtype(high_dtype).kind == "c":
        # The reference's complex systems are its frequency-domain Maxwell
        # operators (frequency_domain.py:178); this path is real-valued.
        raise NotImplementedError("complex systems are not on the HIP path; GMRES here takes real operators")
    rhs_high = np.asarray(rhs, dtype=high_dtype).reshape(-1)
    if rhs_high.size != system.size:
        raise ValueError(f"rhs has size {rhs_high.size}, expected {system.size}")
    if not np.all(np.isfinite(rhs_high)):
        raise ValueError("rhs must contain only finite values")
    guess = None
    if initial_guess is not None:
        guess = np.asarray(initial_guess, dtype=high_dtype).reshape(-1)
        if guess.size != system.size:
            raise ValueError(f"initial_guess has size {guess.size}, expected {system.size}")
    native = getattr(system, "native_solve_mpir", None)
    if native is not None:
        out = native(rhs_high, config, initial_guess)
        if out is not None:
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
    if not getattr(getattr(system, "runtime", None), "is_hip", False):
        raise TypeError(
            "solve_mpir runs on the HIP device: the system needs native_solve_mpir or the HIP FP32 runtime"
        )
    return _refine(system, rhs_high, config, guess)


def _low_preconditioner(system: MatrixFreeMPIRSystem) -> Any:
    """The FP32 preconditioner on runtime vectors (solver.py:41-49): the
    system's ``precondition_low``, else division by ``diagonal_low()``
    (uploaded once)."""

    own = getattr(system, "precondition_low", None)
    if own is not None:
        return own
    runtime = system.runtime
    diagonal = runtime.from_host(np.asarray(system.diagonal_low(), dtype=np.float32).reshape(-1))
    return lambda vector: runtime.divide(vector, diagonal)


def _refine(
    system: MatrixFreeMPIRSystem, rhs_high: np.ndarray, config: MPIRConfig, guess: np.ndarray | None
) -> MPIRResult:
    """The outer FP64 loop of solver.py:636-698 for a system without a device
    solve of its own: the FP64 residual from ``apply_high``, the correction
    from the inner solve the system names."""

    inner_solver = getattr(system, "inner_solver", "pcg")
    if inner_solver == "pcg":
        inner = _inner_pcg
    elif inner_solver == "gmres":
        inner = _inner_gmres
    else:
        raise ValueError(f"unsupported inner solver {inner_solver!r}")
    x = np.zeros_like(rhs_high) if guess is None else guess.copy()
    rhs_norm = float(np.linalg.norm(rhs_high))
    scale = rhs_norm if rhs_norm > 0.0 else 1.0
    target = config.absolute_tolerance + config.relative_tolerance * scale
    steps: list[MPIRStep] = []
    n_high = n_low = inner_total = 0
    relative = math.inf
    for outer in range(config.max_outer_iterations + 1):
        residual = rhs_high - np.asarray(system.apply_high(x), dtype=rhs_high.dtype).reshape(-1)
        n_high += 1
        residual_norm = float(np.linalg.norm(residual))
        relative = residual_norm / scale
        if residual_norm <= target or outer == config.max_outer_iterations:
            return MPIRResult(
                solution=x,
                converged=residual_norm <= target,
                outer_iterations=outer,
                inner_iterations=inner_total,
                relative_residual=relative,
                high_operator_applications=n_high,
                low_operator_applications=n_low,
                low_runtime=system.runtime.name,
                history=tuple(steps),
            )
        correction, count, inner_relative, applications = inner(system, residual, config)
        x += np.asarray(correction, dtype=x.dtype).reshape(-1)
        inner_total += count
        n_low += applications
        steps.append(MPIRStep(outer + 1, relative, count, inner_relative))
    raise AssertionError("unreachable")


def _inner_pcg(system: MatrixFreeMPIRSystem, rhs_high: np.ndarray, config: MPIRConfig):
    """One FP32 Jacobi-PCG correction on the runtime (solver.py:150-232).

    The right-hand side is scaled by a power of two to a peak in [0.5, 1), so
    that the FP32 products of a late, tiny outer residual do not underflow;
    the scaling is exact and undone on the way out.  A non-positive or
    non-finite curvature stops the solve with the reference's error.
    Returns (correction, iterations, relative residual, applications)."""

    runtime = system.runtime
    peak = float(np.max(np.abs(rhs_high))) if rhs_high.size else 0.0
    shift = 0 if peak == 0.0 or not np.isfinite(peak) else -math.frexp(peak)[1]
    scale = math.ldexp(1.0, shift)
    b = runtime.from_host(np.ldexp(rhs_high, shift) if shift else rhs_high)
    b_norm = runtime.norm(b)
    if b_norm == 0.0:
        return np.zeros_like(rhs_high), 0, 0.0, 0
    precondition = _low_preconditioner(system)
    x = runtime.zeros_like(b)
    r = runtime.copy(b)
    z = precondition(r)
    p = runtime.copy(z)
    rz = runtime.dot(r, z)
    relative = 1.0
    applications = 0
    iteration = 0
    while iteration < config.max_inner_iterations:
        iteration += 1
        q = system.apply_low(p)
        applications += 1
        curvature = runtime.dot(p, q)
        if not math.isfinite(curvature) or curvature <= 0.0:
            raise RuntimeError("inner PCG requires a finite symmetric positive-definite operator")
        alpha = rz / curvature
        x = runtime.axpy(alpha, p, x)
        r = runtime.axpy(-alpha, q, r)
        relative = runtime.norm(r) / b_norm
        if relative <= config.inner_relative_tolerance:
            break
        z = precondition(r)
        rz_next = runtime.dot(r, z)
        if not math.isfinite(rz_next) or rz == 0.0:
            break
        p = runtime.axpy(rz_next / rz, p, z)
        rz = rz_next
    return runtime.to_host(x) / scale, iteration, relative, applications


def _inner_gmres(system: MatrixFreeMPIRSystem, rhs_high: np.ndarray, config: MPIRConfig):
    """One FP32 correction by restarted right-preconditioned GMRES
    (solver.py:235-397 on the host, 400-577 on the device), for real
    operators on the device runtime.

    Each restart cycle of at most ``gmres_restart`` columns keeps its basis
    V and the preconditioned directions Z = M^-1 V as device vectors.  A
    column is orthogonalised by two passes of classical Gram-Schmidt, each
    one projection of the candidate on all of V and one combination
    (``runtime.project`` / ``runtime.combine``), as the reference's device
    path does.  The Hessenberg least-squares problem (at most restart + 1 by
    restart) is solved on the host after every column, and its residual
    estimate decides convergence.  A cycle stops early when the candidate's
    norm falls to FP32 epsilon of the cycle's starting residual (a lucky
    breakdown).  At a restart the FP32 residual is formed again from the
    operator, and the correction is the sum of Z weighted by the
    least-squares solution, accumulated in FP64 on the device.
    Returns (correction, iterations, relative residual, applications)."""

    runtime = system.runtime
    b = runtime.from_host(rhs_high)
    b_norm = runtime.norm(b)
    if b_norm == 0.0:
        return np.zeros_like(rhs_high), 0, 0.0, 0
    precondition = _low_preconditioner(system)
    tolerance = config.inner_relative_tolerance
    tiny = float(np.finfo(np.float32).eps)
    x = runtime.zeros_like(b)
    r = b
    done = 0
    applications = 0
    relative = 1.0
    while done < config.max_inner_iterations:
        if done:
            r = runtime.axpy(-1.0, system.apply_low(x), b)
            applications += 1
        beta = runtime.norm(r)
        relative = beta / b_norm
        if relative <= tolerance:
            break
        width = min(config.gmres_restart, config.max_inner_iterations - done)
        basis = [runtime.combine([1.0 / beta], [r])]
        directions = []
        hessenberg = np.zeros((width + 1, width))
        target = np.zeros(width + 1)
        target[0] = beta
        weights = np.zeros(0)
        for j in range(width):
            z = precondition(basis[j])
            directions.append(z)
            w = system.apply_low(z)
            applications += 1
            h = runtime.project(basis, w)
            w = runtime.combine(-h, basis, w)
            h2 = runtime.project(basis, w)
            w = runtime.combine(-h2, basis, w)
            hessenberg[: j + 1, j] = h + h2
            w_norm = runtime.norm(w)
            hessenberg[j + 1, j] = w_norm
            done += 1
            weights, *_ = np.linalg.lstsq(hessenberg[: j + 2, : j + 1], target[: j + 2], rcond=None)
            relative = float(np.linalg.norm(target[: j + 2] - hessenberg[: j + 2, : j + 1] @ weights)) / b_norm
            breakdown = w_norm <= tiny * beta
            if relative <= tolerance or breakdown:
                break
            if j + 1 < width:
                basis.append(runtime.combine([1.0 / w_norm], [w]))
        x = runtime.combine(weights, directions, x)
        if relative <= tolerance:
            break
    return runtime.to_host(x), done, relative, applications


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
