"""Host logic of the refinement loop and its two inner solves
(src/electrical/matrix_free_mpir_fem/solver.py:150-397, 636-698), run on the
CPU against a NumPy stand-in for the device FP32 runtime.

The stand-in below is test infrastructure: it gives the same operations as
the HIP runtime (FP32 vectors, FP64 dot products, the Gram-Schmidt
``project`` / ``combine``), so the solver's bookkeeping -- Arnoldi columns,
Hessenberg least squares, restarts, scaling, the outer loop -- is checked
against the mathematics here, and the device runs the same code in
``tests/test_gpu_solver.py``.  ``solve_mpir`` itself refuses this stand-in:
the product has no host runtime."""

from __future__ import annotations

import math

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import solver
from electrical.matrix_free_mpir_fem.solver import MPIRConfig


class NumpyStandIn:
    name = "numpy-stand-in"
    is_hip = False

    def from_host(self, v):
        return np.asarray(v, dtype=np.float32).reshape(-1).copy()

    def to_host(self, v):
        return np.asarray(v, dtype=np.float64)

    def zeros_like(self, v):
        return np.zeros(np.size(v), dtype=np.float32)

    def copy(self, v):
        return np.array(v, dtype=np.float32)

    def dot(self, a, b):
        return float(np.dot(np.asarray(a, np.float64), np.asarray(b, np.float64)))

    def norm(self, v):
        return math.sqrt(self.dot(v, v))

    def axpy(self, alpha, x, y):
        return (y + np.float32(alpha) * x).astype(np.float32)

    def divide(self, a, b):
        return (np.asarray(a, np.float32) / np.asarray(b, np.float32)).astype(np.float32)

    def project(self, rows, w):
        return np.array([self.dot(r, w) for r in rows])

    def combine(self, coefficients, rows, base=None):
        acc = np.zeros(np.size(rows[0])) if base is None else np.asarray(base, np.float64).copy()
        for c, r in zip(np.asarray(coefficients, np.float64), rows):
            acc += c * np.asarray(r, np.float64)
        return acc.astype(np.float32)


class DenseSystem:
    """A dense matrix, FP64 high and FP32 low, Jacobi preconditioned."""

    high_dtype = np.float64

    def __init__(self, matrix, inner_solver="gmres"):
        self.matrix = np.asarray(matrix, dtype=np.float64)
        self.size = self.matrix.shape[0]
        self.low = self.matrix.astype(np.float32)
        self.runtime = NumpyStandIn()
        self.inner_solver = inner_solver
        self.low_applications = 0

    def apply_high(self, v):
        return self.matrix @ np.asarray(v, np.float64)

    def apply_low(self, v):
        self.low_applications += 1
        return (self.low @ np.asarray(v, np.float32)).astype(np.float32)

    def diagonal_low(self):
        return np.diag(self.low).copy()


def nonsymmetric(n, seed=0, skew=0.6):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) / math.sqrt(n)
    return np.diag(rng.uniform(2.0, 6.0, n)) + a + skew * (np.triu(a, 1) - np.triu(a, 1).T)


def spd(n, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, n))
    return c @ c.T + n * np.eye(n)


def test_gmres_iterate_minimises_the_residual_over_the_krylov_space():
    """One cycle of k columns gives x = M^-1 V y with y the least-squares
    minimiser of ||b - A M^-1 V y|| over the k-dimensional Krylov space of
    A M^-1 and b: checked against that minimiser built directly (an
    orthonormal Krylov basis by QR in FP64).  A wrong sign or index in the
    Gram-Schmidt or the Hessenberg fails this."""

    n, k = 50, 6
    a = nonsymmetric(n, seed=1)
    system = DenseSystem(a)
    b = np.random.default_rng(2).standard_normal(n)
    config = MPIRConfig(inner_relative_tolerance=1e-12, max_inner_iterations=k, gmres_restart=k)
    x, iterations, relative, applications = solver._inner_gmres(system, b, config)
    assert iterations == k and applications == k
    m_inv = 1.0 / np.diag(a)
    op = a * m_inv[None, :]  # A M^-1
    krylov = np.empty((n, k))
    v = b.copy()
    for j in range(k):
        krylov[:, j] = v
        v = op @ v
    q, _ = np.linalg.qr(krylov)
    y, *_ = np.linalg.lstsq(op @ q, b, rcond=None)
    expected = m_inv * (q @ y)
    np.testing.assert_allclose(x, expected, rtol=0, atol=2e-5 * np.abs(expected).max())
    true_relative = np.linalg.norm(b - a @ expected) / np.linalg.norm(b)
    assert relative == pytest.approx(true_relative, rel=1e-3)


@pytest.mark.parametrize("restart", [2, 5, 24])
def test_restarted_gmres_solves_a_nonsymmetric_system(restart):
    n = 60
    a = nonsymmetric(n, seed=restart)
    system = DenseSystem(a)
    b = np.random.default_rng(7).standard_normal(n)
    config = MPIRConfig(inner_relative_tolerance=1e-5, max_inner_iterations=400, gmres_restart=restart)
    x, iterations, relative, applications = solver._inner_gmres(system, b, config)
    assert relative <= 1e-5
    assert np.linalg.norm(b - a @ x) <= 3e-5 * np.linalg.norm(b)
    # One operator application per column and one per restart's residual.
    cycles = math.ceil(iterations / restart)
    assert iterations + cycles - 1 <= applications <= iterations + cycles
    assert system.low_applications == applications


def test_gmres_stops_at_its_iteration_limit_and_on_a_zero_right_hand_side():
    a = nonsymmetric(40, seed=3, skew=2.0)
    system = DenseSystem(a)
    b = np.random.default_rng(4).standard_normal(40)
    config = MPIRConfig(inner_relative_tolerance=1e-9, max_inner_iterations=7, gmres_restart=3)
    _, iterations, relative, _ = solver._inner_gmres(system, b, config)
    assert iterations == 7 and relative > 1e-9
    x, iterations, relative, applications = solver._inner_gmres(system, np.zeros(40), config)
    assert not x.any() and (iterations, relative, applications) == (0, 0.0, 0)


def test_gmres_breakdown_on_an_invariant_subspace_is_exact():
    """With M^-1 b an eigenvector of A M^-1 (A diagonal, b one unit vector)
    the first column's candidate is zero: the cycle stops at once with the
    exact answer."""

    a = np.diag(np.arange(1.0, 11.0))
    system = DenseSystem(a)
    b = np.zeros(10)
    b[3] = 2.0
    x, iterations, relative, _ = solver._inner_gmres(system, b, MPIRConfig(gmres_restart=4))
    assert iterations == 1 and relative <= 1e-7
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-6)


def test_inner_pcg_matches_the_solution_and_scales_tiny_corrections_exactly():
    a = spd(12, seed=0)
    system = DenseSystem(a, inner_solver="pcg")
    rng = np.random.default_rng(0)
    unit = rng.standard_normal(12)
    config = MPIRConfig(inner_relative_tolerance=1e-4)
    big, iterations, relative, applications = solver._inner_pcg(system, unit, config)
    assert relative <= 1e-4 and iterations == applications > 0
    exact = np.linalg.solve(a, unit)
    np.testing.assert_allclose(big, exact, rtol=0, atol=1e-3 * np.abs(exact).max())
    small, *_ = solver._inner_pcg(system, unit * 2.0**-70, config)
    np.testing.assert_array_equal(small * 2.0**70, big)


def test_inner_pcg_refuses_an_indefinite_operator():
    a = spd(10, seed=1)
    a[0, 0] = -50.0
    system = DenseSystem(a, inner_solver="pcg")
    b = np.zeros(10)
    b[0] = 1.0
    with pytest.raises(RuntimeError, match="positive-definite"):
        solver._inner_pcg(system, b, MPIRConfig())


@pytest.mark.parametrize("inner", ["gmres", "pcg"])
def test_refinement_reaches_fp64_accuracy_through_fp32_inner_solves(inner):
    a = nonsymmetric(80, seed=5) if inner == "gmres" else spd(80, seed=5)
    system = DenseSystem(a, inner_solver=inner)
    b = np.random.default_rng(6).standard_normal(80)
    config = MPIRConfig(max_inner_iterations=300, gmres_restart=16)
    result = solver._refine(system, b, config, None)
    assert result.converged and result.relative_residual <= 1e-10
    assert np.linalg.norm(b - a @ result.solution) <= 1e-10 * np.linalg.norm(b)
    assert result.outer_iterations == len(result.history) >= 2
    assert result.high_operator_applications == result.outer_iterations + 1
    assert result.inner_iterations == sum(step.inner_iterations for step in result.history)
    assert result.low_operator_applications == system.low_applications
    assert result.history[0].high_relative_residual == pytest.approx(1.0)
    # A guess that already solves the system costs one FP64 residual.
    again = solver._refine(system, b, config, result.solution.copy())
    assert again.converged and again.outer_iterations == 0 and again.inner_iterations == 0


def test_refinement_reports_non_convergence_at_its_outer_limit():
    a = nonsymmetric(30, seed=8)
    system = DenseSystem(a)
    b = np.ones(30)
    config = MPIRConfig(max_outer_iterations=1, inner_relative_tolerance=0.5, relative_tolerance=1e-14)
    result = solver._refine(system, b, config, None)
    assert not result.converged and result.outer_iterations == 1 and len(result.history) == 1


def test_solve_mpir_refuses_systems_off_the_device_and_unknown_inner_solvers():
    system = DenseSystem(spd(6))
    with pytest.raises(TypeError, match="HIP device"):
        solver.solve_mpir(system, np.ones(6))
    system.inner_solver = "bicgstab"
    with pytest.raises(ValueError, match="unsupported inner solver"):
        solver._refine(system, np.ones(6), MPIRConfig(), None)
