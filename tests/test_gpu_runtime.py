"""The FP32 inner-solve runtime on the device (the reference's
LowPrecisionRuntime contract, src/electrical/matrix_free_mpir_fem/runtime.py:
20-44, 121-170): element-wise operations bit for bit against NumPy's float32
arithmetic, the dot product against an exactly rounded sum and stable from
call to call, operator actions on device vectors equal to the host path, and
the reference's host-driven inner PCG (solver.py:180-230) run through it."""

from __future__ import annotations

import math

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import HipFloat32Runtime
from pcbhip import _capi
from tests.helpers import stack
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def runtime():
    _capi.require_device(0)
    return HipFloat32Runtime(device_id=0)


def _vectors(n, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32))


@pytest.mark.parametrize("n", [1, 7, 4096, 1_000_003])
def test_elementwise_operations_are_numpy_float32_bit_for_bit(runtime, n):
    print(f"runtime elementwise n={n}", flush=True)
    x, y = _vectors(n)
    dx, dy = runtime.from_host(x), runtime.from_host(y)
    assert dx.to_host().tobytes() == x.tobytes()
    for alpha in (0.37, -2.5, 1e-3):
        expected = y + np.float32(alpha) * x
        assert runtime.axpy(alpha, dx, dy).to_host().tobytes() == expected.tobytes(), alpha
    den = np.where(np.abs(y) < 1e-3, np.float32(1.5), y)
    assert runtime.divide(dx, runtime.from_host(den)).to_host().tobytes() == (x / den).tobytes()
    assert runtime.divide(dx, den).to_host().tobytes() == (x / den).tobytes()  # host denominator
    assert runtime.copy(dx).to_host().tobytes() == x.tobytes()
    assert not runtime.zeros_like(dx).to_host().any()
    assert runtime.to_host(dx).dtype == np.float64


@pytest.mark.parametrize("n", [1, 4095, 4097, 3_000_001])
def test_dot_is_accurate_and_the_same_on_every_call(runtime, n):
    print(f"runtime dot n={n}", flush=True)
    x, y = _vectors(n, seed=n)
    dx, dy = runtime.from_host(x), runtime.from_host(y)
    products = x.astype(np.float64) * y.astype(np.float64)  # exact
    exact = math.fsum(products)
    got = runtime.dot(dx, dy)
    assert abs(got - exact) <= 1e-13 * float(np.abs(products).sum())
    assert all(runtime.dot(dx, dy) == got for _ in range(3))
    assert runtime.norm(dx) == pytest.approx(math.sqrt(math.fsum(x.astype(np.float64) ** 2)), rel=1e-14)
    assert runtime.dot(runtime.zeros_like(dx), dy) == 0.0


def test_mismatched_sizes_are_refused(runtime):
    with pytest.raises(ValueError):
        runtime.axpy(1.0, runtime.from_host(np.ones(5)), runtime.from_host(np.ones(6)))


def test_operator_actions_on_device_vectors_equal_the_host_path():
    print("runtime operator actions", flush=True)
    for preconditioner in ("jacobi", "two-level"):
        operator = MatrixFreeThermalOperator(stack(20, 30, fixed=True), preconditioner=preconditioner)
        rt = operator.runtime
        x = np.random.default_rng(3).standard_normal(operator.size).astype(np.float32)
        dx = rt.from_host(x)
        assert operator.apply_low(dx).to_host().tobytes() == operator.apply_low(x).tobytes()
        assert operator.precondition_low(dx).to_host().tobytes() == operator.precondition_low(x).tobytes()


def _pcg_low(operator, runtime, rhs_high, tol, max_iterations):
    """The reference's host-driven FP32 PCG (solver.py:180-230): every vector
    lives in the runtime, every product is a runtime operation."""

    rhs = runtime.from_host(rhs_high)
    correction = runtime.zeros_like(rhs)
    residual = runtime.copy(rhs)
    rhs_norm = runtime.norm(rhs)
    preconditioned = operator.precondition_low(residual)
    direction = runtime.copy(preconditioned)
    rz = runtime.dot(residual, preconditioned)
    for iteration in range(1, max_iterations + 1):
        applied = operator.apply_low(direction)
        curvature = runtime.dot(direction, applied)
        alpha = rz / curvature
        correction = runtime.axpy(alpha, direction, correction)
        residual = runtime.axpy(-alpha, applied, residual)
        if runtime.norm(residual) / rhs_norm <= tol:
            return runtime.to_host(correction), iteration
        preconditioned = operator.precondition_low(residual)
        next_rz = runtime.dot(residual, preconditioned)
        direction = runtime.axpy(next_rz / rz, direction, preconditioned)
        rz = next_rz
    return runtime.to_host(correction), max_iterations


@pytest.mark.parametrize("preconditioner", ["jacobi", "two-level"])
def test_host_driven_pcg_through_the_runtime_solves_to_its_tolerance(preconditioner):
    print(f"runtime host-driven PCG {preconditioner}", flush=True)
    operator = MatrixFreeThermalOperator(stack(24, 40, fixed=True), preconditioner=preconditioner)
    rhs = operator.build_rhs(operator.default_reference_temperature())
    x, iterations = _pcg_low(operator, operator.runtime, rhs, 1e-5, 5000)
    assert iterations < 5000
    # FP32 PCG: the FP64 residual of its answer sits at the tolerance, up to
    # the FP32 drift of the recursive residual.
    residual = rhs - operator.apply_high(x)
    assert np.linalg.norm(residual) <= 5e-5 * np.linalg.norm(rhs)
