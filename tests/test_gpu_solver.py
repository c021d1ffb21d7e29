"""The refinement loop with host-driven inner solves on the device runtime
(src/electrical/matrix_free_mpir_fem/solver.py:150-577, 636-698): the
Gram-Schmidt products bit for bit against NumPy's FP64 arithmetic, and
restarted GMRES and PCG corrections through ``solve_mpir`` for a system
that has no device solve of its own, against the fused device solve of the
same operator."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import HipFloat32Runtime, MPIRConfig, solve_mpir
from pcbhip import _capi
from tests.helpers import stack
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def runtime():
    _capi.require_device(0)
    return HipFloat32Runtime(device_id=0)


@pytest.mark.parametrize("n,count", [(1, 1), (4097, 3), (1_000_003, 64), (20_000, 70)])
def test_basis_products_are_numpy_fp64_bit_for_bit(runtime, n, count):
    print(f"basis products n={n} rows={count}", flush=True)
    rng = np.random.default_rng(n + count)
    rows = rng.standard_normal((count, n)).astype(np.float32)
    w = rng.standard_normal(n).astype(np.float32)
    coef = rng.standard_normal(count)
    d_rows = [runtime.from_host(r) for r in rows]
    d_w = runtime.from_host(w)
    # Each projection has the bits of the dot product of the same pair.
    got = runtime.project(d_rows, d_w)
    assert got.tolist() == [runtime.dot(r, d_w) for r in d_rows]
    # The combination: FP64 sum in row order, rounded once per 64 rows.
    acc = w.astype(np.float64)
    for lo in range(0, count, 64):
        for k in range(lo, min(lo + 64, count)):
            acc = acc + coef[k] * rows[k].astype(np.float64)
        acc = acc.astype(np.float32).astype(np.float64)
    assert runtime.combine(coef, d_rows, d_w).to_host().tobytes() == acc.astype(np.float32).tobytes()
    scaled = runtime.combine([0.5], d_rows[:1])  # no base: zero
    assert scaled.to_host().tobytes() == (rows[0].astype(np.float64) * 0.5).astype(np.float32).tobytes()


def test_basis_products_refuse_bad_arguments(runtime):
    v = runtime.from_host(np.ones(8, np.float32))
    with pytest.raises(ValueError):
        runtime.project([runtime.from_host(np.ones(9, np.float32))], v)
    with pytest.raises(ValueError):
        runtime.combine([1.0, 2.0], [v])
    with pytest.raises(ValueError, match="finite"):
        runtime.combine([float("nan")], [v])


class HostDrivenSystem:
    """An operator's actions with no device solve of its own: solve_mpir
    refines it here, the inner solve driving the device runtime."""

    high_dtype = np.float64

    def __init__(self, operator, inner_solver, own_preconditioner=True):
        self.operator = operator
        self.size = operator.size
        self.runtime = operator.runtime
        self.inner_solver = inner_solver
        if own_preconditioner:
            self.precondition_low = operator.precondition_low

    def apply_high(self, v):
        return self.operator.apply_high(v)

    def apply_low(self, v):
        return self.operator.apply_low(v)

    def diagonal_low(self):
        return self.operator.diagonal_low()


@pytest.mark.parametrize(
    "inner,preconditioner,own",
    [("gmres", "two-level", True), ("gmres", "jacobi", False), ("pcg", "two-level", True), ("pcg", "jacobi", False)],
)
def test_host_driven_refinement_matches_the_fused_device_solve(inner, preconditioner, own):
    print(f"host-driven refinement {inner} {preconditioner} own={own}", flush=True)
    operator = MatrixFreeThermalOperator(stack(20, 28, fixed=True), preconditioner=preconditioner)
    rhs = operator.build_rhs(operator.default_reference_temperature())
    config = MPIRConfig(max_inner_iterations=4000, gmres_restart=32)
    fused = solve_mpir(operator, rhs, config=config)
    system = HostDrivenSystem(operator, inner, own)
    result = solve_mpir(system, rhs, config=config)
    assert fused.converged and result.converged
    assert result.relative_residual <= 1e-10
    assert np.linalg.norm(rhs - operator.apply_high(result.solution)) <= 1e-10 * np.linalg.norm(rhs)
    # Both answers hold the FP64 residual below 1e-10; restarted GMRES with
    # Jacobi stalls on this stiff copper/FR-4 stack and reaches it only in
    # the last outer steps, so the answers differ by what the operator's
    # conditioning allows at that residual (about 1e-7 of the peak).
    scale = np.abs(fused.solution).max()
    assert np.abs(result.solution - fused.solution).max() <= 1e-6 * scale
    assert result.low_runtime == operator.runtime.name
    assert result.inner_iterations == sum(step.inner_iterations for step in result.history)


def test_gmres_refinement_of_a_dc_bar_gives_the_closed_form_resistance():
    """The hexahedral DC operator refined through host-driven GMRES: a
    uniform copper bar one element across, driven by +-1 A end faces,
    drops exactly I L / (sigma w t) (the reference's bar check,
    tests/test_matrix_free_mpir_fem.py:84-101)."""

    from electrical.matrix_free_mpir_fem.hex_dc import CurrentTerminal, MatrixFreeHexDCOperator
    from thermal.matrix_free_mpir_fem import LayeredThermalMesh

    print("host-driven GMRES dc bar", flush=True)
    sigma, cells = 5.8e7, 16
    mesh = LayeredThermalMesh((35e-6,), 1.0e-3, 1.0e-3, sigma, element_shape=(1, cells))
    face = lambda x: tuple((z, y, x) for z in range(mesh.node_shape[0]) for y in range(mesh.node_shape[1]))
    operator = MatrixFreeHexDCOperator(mesh, reference_node=(0, 0, cells // 2))
    rhs = operator.build_rhs((CurrentTerminal(face(0), 1.0), CurrentTerminal(face(cells), -1.0)))
    result = solve_mpir(HostDrivenSystem(operator, "gmres"), rhs,
                        config=MPIRConfig(max_inner_iterations=2000, gmres_restart=8))
    assert result.converged and result.relative_residual <= 1e-10
    v = result.solution.reshape(mesh.node_shape)
    resistance = cells * 1.0e-3 / (sigma * 1.0e-3 * 35e-6)
    np.testing.assert_allclose(v[..., 0] - v[..., cells], resistance, rtol=1e-9)
    np.testing.assert_allclose(v[..., cells // 2], 0.0, atol=1e-9 * resistance)
