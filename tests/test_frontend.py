"""Host-side front end: the problem contract of the reference
(tests/test_thermal_conduction.py:197-236) and the workload generators."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import problem as P
from pcbhip import workloads
from tests.helpers import random_body, to_oracle
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    ExposedFaceConvection,
    HeatSource,
    LayeredThermalMesh,
    ThermalConductionProblem,
)


def _uniform(slabs, t, shape, k):
    return LayeredThermalMesh((t,) * slabs, 1.0e-3, 1.0e-3, k, element_shape=shape)


def test_floating_problem_and_bad_inputs_are_rejected():
    mesh = _uniform(1, 1.0e-3, (1, 1), 1.0)
    with pytest.raises(ValueError, match="positive film coefficient, a radiating face or a fixed node"):
        ThermalConductionProblem(mesh)
    with pytest.raises(ValueError, match="positive film coefficient, a radiating face or a fixed node"):
        ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 0.0, 300.0),))
    with pytest.raises(ValueError, match="element_shape is required"):
        LayeredThermalMesh((1.0e-3,), 1.0e-3, 1.0e-3, 1.0)
    with pytest.raises(ValueError, match="finite and positive"):
        LayeredThermalMesh((1.0e-3,), 1.0e-3, 1.0e-3, 0.0, element_shape=(1, 1))
    with pytest.raises(ValueError, match="side"):
        ConvectionBoundary("left", 1.0, 300.0)  # type: ignore[arg-type]
    with pytest.raises(ValueError, match="outside mesh shape"):
        ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 1.0, 300.0),),
                                 heat_sources=(HeatSource(((5, 0, 0),), 1.0),))
    with pytest.raises(ValueError, match="element_heat_w"):
        ThermalConductionProblem(mesh, convection=(ConvectionBoundary("top", 1.0, 300.0),),
                                 element_heat_w=np.zeros((2, 1, 1)))


def test_mesh_broadcasts_per_slab_conductivity_and_reports_volumes():
    mesh = LayeredThermalMesh((35.0e-6, 1.5e-3), 0.2e-3, 0.1e-3, (385.0, 0.8), element_shape=(2, 3),
                              through_plane_conductivity_w_per_m_k=(385.0, 0.3))
    assert mesh.element_grid_shape == (2, 2, 3)
    assert mesh.node_shape == (3, 3, 4)
    np.testing.assert_allclose(mesh.conductivity_w_per_m_k[1], 0.8)
    np.testing.assert_allclose(mesh.through_plane_conductivity_w_per_m_k[1], 0.3)
    np.testing.assert_allclose(mesh.element_volume_m3[0], 35.0e-6 * 0.2e-3 * 0.1e-3)


@pytest.mark.parametrize("seed", [0, 1])
def test_front_end_arrays_match_the_oracle(seed):
    # The host preparation the device receives (coefficients, Robin of both
    # face kinds, fixed rows, loads) equals the oracle's, bit for bit.
    problem = random_body((3, 4, 5), seed=seed, void=True)
    ops = P.operator_arrays(to_oracle(problem))
    mesh = problem.mesh
    np.testing.assert_array_equal(np.stack(mesh.element_coefficients()), ops["coef"])
    robin = sum(b.lumped_nodal_weights(mesh)[0] for b in problem.convection)
    np.testing.assert_allclose(robin, ops["robin_total"], rtol=1e-15, atol=0)
    fixed = (problem.fixed_temperature_mask | ~mesh.active_nodes).reshape(-1)
    np.testing.assert_array_equal(fixed, ops["fixed"])
    face = ExposedFaceConvection(5.0, 295.0)
    assert face.exposed_area_m2(mesh) > 0


def test_board_generators_have_the_stated_shapes_and_power():
    b1 = workloads.board_64x64x8()
    assert b1.mesh.element_grid_shape == (8, 64, 64)
    assert b1.mesh.size == 38025
    assert b1.element_heat_w.sum() == pytest.approx(1.0)
    k = b1.mesh.conductivity_w_per_m_k
    assert set(np.unique(k)) == {0.3, 385.0}
    assert np.all(k[1:-1] == 0.3)
    assert 0.2 < np.mean(k[0] == 385.0) < 0.6
    assert np.all(b1.fixed_temperature_mask[0]) and not np.any(b1.fixed_temperature_mask[1:])
    again = workloads.board_64x64x8()
    np.testing.assert_array_equal(again.mesh.conductivity_w_per_m_k, k)
    shapes = workloads.uneven_board_shapes()
    assert len(shapes) == 200
    assert all(4 <= s <= 16 and 32 <= r <= 384 and 32 <= c <= 384 for s, r, c in shapes)


def test_mpir_dispatch_refuses_complex_and_host_systems_and_validates_inputs():
    from electrical.matrix_free_mpir_fem import MPIRConfig, solve_mpir

    class ComplexSystem:
        size = 4
        inner_solver = "gmres"
        high_dtype = np.complex128

    class NoDevice:
        size = 4
        inner_solver = "pcg"
        high_dtype = np.float64

    with pytest.raises(NotImplementedError, match="complex"):
        solve_mpir(ComplexSystem(), np.ones(4), config=MPIRConfig())
    with pytest.raises(TypeError, match="native_solve_mpir"):
        solve_mpir(NoDevice(), np.ones(4))
    with pytest.raises(ValueError, match="finite"):
        solve_mpir(NoDevice(), np.array([1.0, np.nan, 0.0, 0.0]))
    with pytest.raises(ValueError, match="size"):
        solve_mpir(NoDevice(), np.ones(3))
