"""Radiation boundaries on the CPU: the Newton linearisation is exact at its
point, and invalid boundaries are refused (the reference's
tests/test_thermal_radiation.py:60-112 checks; no device needed)."""

from __future__ import annotations

import numpy as np
import pytest

from thermal.matrix_free_mpir_fem import (
    STEFAN_BOLTZMANN_W_PER_M2_K4 as SIGMA,
    ExposedFaceRadiation,
    LayeredThermalMesh,
    RadiationBoundary,
    ThermalConductionProblem,
    element_temperature_k,
    face_temperature_k,
    newton_linearisation,
)

AMBIENT = 300.0


def test_newton_linearisation_is_exact_at_its_point():
    emissivity, surface = 0.7, np.array([350.0, 420.0, 300.0])
    h, t_eff = newton_linearisation(emissivity, surface, AMBIENT)
    exact = emissivity * SIGMA * (surface**4 - AMBIENT**4)
    np.testing.assert_allclose(h * (surface - t_eff), exact, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(h, 4.0 * emissivity * SIGMA * surface**3)
    # The tangent: the linearised flux has the true flux's slope at T_k.
    d = 1e-3
    slope = emissivity * SIGMA * ((surface + d) ** 4 - (surface - d) ** 4) / (2 * d)
    np.testing.assert_allclose(h, slope, rtol=1e-9)
    # A non-finite surface value falls back to the ambient (zero flux).
    h0, t0 = newton_linearisation(emissivity, np.array([np.nan]), AMBIENT)
    assert h0[0] * (AMBIENT - t0[0]) == pytest.approx(0.0, abs=1e-12)


def test_face_and_element_means():
    mesh = LayeredThermalMesh((1e-3, 1e-3), 1e-3, 1e-3, 10.0, element_shape=(2, 3))
    t = np.arange(mesh.size, dtype=float).reshape(mesh.node_shape)
    top = face_temperature_k(mesh, t, "top")
    assert top.shape == (2, 3)
    assert top[0, 0] == pytest.approx(np.mean(t[-1, :2, :2]))
    elem = element_temperature_k(mesh, t)
    assert elem[1, 1, 2] == pytest.approx(np.mean(t[1:3, 1:3, 2:4]))


def test_invalid_radiation_is_refused():
    mesh = LayeredThermalMesh((1e-3,), 1e-3, 1e-3, 200.0, element_shape=(8, 10))
    with pytest.raises(ValueError, match="radiating face"):
        ThermalConductionProblem(mesh, element_heat_w=np.zeros((1, 8, 10)))
    with pytest.raises(ValueError, match="emissivity"):
        RadiationBoundary("top", 1.2, AMBIENT)
    with pytest.raises(ValueError, match="kelvin"):
        ExposedFaceRadiation(0.5, 0.0)
    with pytest.raises(ValueError, match="rows, cols"):
        ThermalConductionProblem(mesh, radiation=(RadiationBoundary("top", np.ones((3, 3)), AMBIENT),))
    with pytest.raises(ValueError, match="side"):
        RadiationBoundary("left", 0.5, AMBIENT)
    # A radiating face alone is enough to fix the level.
    ThermalConductionProblem(mesh, radiation=(RadiationBoundary("top", 0.8, AMBIENT),))
