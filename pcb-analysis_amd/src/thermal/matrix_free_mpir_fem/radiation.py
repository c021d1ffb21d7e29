"""Grey-body radiation to an ambient as a Newton-linearised Robin boundary.

The reference interface is src/thermal/matrix_free_mpir_fem/radiation.py:43-196:
``RadiationBoundary`` (top or bottom face), ``ExposedFaceRadiation`` (exposed
faces in chosen directions), ``newton_linearisation`` and the face and
element mean temperatures, with the reference's arguments and messages.

A face at ``T`` radiating to ``T_a`` loses ``eps sigma (T^4 - T_a^4)`` per
unit area.  Its tangent at the current iterate ``T_k`` is a film
coefficient ``h_k = 4 eps sigma T_k^3`` with an effective ambient
``T_eff = (3 T_k^4 + T_a^4) / (4 T_k^3)``, so ``h_k (T - T_eff)`` equals the
loss and its slope at ``T_k``.  Each radiating boundary therefore becomes a
convection boundary per Newton iterate; because conduction is linear,
solving those to a fixed point is Newton's method on the whole problem.  No
view factors: every face sees only its ambient.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._fields import Field, average_to_cells, require_shape, unique_directions
from .boundaries import ConvectionBoundary, ExposedFaceConvection, check_side
from .mesh import FACE_DIRECTIONS, FaceDirection, LayeredThermalMesh, Side

STEFAN_BOLTZMANN_W_PER_M2_K4 = 5.670374419e-8


def _emissivity_field(ranks: tuple[int, ...]) -> Field:
    return Field(ranks, "emissivity must be a scalar or an array matching the faces", "emissivity must lie in [0, 1]",
                 low=0.0, high=1.0)


def _ambient_field(ranks: tuple[int, ...]) -> Field:
    return Field(ranks, "ambient must be a scalar or an array matching the faces",
                 "radiative ambient temperature must be finite and positive (kelvin)", low=0.0, low_open=True)


def newton_linearisation(emissivity, surface_temperature_k, ambient_temperature_k) -> tuple[np.ndarray, np.ndarray]:
    """``(h_k, T_eff)`` of the Newton step at the given surface temperature
    (the reference's radiation.py:60-73).  A surface value that is not finite
    or not positive is replaced by the ambient, where the step is exact and
    carries no heat."""

    ambient = np.asarray(ambient_temperature_k, dtype=np.float64)
    surface = np.asarray(surface_temperature_k, dtype=np.float64)
    surface = np.where(np.isfinite(surface) & (surface > 0.0), surface, ambient)
    cube = surface**3
    coefficient = 4.0 * STEFAN_BOLTZMANN_W_PER_M2_K4 * np.asarray(emissivity, dtype=np.float64) * cube
    effective = (3.0 * surface**4 + np.broadcast_to(ambient, surface.shape) ** 4) / (4.0 * cube)
    return coefficient, np.array(np.broadcast_to(effective, surface.shape))


def face_temperature_k(mesh: LayeredThermalMesh, temperature_k, side: Side) -> np.ndarray:
    """Mean of the four corner temperatures of every top or bottom face, ``(rows, cols)``."""

    grid = np.asarray(temperature_k, dtype=np.float64).reshape(mesh.node_shape)
    return average_to_cells(grid[-1 if side == "top" else 0], (0, 1))


def element_temperature_k(mesh: LayeredThermalMesh, temperature_k) -> np.ndarray:
    """Mean of the eight corner temperatures of every element, ``(slabs, rows, cols)``."""

    return average_to_cells(np.asarray(temperature_k, dtype=np.float64).reshape(mesh.node_shape), (0, 1, 2))


class RadiatingFaces:
    """Shared behaviour: emissivity and ambient over the boundary's faces
    (scalars or face arrays of ``RANK`` dimensions), and the Newton step as a
    convection boundary of the same faces."""

    RANK = 0
    SHAPE_TEXT = ""

    def read_inputs(self) -> None:
        object.__setattr__(self, "emissivity", _emissivity_field((0, self.RANK)).read_compact(self.emissivity))
        object.__setattr__(self, "ambient_temperature_k",
                           _ambient_field((0, self.RANK)).read_compact(self.ambient_temperature_k))

    def face_shape(self, mesh: LayeredThermalMesh) -> tuple[int, ...]:
        return mesh.element_grid_shape[3 - self.RANK:]

    def check_shape(self, mesh: LayeredThermalMesh) -> None:
        for name in ("emissivity", "ambient_temperature_k"):
            require_shape(getattr(self, name), self.face_shape(mesh),
                          f"{name} array must match {self.SHAPE_TEXT} of the mesh", when_rank=self.RANK)

    @property
    def radiates(self) -> bool:
        return bool(np.any(np.asarray(self.emissivity) > 0.0))

    def mean_ambient_k(self) -> float:
        return float(np.mean(self.ambient_temperature_k))

    def surface_temperature_k(self, mesh: LayeredThermalMesh, temperature_k) -> np.ndarray:
        raise NotImplementedError

    def tangent(self, mesh: LayeredThermalMesh, temperature_k) -> tuple[np.ndarray, np.ndarray]:
        ambient = np.broadcast_to(np.asarray(self.ambient_temperature_k, dtype=np.float64), self.face_shape(mesh))
        return newton_linearisation(self.emissivity, self.surface_temperature_k(mesh, temperature_k), ambient)


@dataclass(frozen=True)
class RadiationBoundary(RadiatingFaces):
    """Radiation from the top or bottom face of the stack; emissivity and
    ambient are scalars or ``(rows, cols)`` arrays; faces of inactive
    elements radiate nothing."""

    side: Side
    emissivity: float | np.ndarray
    ambient_temperature_k: float | np.ndarray

    RANK = 2
    SHAPE_TEXT = "(rows, cols)"

    def __post_init__(self) -> None:
        check_side(self.side)
        self.read_inputs()

    def surface_temperature_k(self, mesh: LayeredThermalMesh, temperature_k) -> np.ndarray:
        return face_temperature_k(mesh, temperature_k, self.side)

    def linearize(self, mesh: LayeredThermalMesh, temperature_k) -> ConvectionBoundary:
        h, t_eff = self.tangent(mesh, temperature_k)
        return ConvectionBoundary(self.side, h, t_eff)


@dataclass(frozen=True)
class ExposedFaceRadiation(RadiatingFaces):
    """Radiation from every exposed face of the active elements facing one
    of ``directions``; emissivity and ambient are scalars or one value per
    element."""

    emissivity: float | np.ndarray
    ambient_temperature_k: float | np.ndarray
    directions: tuple[FaceDirection, ...] = FACE_DIRECTIONS

    RANK = 3
    SHAPE_TEXT = "(slabs, rows, cols)"

    def __post_init__(self) -> None:
        object.__setattr__(self, "directions", unique_directions(self.directions, FACE_DIRECTIONS))
        self.read_inputs()

    def surface_temperature_k(self, mesh: LayeredThermalMesh, temperature_k) -> np.ndarray:
        return element_temperature_k(mesh, temperature_k)

    def linearize(self, mesh: LayeredThermalMesh, temperature_k) -> ExposedFaceConvection:
        h, t_eff = self.tangent(mesh, temperature_k)
        return ExposedFaceConvection(h, t_eff, directions=self.directions)


Radiation = RadiationBoundary | ExposedFaceRadiation

__all__ = [
    "ExposedFaceRadiation",
    "Radiation",
    "RadiationBoundary",
    "STEFAN_BOLTZMANN_W_PER_M2_K4",
    "element_temperature_k",
    "face_temperature_k",
    "newton_linearisation",
]
