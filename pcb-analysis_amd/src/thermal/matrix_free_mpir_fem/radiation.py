"""Grey-body radiation to an ambient as a Newton-linearised Robin boundary
(reference: src/thermal/matrix_free_mpir_fem/radiation.py:1-196).

A face at ``T`` radiating to ``T_amb`` loses ``eps sigma (T^4 - T_amb^4)``
per unit area.  Expanding about the current iterate ``T_k`` gives a film
coefficient ``h_k = 4 eps sigma T_k^3`` and an effective ambient
``T_eff = T_k - (T_k^4 - T_amb^4) / (4 T_k^3)``, exact at ``T_k``: an
ordinary convection face per iteration.  Iterating the linear solve to a
fixed point is Newton's method on the whole problem, because conduction is
linear.  No view factors: every face sees only its ambient.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .boundaries import ConvectionBoundary, ExposedFaceConvection
from .mesh import FACE_DIRECTIONS, FaceDirection, LayeredThermalMesh, Side, corner_views

STEFAN_BOLTZMANN_W_PER_M2_K4 = 5.670374419e-8


def _emissivity(value, ndims: tuple[int, ...]):
    e = np.asarray(value, dtype=np.float64)
    if e.ndim not in ndims:
        raise ValueError("emissivity must be a scalar or an array matching the faces")
    if not np.all(np.isfinite(e)) or np.any((e < 0.0) | (e > 1.0)):
        raise ValueError("emissivity must lie in [0, 1]")
    return float(e) if e.ndim == 0 else e.copy()


def _ambient(value, ndims: tuple[int, ...]):
    t = np.asarray(value, dtype=np.float64)
    if t.ndim not in ndims:
        raise ValueError("ambient must be a scalar or an array matching the faces")
    if not np.all(np.isfinite(t)) or np.any(t <= 0.0):
        raise ValueError("radiative ambient temperature must be finite and positive (kelvin)")
    return float(t) if t.ndim == 0 else t.copy()


def newton_linearisation(emissivity, surface_temperature_k, ambient_temperature_k) -> tuple[np.ndarray, np.ndarray]:
    """``(h_k, T_eff)`` of the Newton step at the given surface temperature
    (radiation.py:60-73); non-finite or non-positive surface values fall back
    to the ambient."""

    ambient = np.asarray(ambient_temperature_k, dtype=np.float64)
    surface = np.asarray(surface_temperature_k, dtype=np.float64)
    surface = np.where(np.isfinite(surface) & (surface > 0.0), surface, ambient)
    t3 = surface * surface * surface
    coefficient = 4.0 * np.asarray(emissivity, dtype=np.float64) * STEFAN_BOLTZMANN_W_PER_M2_K4 * t3
    effective = surface - (surface * t3 - ambient**4) / (4.0 * t3)
    return coefficient, np.array(np.broadcast_to(effective, surface.shape))


def face_temperature_k(mesh: LayeredThermalMesh, temperature_k, side: Side) -> np.ndarray:
    """Mean of the four corner temperatures of every top or bottom face, ``(rows, cols)``."""

    grid = np.asarray(temperature_k, dtype=np.float64).reshape(mesh.node_shape)
    plane = grid[-1] if side == "top" else grid[0]
    return 0.25 * ((plane[:-1, :-1] + plane[:-1, 1:]) + (plane[1:, :-1] + plane[1:, 1:]))


def element_temperature_k(mesh: LayeredThermalMesh, temperature_k) -> np.ndarray:
    """Mean of the eight corner temperatures of every element, ``(slabs, rows, cols)``."""

    grid = np.asarray(temperature_k, dtype=np.float64).reshape(mesh.node_shape)
    total = np.zeros(mesh.element_grid_shape)
    for view in corner_views(grid):
        total += view
    return total / 8.0


@dataclass(frozen=True)
class RadiationBoundary:
    """Radiation from the top or bottom face of the stack (radiation.py:98-138).
    Emissivity and ambient are scalars or ``(rows, cols)`` arrays; faces of
    inactive elements radiate nothing."""

    side: Side
    emissivity: float | np.ndarray
    ambient_temperature_k: float | np.ndarray

    def __post_init__(self) -> None:
        if self.side not in ("top", "bottom"):
            raise ValueError("side must be 'top' or 'bottom'")
        object.__setattr__(self, "emissivity", _emissivity(self.emissivity, (0, 2)))
        object.__setattr__(self, "ambient_temperature_k", _ambient(self.ambient_temperature_k, (0, 2)))

    def check_shape(self, mesh: LayeredThermalMesh) -> None:
        for name in ("emissivity", "ambient_temperature_k"):
            value = np.asarray(getattr(self, name))
            if value.ndim == 2 and value.shape != mesh.element_grid_shape[1:]:
                raise ValueError(f"{name} array must match (rows, cols) of the mesh")

    @property
    def radiates(self) -> bool:
        return bool(np.any(np.asarray(self.emissivity) > 0.0))

    def mean_ambient_k(self) -> float:
        return float(np.mean(self.ambient_temperature_k))

    def linearize(self, mesh: LayeredThermalMesh, temperature_k) -> ConvectionBoundary:
        plane = mesh.element_grid_shape[1:]
        h, t_eff = newton_linearisation(self.emissivity, face_temperature_k(mesh, temperature_k, self.side),
                                        np.broadcast_to(self.ambient_temperature_k, plane))
        return ConvectionBoundary(self.side, h, t_eff)


@dataclass(frozen=True)
class ExposedFaceRadiation:
    """Radiation from every exposed face of the active elements in
    ``directions`` (radiation.py:141-180); emissivity and ambient are scalars
    or one value per element."""

    emissivity: float | np.ndarray
    ambient_temperature_k: float | np.ndarray
    directions: tuple[FaceDirection, ...] = FACE_DIRECTIONS

    def __post_init__(self) -> None:
        directions = tuple(self.directions)
        if not directions or any(d not in FACE_DIRECTIONS for d in directions):
            raise ValueError(f"directions must be drawn from {FACE_DIRECTIONS}")
        if len(set(directions)) != len(directions):
            raise ValueError("directions must be unique")
        object.__setattr__(self, "emissivity", _emissivity(self.emissivity, (0, 3)))
        object.__setattr__(self, "ambient_temperature_k", _ambient(self.ambient_temperature_k, (0, 3)))
        object.__setattr__(self, "directions", directions)

    def check_shape(self, mesh: LayeredThermalMesh) -> None:
        for name in ("emissivity", "ambient_temperature_k"):
            value = np.asarray(getattr(self, name))
            if value.ndim == 3 and value.shape != mesh.element_grid_shape:
                raise ValueError(f"{name} array must match (slabs, rows, cols) of the mesh")

    @property
    def radiates(self) -> bool:
        return bool(np.any(np.asarray(self.emissivity) > 0.0))

    def mean_ambient_k(self) -> float:
        return float(np.mean(self.ambient_temperature_k))

    def linearize(self, mesh: LayeredThermalMesh, temperature_k) -> ExposedFaceConvection:
        shape = mesh.element_grid_shape
        h, t_eff = newton_linearisation(self.emissivity, element_temperature_k(mesh, temperature_k),
                                        np.broadcast_to(self.ambient_temperature_k, shape))
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
