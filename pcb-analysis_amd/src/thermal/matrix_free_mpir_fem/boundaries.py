"""Convection boundaries and heat sources.

The reference interface is src/thermal/matrix_free_mpir_fem/boundaries.py:27-206:
``ConvectionBoundary`` (top or bottom face of the stack), ``ExposedFaceConvection``
(every exposed face of the active elements in chosen directions) and
``HeatSource``, with the reference's arguments, stored attribute forms and
error messages.

Both convection kinds are film faces: a set of element faces, each with a
film coefficient ``h`` and an ambient.  A boundary only says which faces it
covers (``face_sets``); the lumping onto nodes -- a quarter of every face's
``h A`` on each face node into the Robin conductance ``R`` and likewise
``h A T_amb`` into the load -- is shared, so the operator only ever sees one
Robin array.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Iterator

import numpy as np

from ._fields import Field, require_shape, unique_directions
from .mesh import FACE_DIRECTIONS, FaceDirection, LayeredThermalMesh, Node, Side, face_area_m2, lump_face_weights


class FilmFaces:
    """Shared behaviour of boundaries made of element faces with a film
    coefficient and an ambient.  Subclasses yield ``(direction, faces)``
    pairs, ``faces`` a boolean ``(slabs, rows, cols)`` array of the element
    faces facing ``direction`` that the boundary covers, and broadcast their
    coefficient and ambient over the element grid."""

    def face_sets(self, mesh: LayeredThermalMesh) -> Iterator[tuple[str, np.ndarray]]:
        raise NotImplementedError

    def element_values(self, mesh: LayeredThermalMesh, value) -> np.ndarray:
        raise NotImplementedError

    def face_conductance(self, mesh: LayeredThermalMesh, direction: str, faces: np.ndarray,
                         coefficient) -> np.ndarray:
        """``h A`` of every covered face facing ``direction`` (zero elsewhere)."""

        return self.element_values(mesh, coefficient) * face_area_m2(mesh, direction) * faces

    def lumped_weights(self, mesh: LayeredThermalMesh, coefficient, ambient) -> tuple[np.ndarray, np.ndarray]:
        amb = self.element_values(mesh, ambient)
        conductance = np.zeros(mesh.node_shape)
        load = np.zeros(mesh.node_shape)
        for direction, faces in self.face_sets(mesh):
            ha = self.face_conductance(mesh, direction, faces, coefficient)
            conductance += lump_face_weights(ha, direction, mesh.node_shape)
            load += lump_face_weights(ha * amb, direction, mesh.node_shape)
        return conductance.reshape(-1), load.reshape(-1)

    def covered_area_m2(self, mesh: LayeredThermalMesh) -> float:
        return float(sum(np.sum(face_area_m2(mesh, d) * faces) for d, faces in self.face_sets(mesh)))


# Inputs of the two convection kinds (ranks: scalar or face array).
_TOP_COEFFICIENT = Field((0, 2), "coefficient must be a scalar or a (rows, cols) array",
                         "film coefficients must be finite and non-negative", low=0.0)
_TOP_AMBIENT = Field((0, 2), "ambient must be a scalar or a (rows, cols) array", "ambient temperature must be finite")
_EXPOSED_COEFFICIENT = Field((0, 3), "film coefficient must be a scalar or a (slabs, rows, cols) array",
                             "film coefficient must be finite and non-negative", low=0.0)
_EXPOSED_AMBIENT = Field((0, 3), "ambient must be a scalar or a (slabs, rows, cols) array",
                         "ambient temperature must be finite")


def check_side(side: str) -> None:
    if side not in ("top", "bottom"):
        raise ValueError("side must be 'top' or 'bottom'")


def side_faces(mesh: LayeredThermalMesh, side: str) -> tuple[str, np.ndarray]:
    """The ``+z`` faces of the top slab's active elements, or the ``-z``
    faces of the bottom slab's."""

    faces = np.zeros(mesh.element_grid_shape, dtype=bool)
    slab = -1 if side == "top" else 0
    faces[slab] = np.asarray(mesh.active)[slab]
    return ("+z" if side == "top" else "-z"), faces


@dataclass(frozen=True)
class ConvectionBoundary(FilmFaces):
    """Newton cooling on the top or bottom face of the stack.

    Coefficient and ambient are scalars or ``(rows, cols)`` arrays; a zero
    coefficient switches a face off; faces of inactive elements do not cool.
    """

    side: Side
    coefficient_w_per_m2_k: float | np.ndarray
    ambient_temperature_k: float | np.ndarray

    def __post_init__(self) -> None:
        check_side(self.side)
        object.__setattr__(self, "coefficient_w_per_m2_k", _TOP_COEFFICIENT.read(self.coefficient_w_per_m2_k).copy())
        object.__setattr__(self, "ambient_temperature_k", _TOP_AMBIENT.read_compact(self.ambient_temperature_k))

    def check_shape(self, mesh: LayeredThermalMesh) -> None:
        plane = mesh.element_grid_shape[1:]
        require_shape(self.coefficient_w_per_m2_k, plane,
                      "convection coefficient array must match (rows, cols) of the mesh", when_rank=2)
        require_shape(self.ambient_temperature_k, plane,
                      "convection ambient array must match (rows, cols) of the mesh", when_rank=2)

    def face_sets(self, mesh: LayeredThermalMesh) -> Iterator[tuple[str, np.ndarray]]:
        yield side_faces(mesh, self.side)

    def element_values(self, mesh: LayeredThermalMesh, value) -> np.ndarray:
        # A (rows, cols) face array holds for every slab; only the side slab's
        # faces are covered.
        return np.array(np.broadcast_to(np.asarray(value, dtype=np.float64), mesh.element_grid_shape))

    @property
    def cools(self) -> bool:
        return bool(np.any(self.coefficient_w_per_m2_k > 0.0))

    def mean_ambient_k(self) -> float:
        return float(np.mean(self.ambient_temperature_k))

    def lumped_nodal_weights(self, mesh: LayeredThermalMesh) -> tuple[np.ndarray, np.ndarray]:
        """Nodal conductance ``R`` and load ``R T_amb`` of this face, flat
        (the reference's boundaries.py:82-102)."""

        return self.lumped_weights(mesh, self.coefficient_w_per_m2_k, self.ambient_temperature_k)


@dataclass(frozen=True)
class ExposedFaceConvection(FilmFaces):
    """Newton cooling on every exposed face of the active elements facing one
    of ``directions``; coefficient and ambient are scalars or one value per
    element."""

    coefficient_w_per_m2_k: float | np.ndarray
    ambient_temperature_k: float | np.ndarray
    directions: tuple[FaceDirection, ...] = FACE_DIRECTIONS

    def __post_init__(self) -> None:
        coefficient = _EXPOSED_COEFFICIENT.read_compact(self.coefficient_w_per_m2_k)
        ambient = _EXPOSED_AMBIENT.read_compact(self.ambient_temperature_k)
        object.__setattr__(self, "directions", unique_directions(self.directions, FACE_DIRECTIONS))
        object.__setattr__(self, "coefficient_w_per_m2_k", coefficient)
        object.__setattr__(self, "ambient_temperature_k", ambient)

    def check_shape(self, mesh: LayeredThermalMesh) -> None:
        for name in ("coefficient_w_per_m2_k", "ambient_temperature_k"):
            require_shape(getattr(self, name), mesh.element_grid_shape,
                          f"{name} array must match (slabs, rows, cols) of the mesh", when_rank=3)

    def face_sets(self, mesh: LayeredThermalMesh) -> Iterator[tuple[str, np.ndarray]]:
        exposed = mesh.exposed_faces()
        for direction in self.directions:
            yield direction, exposed[direction]

    def element_values(self, mesh: LayeredThermalMesh, value) -> np.ndarray:
        return np.array(np.broadcast_to(np.asarray(value, dtype=np.float64), mesh.element_grid_shape))

    @property
    def cools(self) -> bool:
        return bool(np.any(np.asarray(self.coefficient_w_per_m2_k) > 0.0))

    def mean_ambient_k(self) -> float:
        return float(np.mean(self.ambient_temperature_k))

    def exposed_area_m2(self, mesh: LayeredThermalMesh) -> float:
        return self.covered_area_m2(mesh)

    def lumped_nodal_weights(self, mesh: LayeredThermalMesh) -> tuple[np.ndarray, np.ndarray]:
        """Nodal conductance and load of every covered face, flat (the
        reference's boundaries.py:170-182)."""

        return self.lumped_weights(mesh, self.coefficient_w_per_m2_k, self.ambient_temperature_k)


Convection = ConvectionBoundary | ExposedFaceConvection


@dataclass(frozen=True)
class HeatSource:
    """A total power spread uniformly over a set of mesh nodes (the
    reference's boundaries.py:188-206)."""

    nodes: tuple[Node, ...]
    power_w: float
    name: str = "source"

    def __post_init__(self) -> None:
        nodes = tuple(tuple(int(i) for i in node) for node in self.nodes)
        problems = (
            (not nodes, "a heat source needs at least one node"),
            (any(len(node) != 3 for node in nodes), "heat source nodes must be (slab_face, row, column)"),
            (len(set(nodes)) != len(nodes), "heat source nodes must be unique"),
            (not np.isfinite(self.power_w), "heat source power must be finite"),
        )
        for failed, message in problems:
            if failed:
                raise ValueError(message)
        object.__setattr__(self, "nodes", nodes)
