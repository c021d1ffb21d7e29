"""Structured hexahedral Q1 mesh of a layered board (reference:
src/thermal/matrix_free_mpir_fem/mesh.py).

Same constructor, validation, properties and node/element layout as the
reference: nodes ``(slabs + 1, rows + 1, cols + 1)`` in C order with x
fastest, local corners ``4 dz + 2 dy + dx``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Literal, Sequence

import numpy as np

from electrical.matrix_free_mpir_fem.grid import check_pitch_axis

COPPER_THERMAL_CONDUCTIVITY_W_PER_M_K = 385.0
FR4_IN_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K = 0.8
FR4_THROUGH_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K = 0.3

Node = tuple[int, int, int]
Side = Literal["top", "bottom"]
Preconditioner = Literal["two-level", "jacobi"]
FaceDirection = Literal["-x", "+x", "-y", "+y", "-z", "+z"]
FACE_DIRECTIONS: tuple[FaceDirection, ...] = ("-x", "+x", "-y", "+y", "-z", "+z")

# Local corners of each element face (mesh.py:45-53).
FACE_CORNERS: dict[str, tuple[int, int, int, int]] = {
    "-z": (0, 1, 2, 3),
    "+z": (4, 5, 6, 7),
    "-y": (0, 1, 4, 5),
    "+y": (2, 3, 6, 7),
    "-x": (0, 2, 4, 6),
    "+x": (1, 3, 5, 7),
}


def corner_views(grid: Any) -> tuple[Any, ...]:
    """The eight ``(slabs, rows, cols)`` corner views of a node grid, ordered
    ``4 dz + 2 dy + dx`` (mesh.py:28-37)."""

    s, r, c = (n - 1 for n in grid.shape)
    return tuple(grid[dz : dz + s, dy : dy + r, dx : dx + c] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1))


def exposed_element_faces(active: np.ndarray) -> dict[FaceDirection, np.ndarray]:
    """Faces of active elements bordering a void element or the grid edge (mesh.py:53-77)."""

    mask = np.asarray(active, dtype=bool)
    if mask.ndim != 3:
        raise ValueError("active must have shape (slabs, rows, cols)")
    out: dict[FaceDirection, np.ndarray] = {}
    for direction in FACE_DIRECTIONS:
        axis = "zyx".index(direction[1])
        neighbour = np.zeros_like(mask)
        head = [slice(None)] * 3
        tail = [slice(None)] * 3
        head[axis], tail[axis] = slice(None, -1), slice(1, None)
        if direction[0] == "+":
            neighbour[tuple(head)] = mask[tuple(tail)]
        else:
            neighbour[tuple(tail)] = mask[tuple(head)]
        out[direction] = mask & ~neighbour
    return out


def lump_faces_onto_nodes(face_weight: np.ndarray, direction: str, node_shape) -> np.ndarray:
    """A quarter of every face weight onto each of the face's four nodes (mesh.py:80-92)."""

    weights = np.zeros(node_shape)
    views = corner_views(weights)
    quarter = face_weight / 4.0
    for corner in FACE_CORNERS[direction]:
        views[corner][...] += quarter
    return weights


def face_area_m2(mesh: "LayeredThermalMesh", direction: str) -> np.ndarray:
    """Area of every element face facing ``direction`` (mesh.py:95-106)."""

    shape = mesh.element_grid_shape
    hz = np.asarray(mesh.slab_thickness_m, dtype=np.float64)[:, None, None]
    hx = mesh.pitch_x_m[None, None, :]
    hy = mesh.pitch_y_m[None, :, None]
    area = {"z": hx * hy, "y": hx * hz, "x": hy * hz}[direction[1]]
    return np.broadcast_to(area, shape).copy()


def _element_field(value: Any, shape: tuple[int, int, int], name: str) -> np.ndarray:
    array = np.asarray(value, dtype=np.float64)
    if array.ndim == 0:
        return np.full(shape, float(array))
    if array.shape == (shape[0],):
        return np.broadcast_to(array[:, None, None], shape).copy()
    if array.shape == shape:
        return array.copy()
    raise ValueError(f"{name} must be a scalar, one value per slab, or match shape {shape}")


@dataclass(frozen=True)
class LayeredThermalMesh:
    """Structured hexahedral Q1 mesh of a PCB stack (mesh.py:126-330).

    ``slab_thickness_m`` lists element slabs bottom to top; pitches are one
    width or one value per column / row; conductivities are a scalar, one
    value per slab or one per element; ``active`` carves void elements.
    """

    slab_thickness_m: Sequence[float]
    pitch_x_m: float | Sequence[float] | np.ndarray
    pitch_y_m: float | Sequence[float] | np.ndarray
    conductivity_w_per_m_k: float | Sequence[float] | np.ndarray
    element_shape: tuple[int, int] | None = None
    through_plane_conductivity_w_per_m_k: float | Sequence[float] | np.ndarray | None = None
    active: np.ndarray | None = None
    volumetric_heat_capacity_j_per_m3_k: float | Sequence[float] | np.ndarray | None = None

    def __post_init__(self) -> None:
        thickness = np.asarray(self.slab_thickness_m, dtype=np.float64)
        if thickness.ndim != 1 or thickness.size < 1:
            raise ValueError("slab_thickness_m must list at least one slab")
        if not np.all(np.isfinite(thickness)) or np.any(thickness <= 0.0):
            raise ValueError("slab thicknesses must be finite and positive")
        k_in = np.asarray(self.conductivity_w_per_m_k, dtype=np.float64)
        px = np.asarray(self.pitch_x_m, dtype=np.float64)
        py = np.asarray(self.pitch_y_m, dtype=np.float64)
        if self.element_shape is not None:
            rows, cols = (int(v) for v in self.element_shape)
        elif k_in.ndim == 3:
            rows, cols = int(k_in.shape[1]), int(k_in.shape[2])
        elif px.ndim == 1 and py.ndim == 1:
            rows, cols = int(py.size), int(px.size)
        else:
            raise ValueError(
                "element_shape is required unless conductivity has shape "
                "(slabs, rows, cols) or both pitches are per-cell arrays"
            )
        if rows < 1 or cols < 1:
            raise ValueError("element_shape must have positive axes")
        shape = (int(thickness.size), rows, cols)
        object.__setattr__(self, "pitch_x_m", check_pitch_axis(px, cols, "pitch_x_m"))
        object.__setattr__(self, "pitch_y_m", check_pitch_axis(py, rows, "pitch_y_m"))
        k_in = _element_field(k_in, shape, "conductivity_w_per_m_k")
        if self.through_plane_conductivity_w_per_m_k is None:
            k_z = k_in.copy()
        else:
            k_z = _element_field(self.through_plane_conductivity_w_per_m_k, shape,
                                 "through_plane_conductivity_w_per_m_k")
        if self.active is None:
            active = np.ones(shape, dtype=bool)
        else:
            active = np.asarray(self.active, dtype=bool)
            if active.shape != shape:
                raise ValueError("active must match (slabs, rows, cols)")
            if not np.any(active):
                raise ValueError("at least one element must be active")
            active = active.copy()
        for name, array in (("conductivity_w_per_m_k", k_in), ("through_plane_conductivity_w_per_m_k", k_z)):
            if not np.all(np.isfinite(array)) or np.any(array[active] <= 0.0):
                raise ValueError(f"{name} must be finite and positive on active elements")
        capacity = None
        if self.volumetric_heat_capacity_j_per_m3_k is not None:
            capacity = _element_field(self.volumetric_heat_capacity_j_per_m3_k, shape,
                                      "volumetric_heat_capacity_j_per_m3_k")
            if not np.all(np.isfinite(capacity)) or np.any(capacity[active] <= 0.0):
                raise ValueError(
                    "volumetric_heat_capacity_j_per_m3_k must be finite and positive on active elements")
            capacity = np.where(active, capacity, 0.0)
        object.__setattr__(self, "slab_thickness_m", tuple(thickness.tolist()))
        object.__setattr__(self, "element_shape", (rows, cols))
        object.__setattr__(self, "conductivity_w_per_m_k", np.where(active, k_in, 0.0))
        object.__setattr__(self, "through_plane_conductivity_w_per_m_k", np.where(active, k_z, 0.0))
        object.__setattr__(self, "active", active)
        object.__setattr__(self, "volumetric_heat_capacity_j_per_m3_k", capacity)

    @property
    def element_grid_shape(self) -> tuple[int, int, int]:
        rows, cols = self.element_shape  # type: ignore[misc]
        return len(self.slab_thickness_m), rows, cols

    @property
    def node_shape(self) -> tuple[int, int, int]:
        s, r, c = self.element_grid_shape
        return s + 1, r + 1, c + 1

    @property
    def size(self) -> int:
        return int(np.prod(self.node_shape))

    @property
    def uniform_pitch(self) -> bool:
        return bool(np.all(self.pitch_x_m == self.pitch_x_m[0]) and np.all(self.pitch_y_m == self.pitch_y_m[0]))

    @property
    def cell_area_m2(self) -> np.ndarray:
        return self.pitch_y_m[:, None] * self.pitch_x_m[None, :]

    @property
    def x_edges_m(self) -> np.ndarray:
        return np.concatenate(([0.0], np.cumsum(self.pitch_x_m)))

    @property
    def y_edges_m(self) -> np.ndarray:
        return np.concatenate(([0.0], np.cumsum(self.pitch_y_m)))

    @property
    def element_volume_m3(self) -> np.ndarray:
        hz = np.asarray(self.slab_thickness_m, dtype=np.float64)
        return hz[:, None, None] * self.cell_area_m2[None, :, :]

    def element_coefficients(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """``a_x = k_in hy hz / hx``, ``a_y = k_in hx hz / hy``, ``a_z = k_z hx hy / hz``
        per element, zero on void (mesh.py:276-293)."""

        hz = np.asarray(self.slab_thickness_m, dtype=np.float64)[:, None, None]
        hx = self.pitch_x_m[None, None, :]
        hy = self.pitch_y_m[None, :, None]
        k_in = np.asarray(self.conductivity_w_per_m_k)
        k_z = np.asarray(self.through_plane_conductivity_w_per_m_k)
        return (
            np.ascontiguousarray(k_in * hy * hz / hx),
            np.ascontiguousarray(k_in * hx * hz / hy),
            np.ascontiguousarray(k_z * hx * hy / hz),
        )

    @property
    def has_heat_capacity(self) -> bool:
        return self.volumetric_heat_capacity_j_per_m3_k is not None

    def nodal_heat_capacity_j_per_k(self) -> np.ndarray:
        if self.volumetric_heat_capacity_j_per_m3_k is None:
            raise ValueError("the mesh has no volumetric_heat_capacity_j_per_m3_k; a transient solve needs one")
        element = self.volumetric_heat_capacity_j_per_m3_k * self.element_volume_m3 / 8.0
        nodal = np.zeros(self.node_shape)
        for view in corner_views(nodal):
            view[...] += element
        return nodal

    @property
    def is_full(self) -> bool:
        return bool(np.all(self.active))

    @property
    def active_nodes(self) -> np.ndarray:
        nodes = np.zeros(self.node_shape, dtype=bool)
        for view in corner_views(nodes):
            view[...] |= self.active  # type: ignore[operator]
        return nodes

    def exposed_faces(self) -> dict[FaceDirection, np.ndarray]:
        return exposed_element_faces(self.active)  # type: ignore[arg-type]


def flat_index(node: Node, shape: tuple[int, int, int]) -> int:
    if len(node) != 3 or any(index < 0 for index in node):
        raise ValueError(f"invalid node {node!r}")
    try:
        return int(np.ravel_multi_index(node, shape))
    except ValueError as exc:
        raise ValueError(f"node {node!r} lies outside mesh shape {shape}") from exc


def unit_hexahedron_matrices() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The three (8, 8) unit trilinear stiffness matrices (mesh.py:346-359):
    ``U_x = M(z) M(y) S(x)``, ``U_y = M S M``, ``U_z = S M M`` with
    ``S = [[1, -1], [-1, 1]]`` and ``M = [[2, 1], [1, 2]] / 6``."""

    s = np.array([[1.0, -1.0], [-1.0, 1.0]])
    m = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    return np.kron(m, np.kron(m, s)), np.kron(m, np.kron(s, m)), np.kron(s, np.kron(m, m))
