"""Structured hexahedral Q1 mesh of a layered board.

The reference interface is ``LayeredThermalMesh`` of
src/thermal/matrix_free_mpir_fem/mesh.py:126-330: the same constructor
arguments, accepted shapes, properties and error messages, and the same
node and element layout (nodes ``(slabs + 1, rows + 1, cols + 1)`` in C order
with x fastest, local corners ``4 dz + 2 dy + dx``), since arrays cross to the
device in that layout.

Geometry is kept as three broadcastable cell-length arrays (z, y, x).  Face
areas, volumes and coefficients are products of them; exposed faces come
from first differences of the padded activity mask; quantities that live on
nodes are spread from cells with the separable ``spread_to_nodes``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Literal, Sequence

import numpy as np

from electrical.matrix_free_mpir_fem.grid import check_pitch_axis

from ._fields import Field, per_element, spread_to_nodes

COPPER_THERMAL_CONDUCTIVITY_W_PER_M_K = 385.0
FR4_IN_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K = 0.8
FR4_THROUGH_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K = 0.3

Node = tuple[int, int, int]
Side = Literal["top", "bottom"]
Preconditioner = Literal["two-level", "jacobi"]
FaceDirection = Literal["-x", "+x", "-y", "+y", "-z", "+z"]
FACE_DIRECTIONS: tuple[FaceDirection, ...] = ("-x", "+x", "-y", "+y", "-z", "+z")

# Array axis of each coordinate in (slabs, rows, cols) and node arrays.
AXIS = {"z": 0, "y": 1, "x": 2}


def corner_views(grid: Any) -> tuple[Any, ...]:
    """The eight ``(slabs, rows, cols)`` corner views of a node grid, ordered
    ``4 dz + 2 dy + dx`` (the reference's local corner order, mesh.py:28-37);
    views, so they may be written through."""

    s, r, c = (n - 1 for n in grid.shape)
    return tuple(grid[k >> 2 : (k >> 2) + s, (k >> 1) & 1 : ((k >> 1) & 1) + r, k & 1 : (k & 1) + c]
                 for k in range(8))


def exposed_element_faces(active: np.ndarray) -> dict[FaceDirection, np.ndarray]:
    """Faces of active elements that border a void element or the grid edge
    (the reference's mesh.py:53-77): along each axis, the activity of the
    padded grid steps down (+ face) or up (- face) across the face."""

    mask = np.asarray(active, dtype=bool)
    if mask.ndim != 3:
        raise ValueError("active must have shape (slabs, rows, cols)")
    out: dict[FaceDirection, np.ndarray] = {}
    for name, axis in AXIS.items():
        pad = [(0, 0)] * 3
        pad[axis] = (1, 1)
        step = np.diff(np.pad(mask.astype(np.int8), pad), axis=axis)
        head = [slice(None)] * 3
        tail = [slice(None)] * 3
        head[axis], tail[axis] = slice(None, -1), slice(1, None)
        out[f"-{name}"] = step[tuple(head)] == 1  # type: ignore[index]
        out[f"+{name}"] = step[tuple(tail)] == -1  # type: ignore[index]
    return {d: out[d] for d in FACE_DIRECTIONS}


def cell_lengths(mesh: "LayeredThermalMesh") -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Edge lengths of the elements along z, y and x, shaped to broadcast over
    ``(slabs, rows, cols)``."""

    return (
        np.asarray(mesh.slab_thickness_m, dtype=np.float64)[:, None, None],
        np.asarray(mesh.pitch_y_m)[None, :, None],
        np.asarray(mesh.pitch_x_m)[None, None, :],
    )


def face_area_m2(mesh: "LayeredThermalMesh", direction: str) -> np.ndarray:
    """Area of every element face facing ``direction``: the product of the
    two edge lengths across the face normal."""

    lengths = cell_lengths(mesh)
    normal = AXIS[direction[1]]
    a, b = (lengths[i] for i in range(3) if i != normal)
    return np.array(np.broadcast_to(a * b, mesh.element_grid_shape))


def lump_face_weights(face_weight: np.ndarray, direction: str, node_shape) -> np.ndarray:
    """Node values of per-face weights: each face's weight split in quarters
    onto its four nodes (the reference's lumping, mesh.py:80-92).  The faces
    facing ``+a`` lie on the node plane above their element along a, those
    facing ``-a`` on the plane below."""

    normal = AXIS[direction[1]]
    across = tuple(i for i in range(3) if i != normal)
    quarters = spread_to_nodes(np.asarray(face_weight, dtype=np.float64) / 4.0, across)
    pad = [(0, 0)] * 3
    pad[normal] = (1, 0) if direction[0] == "+" else (0, 1)
    nodes = np.pad(quarters, pad)
    assert nodes.shape == tuple(node_shape)
    return nodes


_THICKNESS = Field((1,), "slab_thickness_m must list at least one slab", "slab thicknesses must be finite and positive",
                   low=0.0, low_open=True)


@dataclass(frozen=True)
class LayeredThermalMesh:
    """Structured hexahedral Q1 mesh of a PCB stack.

    ``slab_thickness_m`` lists element slabs bottom to top; pitches are one
    width or one value per column / row; conductivities are a scalar, one
    value per slab or one per element; ``active`` carves void elements;
    ``volumetric_heat_capacity_j_per_m3_k`` (same forms) is needed by
    transient solves only.
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
        thickness = _THICKNESS.read(self.slab_thickness_m)
        if thickness.size < 1:
            raise ValueError(_THICKNESS.rank_message)
        rows, cols = self._footprint()
        shape = (int(thickness.size), rows, cols)
        active = self._activity(shape)
        k_in = per_element(self.conductivity_w_per_m_k, shape, "conductivity_w_per_m_k")
        k_z = k_in.copy() if self.through_plane_conductivity_w_per_m_k is None else per_element(
            self.through_plane_conductivity_w_per_m_k, shape, "through_plane_conductivity_w_per_m_k")
        capacity = None
        if self.volumetric_heat_capacity_j_per_m3_k is not None:
            capacity = per_element(self.volumetric_heat_capacity_j_per_m3_k, shape,
                                   "volumetric_heat_capacity_j_per_m3_k")
        for name, values in (("conductivity_w_per_m_k", k_in), ("through_plane_conductivity_w_per_m_k", k_z),
                             ("volumetric_heat_capacity_j_per_m3_k", capacity)):
            if values is not None and not (np.all(np.isfinite(values)) and np.all(values[active] > 0.0)):
                raise ValueError(f"{name} must be finite and positive on active elements")
        fields = {
            "slab_thickness_m": tuple(float(t) for t in thickness),
            "element_shape": (rows, cols),
            "pitch_x_m": check_pitch_axis(np.asarray(self.pitch_x_m, dtype=np.float64), cols, "pitch_x_m"),
            "pitch_y_m": check_pitch_axis(np.asarray(self.pitch_y_m, dtype=np.float64), rows, "pitch_y_m"),
            "conductivity_w_per_m_k": np.where(active, k_in, 0.0),
            "through_plane_conductivity_w_per_m_k": np.where(active, k_z, 0.0),
            "active": active,
            "volumetric_heat_capacity_j_per_m3_k": None if capacity is None else np.where(active, capacity, 0.0),
        }
        for name, value in fields.items():
            object.__setattr__(self, name, value)

    def _footprint(self) -> tuple[int, int]:
        """(rows, cols): given, else read off a full conductivity array, else
        off two per-cell pitch arrays."""

        k = np.asarray(self.conductivity_w_per_m_k)
        px, py = np.asarray(self.pitch_x_m), np.asarray(self.pitch_y_m)
        if self.element_shape is not None:
            rows, cols = (int(v) for v in self.element_shape)
        elif k.ndim == 3:
            rows, cols = (int(v) for v in k.shape[1:])
        elif px.ndim == 1 and py.ndim == 1:
            rows, cols = int(py.size), int(px.size)
        else:
            raise ValueError(
                "element_shape is required unless conductivity has shape "
                "(slabs, rows, cols) or both pitches are per-cell arrays"
            )
        if min(rows, cols) < 1:
            raise ValueError("element_shape must have positive axes")
        return rows, cols

    def _activity(self, shape: tuple[int, int, int]) -> np.ndarray:
        if self.active is None:
            return np.ones(shape, dtype=bool)
        active = np.array(self.active, dtype=bool)
        if active.shape != shape:
            raise ValueError("active must match (slabs, rows, cols)")
        if not active.any():
            raise ValueError("at least one element must be active")
        return active

    @property
    def element_grid_shape(self) -> tuple[int, int, int]:
        rows, cols = self.element_shape  # type: ignore[misc]
        return len(self.slab_thickness_m), rows, cols

    @property
    def node_shape(self) -> tuple[int, int, int]:
        return tuple(n + 1 for n in self.element_grid_shape)  # type: ignore[return-value]

    @property
    def size(self) -> int:
        return int(np.prod(self.node_shape))

    @property
    def uniform_pitch(self) -> bool:
        return bool(np.ptp(self.pitch_x_m) == 0.0 and np.ptp(self.pitch_y_m) == 0.0)

    @property
    def cell_area_m2(self) -> np.ndarray:
        return np.multiply.outer(self.pitch_y_m, self.pitch_x_m)

    @property
    def x_edges_m(self) -> np.ndarray:
        return np.cumsum(np.r_[0.0, self.pitch_x_m])

    @property
    def y_edges_m(self) -> np.ndarray:
        return np.cumsum(np.r_[0.0, self.pitch_y_m])

    @property
    def element_volume_m3(self) -> np.ndarray:
        hz, hy, hx = cell_lengths(self)
        return hz * (hy * hx)

    def element_coefficients(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """``a_x = k_in hy hz / hx``, ``a_y = k_in hx hz / hy``, ``a_z = k_z hx hy / hz``
        per element, zero on void (the reference's mesh.py:276-293, evaluated
        left to right as there, so the bits agree)."""

        hz, hy, hx = cell_lengths(self)
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
        """An eighth of every element's heat capacity on each of its corners."""

        if self.volumetric_heat_capacity_j_per_m3_k is None:
            raise ValueError("the mesh has no volumetric_heat_capacity_j_per_m3_k; a transient solve needs one")
        return spread_to_nodes(self.volumetric_heat_capacity_j_per_m3_k * self.element_volume_m3 / 8.0, (0, 1, 2))

    @property
    def is_full(self) -> bool:
        return bool(np.all(self.active))

    @property
    def active_nodes(self) -> np.ndarray:
        return spread_to_nodes(np.asarray(self.active, dtype=np.float64), (0, 1, 2)) > 0.0

    def exposed_faces(self) -> dict[FaceDirection, np.ndarray]:
        return exposed_element_faces(self.active)  # type: ignore[arg-type]


def flat_index(node: Node, shape: tuple[int, int, int]) -> int:
    """Flat C-order index of a (slab face, row, column) node, refusing nodes
    outside the grid."""

    if len(node) != 3 or min(node) < 0:
        raise ValueError(f"invalid node {node!r}")
    if any(i >= n for i, n in zip(node, shape)):
        raise ValueError(f"node {node!r} lies outside mesh shape {shape}")
    z, y, x = (int(i) for i in node)
    return (z * shape[1] + y) * shape[2] + x


def unit_hexahedron_matrices() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The three (8, 8) unit trilinear stiffness matrices: with the 1-D
    stiffness ``S = [[1, -1], [-1, 1]]`` and mass ``M = [[2, 1], [1, 2]] / 6``,
    ``U_x = M(z) M(y) S(x)``, ``U_y = M S M`` and ``U_z = S M M`` over the
    corners ``4 dz + 2 dy + dx`` (the reference's mesh.py:346-359)."""

    one_d = {"S": np.array([[1.0, -1.0], [-1.0, 1.0]]), "M": np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0}

    def tensor(z: str, y: str, x: str) -> np.ndarray:
        # z (y x): the factors are associated as in the oracle, so every
        # entry rounds the same way.
        yx = np.einsum("be,cf->becf", one_d[y], one_d[x])
        t = np.einsum("ad,becf->abcdef", one_d[z], yx)
        return t.reshape(8, 8)

    return tensor("M", "M", "S"), tensor("M", "S", "M"), tensor("S", "M", "M")
