"""The steady conduction problem: mesh, boundaries, fixed nodes and loads.

The reference interface is ``ThermalConductionProblem`` of
src/thermal/matrix_free_mpir_fem/problem.py:15-124: the same fields,
normalised forms (a full fixed mask and fixed values, zero-filled loads,
tuples) and error messages.  Here the boundary kinds and the two load
fields are tables that one loop checks.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .boundaries import Convection, ConvectionBoundary, ExposedFaceConvection, HeatSource
from .mesh import LayeredThermalMesh, flat_index
from .radiation import ExposedFaceRadiation, RadiationBoundary

# (attribute, accepted kinds, message for anything else, the property that
# says whether an entry sets the temperature level)
_BOUNDARY_GROUPS = (
    ("convection", (ConvectionBoundary, ExposedFaceConvection),
     "convection entries must be ConvectionBoundary or ExposedFaceConvection", "cools"),
    ("radiation", (RadiationBoundary, ExposedFaceRadiation),
     "radiation entries must be RadiationBoundary or ExposedFaceRadiation", "radiates"),
)

# (attribute, grid, shape message, finiteness message, inactive-support message)
_LOADS = (
    ("element_heat_w", "element", "element_heat_w must match (slabs, rows, cols)", "element heat must be finite",
     "element heat must be zero on inactive elements"),
    ("nodal_heat_w", "node", "nodal_heat_w must match mesh.node_shape", "nodal heat must be finite",
     "nodal heat must be zero on inactive nodes"),
)


def fixed_nodes(mask, values, node_shape) -> tuple[np.ndarray, np.ndarray]:
    """The fixed-temperature mask and a full array of fixed values (zero
    where not fixed); no mask means no fixed node."""

    if mask is None:
        if values is not None:
            raise ValueError("fixed_temperature_k needs a fixed_temperature_mask")
        return np.zeros(node_shape, dtype=bool), np.zeros(node_shape)
    mask = np.asarray(mask, dtype=bool)
    if mask.shape != node_shape:
        raise ValueError("fixed_temperature_mask must match mesh.node_shape")
    if values is None:
        raise ValueError("fixed_temperature_mask needs fixed_temperature_k")
    values = np.asarray(values, dtype=np.float64)
    if values.ndim != 0 and values.shape != node_shape:
        raise ValueError("fixed_temperature_k must be a scalar or match mesh.node_shape")
    values = np.broadcast_to(values, node_shape)
    if not np.isfinite(values[mask]).all():
        raise ValueError("fixed temperatures must be finite")
    return mask.copy(), np.where(mask, values, 0.0)


@dataclass(frozen=True)
class ThermalConductionProblem:
    """Steady conduction with convective and radiating faces and
    fixed-temperature nodes.

    ``element_heat_w`` is each element's power, lumped equally onto its eight
    corners; ``nodal_heat_w`` is power already on nodes.  At least one
    positive film coefficient, one radiating face or one fixed node must set
    the temperature level.  Radiation makes the problem nonlinear;
    ``solve_thermal_conduction`` iterates its Newton linearisation.
    """

    mesh: LayeredThermalMesh
    convection: tuple[Convection, ...] = ()
    fixed_temperature_mask: np.ndarray | None = None
    fixed_temperature_k: float | np.ndarray | None = None
    heat_sources: tuple[HeatSource, ...] = ()
    element_heat_w: np.ndarray | None = None
    nodal_heat_w: np.ndarray | None = None
    radiation: tuple = ()

    def __post_init__(self) -> None:
        mesh = self.mesh
        normalised: dict[str, object] = {}
        sets_level = False
        for name, kinds, message, flag in _BOUNDARY_GROUPS:
            entries = tuple(getattr(self, name))
            for entry in entries:
                if not isinstance(entry, kinds):
                    raise TypeError(message)
                entry.check_shape(mesh)
            sets_level = sets_level or any(getattr(e, flag) for e in entries)
            normalised[name] = entries

        mask, values = fixed_nodes(self.fixed_temperature_mask, self.fixed_temperature_k, mesh.node_shape)
        active_nodes = mesh.active_nodes
        if not (sets_level or (mask & active_nodes).any()):
            raise ValueError("the problem needs a positive film coefficient, a radiating face or a fixed node")
        normalised["fixed_temperature_mask"] = mask
        normalised["fixed_temperature_k"] = values

        support = {"element": np.asarray(mesh.active, dtype=bool), "node": active_nodes}
        shapes = {"element": mesh.element_grid_shape, "node": mesh.node_shape}
        for name, grid, shape_message, finite_message, support_message in _LOADS:
            value = getattr(self, name)
            load = np.zeros(shapes[grid]) if value is None else np.array(value, dtype=np.float64)
            if load.shape != shapes[grid]:
                raise ValueError(shape_message)
            if not np.isfinite(load).all():
                raise ValueError(finite_message)
            if (load[~support[grid]] != 0.0).any():
                raise ValueError(support_message)
            normalised[name] = load

        sources = tuple(self.heat_sources)
        for source in sources:
            for node in source.nodes:
                if not active_nodes.flat[flat_index(node, mesh.node_shape)]:
                    raise ValueError(f"heat source {source.name!r} touches inactive node {node!r}")
        normalised["heat_sources"] = sources

        for name, value in normalised.items():
            object.__setattr__(self, name, value)
