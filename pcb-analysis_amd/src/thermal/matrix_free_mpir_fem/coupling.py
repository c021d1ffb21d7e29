"""Joule loss of a layered-PCB DC solve as the load of a thermal solve.

The reference interface is src/thermal/matrix_free_mpir_fem/coupling.py:25-129:
``element_heat_from_losses``, ``element_joule_heat_w``, ``via_heat_sources``
and ``via_joule_heat_sources`` with the reference's arguments and messages.

The electrical mesh has copper layers; the thermal mesh has element slabs,
copper and dielectric.  ``layer_slabs[L]`` is the thermal slab that holds
electrical layer ``L``.  The two meshes share the in-plane grid, so element
losses map one to one onto thermal elements, and a via's loss goes half to
each end, spread over the two node faces that bound the end's copper slab.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np

from electrical.matrix_free_mpir_fem.pcb import PCBConductionProblem, PCBConductionSolution, ViaConnection

from .boundaries import HeatSource
from .mesh import LayeredThermalMesh


@dataclass(frozen=True)
class LayerSlabMap:
    """Thermal slab of every electrical copper layer, checked against both meshes."""

    slab_of_layer: np.ndarray
    element_grid_shape: tuple[int, int, int]

    @classmethod
    def build(cls, layer_slabs: Sequence[int], layers: int, thermal_mesh: LayeredThermalMesh) -> "LayerSlabMap":
        slab_of_layer = np.array([int(s) for s in layer_slabs], dtype=np.int64)
        slabs = thermal_mesh.element_grid_shape[0]
        if slab_of_layer.size != layers:
            raise ValueError(f"layer_slabs has {slab_of_layer.size} entries for {layers} electrical layers")
        if ((slab_of_layer < 0) | (slab_of_layer >= slabs)).any():
            raise ValueError(f"layer_slabs must lie in [0, {slabs - 1}]")
        return cls(slab_of_layer, tuple(thermal_mesh.element_grid_shape))

    def place(self, per_layer: np.ndarray) -> np.ndarray:
        """``(layers, rows, cols)`` values summed into ``(slabs, rows, cols)``;
        layers sharing a slab add in layer order."""

        placed = np.zeros(self.element_grid_shape)
        np.add.at(placed, self.slab_of_layer, per_layer)
        return placed


def element_heat_from_losses(element_loss_w, thermal_mesh: LayeredThermalMesh,
                             layer_slabs: Sequence[int]) -> np.ndarray:
    """Copper element losses ``(layers, rows, cols)`` placed into the thermal
    element grid ``(slabs, rows, cols)``."""

    loss = np.asarray(element_loss_w, dtype=np.float64)
    if loss.ndim != 3 or loss.shape[1:] != tuple(thermal_mesh.element_grid_shape[1:]):
        raise ValueError("the electrical element grid must match the thermal (rows, cols) footprint")
    return LayerSlabMap.build(layer_slabs, loss.shape[0], thermal_mesh).place(loss)


def element_joule_heat_w(solution: PCBConductionSolution, thermal_mesh: LayeredThermalMesh,
                         layer_slabs: Sequence[int]) -> np.ndarray:
    """``ThermalConductionProblem.element_heat_w`` of a DC solution's element
    losses; their sum is the solution's element loss, and via losses are
    nodal sources instead (``via_joule_heat_sources``)."""

    return element_heat_from_losses(solution.element_joule_loss_w, thermal_mesh, layer_slabs)


def via_heat_sources(vias: Sequence[ViaConnection], via_loss_w, layer_count: int, thermal_mesh: LayeredThermalMesh,
                     layer_slabs: Sequence[int]) -> tuple[HeatSource, ...]:
    """Half of each via's loss at each end; an end on electrical layer ``L``
    heats the two node faces bounding slab ``layer_slabs[L]``, so the heat
    enters the copper slab rather than one of its surfaces.  Sources are
    named ``via<i>-lower`` and ``via<i>-upper``."""

    slab_map = LayerSlabMap.build(layer_slabs, layer_count, thermal_mesh)
    losses = np.asarray(via_loss_w, dtype=np.float64)
    if losses.shape != (len(vias),):
        raise ValueError("via_loss_w must hold one value per via")
    if not vias:
        return ()
    # ends[i, e] = (layer, row, col) of via i's lower (e = 0) or upper (e = 1) end
    ends = np.array([(via.lower, via.upper) for via in vias], dtype=np.int64).reshape(len(vias), 2, 3)
    footprint = np.array(thermal_mesh.element_grid_shape[1:])
    outside = (ends[..., 1:] > footprint).any(axis=-1).reshape(-1)
    if outside.any():
        first = int(np.flatnonzero(outside)[0])
        raise ValueError(f"via node {tuple(int(v) for v in ends.reshape(-1, 3)[first])!r} lies outside the thermal footprint")
    slab = slab_map.slab_of_layer[ends[..., 0]]
    half = 0.5 * losses
    return tuple(
        HeatSource(((int(slab[i, e]), int(ends[i, e, 1]), int(ends[i, e, 2])),
                    (int(slab[i, e]) + 1, int(ends[i, e, 1]), int(ends[i, e, 2]))),
                   float(half[i]), f"via{i}-{name}")
        for i in range(len(vias)) for e, name in enumerate(("lower", "upper"))
    )


def via_joule_heat_sources(problem: PCBConductionProblem, solution: PCBConductionSolution,
                           thermal_mesh: LayeredThermalMesh, layer_slabs: Sequence[int]) -> tuple[HeatSource, ...]:
    """Nodal heat of every via of a current- or voltage-driven DC solve."""

    return via_heat_sources(problem.vias, solution.via_joule_loss_w, problem.mesh.node_shape[0], thermal_mesh,
                            layer_slabs)


__all__ = ["element_heat_from_losses", "element_joule_heat_w", "via_heat_sources", "via_joule_heat_sources"]
