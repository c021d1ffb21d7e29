"""Joule loss of a layered-PCB DC solve as the load of a thermal solve
(reference: src/thermal/matrix_free_mpir_fem/coupling.py).

The electrical mesh has copper layers; the thermal mesh has element slabs,
copper and dielectric.  ``layer_slabs[L]`` is the thermal slab that holds
electrical layer ``L``.  The two meshes share the in-plane grid, so element
losses map one to one onto thermal elements, and a via's loss goes half to
each end, spread over the two node faces that bound the end's copper slab.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from electrical.matrix_free_mpir_fem.pcb import PCBConductionProblem, PCBConductionSolution, ViaConnection

from .boundaries import HeatSource
from .mesh import LayeredThermalMesh


def _layer_slabs(layer_slabs: Sequence[int], layers: int, slabs: int) -> tuple[int, ...]:
    mapping = tuple(int(s) for s in layer_slabs)
    if len(mapping) != layers:
        raise ValueError(f"layer_slabs has {len(mapping)} entries for {layers} electrical layers")
    if any(not 0 <= s < slabs for s in mapping):
        raise ValueError(f"layer_slabs must lie in [0, {slabs - 1}]")
    return mapping


def element_heat_from_losses(element_loss_w, thermal_mesh: LayeredThermalMesh,
                             layer_slabs: Sequence[int]) -> np.ndarray:
    """Copper element losses ``(layers, rows, cols)`` placed into the thermal
    element grid ``(slabs, rows, cols)`` (coupling.py:38-57)."""

    loss = np.asarray(element_loss_w, dtype=np.float64)
    slabs, rows, cols = thermal_mesh.element_grid_shape
    if loss.ndim != 3 or loss.shape[1:] != (rows, cols):
        raise ValueError("the electrical element grid must match the thermal (rows, cols) footprint")
    heat = np.zeros((slabs, rows, cols))
    for layer, slab in enumerate(_layer_slabs(layer_slabs, loss.shape[0], slabs)):
        heat[slab] += loss[layer]
    return heat


def element_joule_heat_w(solution: PCBConductionSolution, thermal_mesh: LayeredThermalMesh,
                         layer_slabs: Sequence[int]) -> np.ndarray:
    """``ThermalConductionProblem.element_heat_w`` of a DC solution's element
    losses (coupling.py:60-73); via losses are nodal sources instead."""

    return element_heat_from_losses(solution.element_joule_loss_w, thermal_mesh, layer_slabs)


def via_heat_sources(vias: Sequence[ViaConnection], via_loss_w, layer_count: int, thermal_mesh: LayeredThermalMesh,
                     layer_slabs: Sequence[int]) -> tuple[HeatSource, ...]:
    """Half of each via's loss at each end, over the two node faces bounding
    that end's copper slab (coupling.py:76-111)."""

    slabs, rows, cols = thermal_mesh.element_grid_shape
    mapping = _layer_slabs(layer_slabs, layer_count, slabs)
    losses = np.asarray(via_loss_w, dtype=np.float64)
    if losses.shape != (len(vias),):
        raise ValueError("via_loss_w must hold one value per via")
    out = []
    for index, (via, loss) in enumerate(zip(vias, losses)):
        for end, (layer, row, col) in (("lower", via.lower), ("upper", via.upper)):
            if row > rows or col > cols:
                raise ValueError(f"via node {(layer, row, col)!r} lies outside the thermal footprint")
            slab = mapping[layer]
            out.append(HeatSource(((slab, row, col), (slab + 1, row, col)), 0.5 * float(loss), f"via{index}-{end}"))
    return tuple(out)


def via_joule_heat_sources(problem: PCBConductionProblem, solution: PCBConductionSolution,
                           thermal_mesh: LayeredThermalMesh, layer_slabs: Sequence[int]) -> tuple[HeatSource, ...]:
    """Nodal heat of every via of a DC solve (coupling.py:114-129)."""

    return via_heat_sources(problem.vias, solution.via_joule_loss_w, problem.mesh.node_shape[0], thermal_mesh,
                            layer_slabs)


__all__ = ["element_heat_from_losses", "element_joule_heat_w", "via_heat_sources", "via_joule_heat_sources"]
