"""Staggered electro-thermal coupling with temperature-dependent copper, on
the HIP device (reference: src/multiphysics/staggered_coupling/
electro_thermal.py).

Copper resistivity rises with temperature, ``sigma(T) = sigma_ref / (1 +
alpha (T - T_ref))``.  The loop alternates a layered-PCB DC solve with the
current conductivity, a thermal solve loaded with that solve's Joule loss,
and a conductivity update from the element temperatures, until temperature
and loss settle.  Each solve is warm-started from the previous iterate, and
the temperature update is relaxed with Aitken's delta-squared factor.

A voltage drive (two or more voltage terminals) gives negative feedback: the
loss V^2 / R(T) falls as the copper heats.  A current drive imposes every
copper current, its loss I^2 R(T) rises with temperature and the loop can run
away; such scenarios still run, with a DeprecationWarning, as in the
reference.

Both solves of every iteration are device MPIR solves; the loop itself is
host work of a few field reductions per iteration.  ``run_electro_thermal_many``
runs many loops at once as a served workload: their DC and thermal solves,
each waiting on the previous one of its loop, stay in flight together on one
continuous-batching server.
"""

from __future__ import annotations

import dataclasses
import warnings
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from electrical.matrix_free_mpir_fem import (
    MPIRConfig,
    PCBConductionProblem,
    PCBConductionSolution,
    ViaConnection,
    solve_pcb_dc,
)
from pcbhip import serving
from thermal.matrix_free_mpir_fem import (
    ConvectionBoundary,
    ExposedFaceRadiation,
    HeatSource,
    LayeredThermalMesh,
    RadiationBoundary,
    ThermalConductionProblem,
    ThermalConductionSolution,
    element_joule_heat_w,
    solve_thermal_conduction,
    via_joule_heat_sources,
)

COPPER_TEMPERATURE_COEFFICIENT_PER_K = 3.93e-3

CURRENT_DRIVEN_DEPRECATION = (
    "current-driven electro-thermal coupling is deprecated: with fewer than two "
    "voltage terminals the current terminals fix every copper current, so the "
    "loss I^2 R(T) rises with temperature and the board can run away; hold at "
    "least two pads with VoltageTerminal (PCBConductionProblem.voltage_terminals)"
)


@dataclass(frozen=True)
class ElectroThermalScenario:
    """One board: its electrical drive, its thermal mesh and environment, and
    the resistivity law (electro_thermal.py:68-124).  ``layer_slabs[L]`` is
    the thermal slab holding electrical layer ``L``."""

    electrical: PCBConductionProblem
    thermal_mesh: LayeredThermalMesh
    layer_slabs: tuple[int, ...]
    convection: tuple[ConvectionBoundary, ...] = ()
    fixed_temperature_mask: np.ndarray | None = None
    fixed_temperature_k: float | np.ndarray | None = None
    extra_heat_sources: tuple[HeatSource, ...] = ()
    extra_element_heat_w: np.ndarray | None = None
    conductivity_reference_temperature_k: float = 293.15
    temperature_coefficient_per_k: float = COPPER_TEMPERATURE_COEFFICIENT_PER_K
    radiation: tuple[RadiationBoundary | ExposedFaceRadiation, ...] = ()

    def __post_init__(self) -> None:
        layers, rows, cols = self.electrical.mesh.element_active.shape
        slabs, t_rows, t_cols = self.thermal_mesh.element_grid_shape
        if (rows, cols) != (t_rows, t_cols):
            raise ValueError("electrical and thermal meshes must share (rows, cols)")
        same_x = np.allclose(self.electrical.mesh.pitch_x_m, self.thermal_mesh.pitch_x_m)
        same_y = np.allclose(self.electrical.mesh.pitch_y_m, self.thermal_mesh.pitch_y_m)
        if not (same_x and same_y):
            raise ValueError("electrical and thermal meshes must share the same grid lines")
        mapping = tuple(int(s) for s in self.layer_slabs)
        if len(mapping) != layers or any(not 0 <= s < slabs for s in mapping):
            raise ValueError("layer_slabs must name one thermal slab per electrical layer")
        alpha = self.temperature_coefficient_per_k
        if not np.isfinite(alpha) or alpha < 0.0:
            raise ValueError("temperature_coefficient_per_k must be finite and non-negative")
        object.__setattr__(self, "layer_slabs", mapping)
        object.__setattr__(self, "convection", tuple(self.convection))
        object.__setattr__(self, "extra_heat_sources", tuple(self.extra_heat_sources))
        object.__setattr__(self, "radiation", tuple(self.radiation))
        # The thermal boundary conditions are checked once, before any heat.
        ThermalConductionProblem(
            self.thermal_mesh, convection=self.convection, fixed_temperature_mask=self.fixed_temperature_mask,
            fixed_temperature_k=self.fixed_temperature_k, heat_sources=self.extra_heat_sources,
            element_heat_w=self.extra_element_heat_w, radiation=self.radiation,
        )
        if len(self.electrical.voltage_terminals) < 2:
            warnings.warn(CURRENT_DRIVEN_DEPRECATION, DeprecationWarning, stacklevel=3)


@dataclass(frozen=True)
class CouplingConfig:
    """Fixed-point limits of a staggered coupling (electro_thermal.py:127-150)."""

    max_iterations: int = 20
    temperature_tolerance_k: float = 1.0e-3
    relative_loss_tolerance: float = 1.0e-6
    relaxation: float = 1.0
    aitken: bool = True
    max_relaxation: float = 4.0
    electrical: MPIRConfig | None = None
    thermal: MPIRConfig | None = None

    def __post_init__(self) -> None:
        if self.max_iterations < 1:
            raise ValueError("max_iterations must be positive")
        if self.temperature_tolerance_k <= 0.0 or self.relative_loss_tolerance <= 0.0:
            raise ValueError("tolerances must be positive")
        if not 0.0 < self.relaxation <= self.max_relaxation:
            raise ValueError("relaxation must lie in (0, max_relaxation]")
        if self.max_relaxation < 1.0:
            raise ValueError("max_relaxation must be at least one")


@dataclass(frozen=True)
class CouplingStep:
    iteration: int
    joule_loss_w: float
    max_temperature_k: float
    temperature_change_k: float
    relative_loss_change: float
    relaxation: float
    electrical_inner_iterations: int
    thermal_inner_iterations: int


@dataclass(frozen=True)
class ElectroThermalResult:
    electrical: PCBConductionSolution
    thermal: ThermalConductionSolution
    conductivity_s_per_m: np.ndarray
    via_resistance_ohm: np.ndarray
    element_temperature_k: np.ndarray
    converged: bool
    iterations: int
    history: tuple[CouplingStep, ...]
    cold_joule_loss_w: float

    @property
    def loss_increase_ratio(self) -> float:
        if not self.cold_joule_loss_w:
            return float("nan")
        return self.electrical.joule_loss_w / self.cold_joule_loss_w


def slab_element_temperature_k(temperature_k, thermal_mesh: LayeredThermalMesh, slab: int) -> np.ndarray:
    """Mean of the eight corner temperatures of every element of one slab."""

    grid = np.asarray(temperature_k, dtype=np.float64).reshape(thermal_mesh.node_shape)
    total = np.zeros((grid.shape[1] - 1, grid.shape[2] - 1))
    for face in (grid[slab], grid[slab + 1]):
        total += face[:-1, :-1] + face[:-1, 1:] + face[1:, :-1] + face[1:, 1:]
    return 0.125 * total


def electrical_layer_temperature_k(temperature_k, thermal_mesh: LayeredThermalMesh,
                                   layer_slabs: Sequence[int]) -> np.ndarray:
    """Element temperature of every electrical layer, ``(layers, rows, cols)``."""

    return np.stack([slab_element_temperature_k(temperature_k, thermal_mesh, s) for s in layer_slabs])


def via_temperature_k(temperature_k, thermal_mesh: LayeredThermalMesh, layer_slabs: Sequence[int],
                      vias: Sequence[ViaConnection]) -> np.ndarray:
    """Mean over a via's two ends, each end the mean of the two node faces
    bounding its copper slab."""

    grid = np.asarray(temperature_k, dtype=np.float64).reshape(thermal_mesh.node_shape)
    out = np.empty(len(vias))
    for k, via in enumerate(vias):
        ends = []
        for layer, row, col in (via.lower, via.upper):
            slab = layer_slabs[layer]
            ends.append(0.5 * (grid[slab, row, col] + grid[slab + 1, row, col]))
        out[k] = 0.5 * (ends[0] + ends[1])
    return out


def via_node_temperature_k(temperature_k, thermal_mesh: LayeredThermalMesh, layer_slabs: Sequence[int],
                           problem: PCBConductionProblem) -> np.ndarray:
    return via_temperature_k(temperature_k, thermal_mesh, layer_slabs, problem.vias)


def conductivity_at_temperature(reference_conductivity, temperature_k, *, reference_temperature_k: float,
                                coefficient_per_k: float) -> np.ndarray:
    """``sigma_ref / (1 + alpha (T - T_ref))``; the linear law must stay positive."""

    factor = 1.0 + coefficient_per_k * (np.asarray(temperature_k) - reference_temperature_k)
    if np.any(factor <= 0.0):
        raise ValueError("the linear resistivity model left the valid range (σ ≤ 0)")
    return np.asarray(reference_conductivity) / factor


def thermal_problem_with_joule_heat(scenario: ElectroThermalScenario,
                                    electrical: PCBConductionSolution) -> ThermalConductionProblem:
    """The scenario's thermal problem loaded with a DC solution's element and via losses."""

    heat = element_joule_heat_w(electrical, scenario.thermal_mesh, scenario.layer_slabs)
    if scenario.extra_element_heat_w is not None:
        heat = heat + np.asarray(scenario.extra_element_heat_w, dtype=np.float64)
    sources = via_joule_heat_sources(scenario.electrical, electrical, scenario.thermal_mesh, scenario.layer_slabs)
    return ThermalConductionProblem(
        scenario.thermal_mesh, convection=scenario.convection,
        fixed_temperature_mask=scenario.fixed_temperature_mask, fixed_temperature_k=scenario.fixed_temperature_k,
        heat_sources=sources + scenario.extra_heat_sources, element_heat_w=heat, radiation=scenario.radiation,
    )


def heated_electrical_problem(scenario: ElectroThermalScenario, temperature_k) -> PCBConductionProblem:
    """The DC problem with copper and via barrels at the board's temperatures."""

    law = dict(reference_temperature_k=scenario.conductivity_reference_temperature_k,
               coefficient_per_k=scenario.temperature_coefficient_per_k)
    layer_t = electrical_layer_temperature_k(temperature_k, scenario.thermal_mesh, scenario.layer_slabs)
    sigma = conductivity_at_temperature(scenario.electrical.mesh.conductivity_s_per_m, layer_t, **law)
    via_t = via_node_temperature_k(temperature_k, scenario.thermal_mesh, scenario.layer_slabs, scenario.electrical)
    via_factor = 1.0 + scenario.temperature_coefficient_per_k * (via_t - scenario.conductivity_reference_temperature_k)
    vias = tuple(dataclasses.replace(v, resistance_ohm=v.resistance_ohm * float(f))
                 for v, f in zip(scenario.electrical.vias, via_factor))
    mesh = dataclasses.replace(scenario.electrical.mesh, conductivity_s_per_m=sigma)
    return dataclasses.replace(scenario.electrical, mesh=mesh, vias=vias)


class TemperatureFixedPoint:
    """Relaxed fixed-point update of a temperature field with Aitken's
    delta-squared factor (electro_thermal.py:295-345).

    With increments d_k = T_proposed - T_k, the factor is
    ``w_k = -w_{k-1} (d_{k-1} . (d_k - d_{k-1})) / |d_k - d_{k-1}|^2``,
    clipped to [0.05, max_relaxation]; an oscillating iteration gets w < 1, a
    slowly contracting one w > 1.  A relaxed step that would take the field
    below the temperature where the linear resistivity law reaches 0.1 % of
    its reference value falls back to the plain step."""

    def __init__(self, scenario: ElectroThermalScenario, config: CouplingConfig) -> None:
        self.config = config
        self.relaxation = config.relaxation
        self.temperature: np.ndarray | None = None
        self._last: np.ndarray | None = None
        alpha = scenario.temperature_coefficient_per_k
        self._floor = (scenario.conductivity_reference_temperature_k - (1.0 - 1.0e-3) / alpha) if alpha > 0.0 \
            else -np.inf

    def update(self, proposed) -> float:
        """Takes a proposed field; returns the largest change of the relaxed iterate."""

        proposed = np.asarray(proposed, dtype=np.float64)
        if self.temperature is None:
            self.temperature = proposed.copy()
            self._last = None
            return float("inf")
        step = proposed - self.temperature
        if self.config.aitken and self._last is not None:
            delta = (step - self._last).reshape(-1)
            norm = float(delta @ delta)
            if norm > 0.0:
                w = -self.relaxation * float(self._last.reshape(-1) @ delta) / norm
                self.relaxation = float(np.clip(w, 0.05, self.config.max_relaxation))
        candidate = self.temperature + self.relaxation * step
        if float(np.min(candidate)) < self._floor <= float(np.min(proposed)):
            self.relaxation = 1.0
            candidate = proposed.copy()
        change = float(np.max(np.abs(candidate - self.temperature)))
        self.temperature = candidate
        self._last = step
        return change


def run_electro_thermal(scenario: ElectroThermalScenario, *, config: CouplingConfig | None = None,
                        backend: str | None = None, device_id: int = 0,
                        native: bool | None = None) -> ElectroThermalResult:
    """Iterates DC and thermal device solves to a self-consistent
    ``sigma(T)`` state (electro_thermal.py:348-466, the solve-by-solve loop).
    ``native=False`` asks for the reference's NumPy path, which this package
    does not ship."""

    if native is False:
        raise ValueError("native=False asks for the NumPy reference path, which this package does not ship")
    config = config or CouplingConfig()
    problem = scenario.electrical
    fixed_point = TemperatureFixedPoint(scenario, config)
    history: list[CouplingStep] = []
    potential = None
    cold_loss = float("nan")
    previous_loss = float("nan")
    converged = False
    electrical = thermal = None
    for iteration in range(1, config.max_iterations + 1):
        electrical = solve_pcb_dc(problem, config=config.electrical, backend=backend, device_id=device_id,
                                  initial_potential_v=potential)
        potential = electrical.potential_v
        loss = electrical.joule_loss_w
        if iteration == 1:
            cold_loss = loss
        thermal = solve_thermal_conduction(thermal_problem_with_joule_heat(scenario, electrical),
                                           config=config.thermal, backend=backend, device_id=device_id,
                                           initial_temperature_k=fixed_point.temperature)
        # Nodes of no element are NaN; they hold no copper and enter the
        # fixed point at the coldest node's temperature.
        proposed = np.where(np.isfinite(thermal.temperature_k), thermal.temperature_k, thermal.min_temperature_k)
        change = fixed_point.update(proposed)
        temperature = fixed_point.temperature
        if np.isfinite(previous_loss) and loss:
            loss_change = abs(loss - previous_loss) / abs(loss)
        else:
            loss_change = float("inf")
        previous_loss = loss
        history.append(CouplingStep(
            iteration=iteration, joule_loss_w=loss, max_temperature_k=float(np.max(temperature)),
            temperature_change_k=change, relative_loss_change=loss_change, relaxation=fixed_point.relaxation,
            electrical_inner_iterations=electrical.solve.inner_iterations,
            thermal_inner_iterations=thermal.solve.inner_iterations,
        ))
        solved = electrical.solve.converged and thermal.solve.converged
        if scenario.temperature_coefficient_per_k == 0.0:
            converged = solved
            break
        if change <= config.temperature_tolerance_k and loss_change <= config.relative_loss_tolerance and solved:
            converged = True
            break
        problem = heated_electrical_problem(scenario, temperature)
    return ElectroThermalResult(
        electrical=electrical, thermal=thermal,
        conductivity_s_per_m=np.asarray(problem.mesh.conductivity_s_per_m, dtype=np.float64),
        via_resistance_ohm=np.asarray([v.resistance_ohm for v in problem.vias], dtype=np.float64),
        element_temperature_k=electrical_layer_temperature_k(temperature, scenario.thermal_mesh,
                                                             scenario.layer_slabs),
        converged=converged, iterations=len(history), history=tuple(history), cold_joule_loss_w=cold_loss,
    )


def run_electro_thermal_many(scenarios: Sequence[ElectroThermalScenario], *, config: CouplingConfig | None = None,
                             device_id: int = 0, max_slots: int = 256, memory_budget_bytes: float = 200e9,
                             on_done: Callable[[int], None] | None = None) -> list[ElectroThermalResult]:
    """Many electro-thermal loops at once: one host thread per loop, every DC
    and thermal solve of every loop continuously batched on one
    :class:`pcbhip.serving.ServedSolves` server.  The reference runs such
    loops one after another; each result here is bitwise what
    :func:`run_electro_thermal` gives the scenario alone.  ``on_done(i)`` is
    called from loop i's thread when it finishes."""

    scenarios = list(scenarios)
    if not scenarios:
        return []
    with serving.ServedSolves(device_id, max_slots=max_slots, memory_budget_bytes=memory_budget_bytes) as served:
        def one(item) -> ElectroThermalResult:
            index, scenario = item
            with serving.route_solves(served):
                result = run_electro_thermal(scenario, config=config, device_id=device_id)
            if on_done is not None:
                on_done(index)
            return result

        with ThreadPoolExecutor(max_workers=len(scenarios), thread_name_prefix="coupling") as pool:
            return list(pool.map(one, enumerate(scenarios)))
