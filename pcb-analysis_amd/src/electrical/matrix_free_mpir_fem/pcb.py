"""Layered-PCB DC conduction on the HIP device (reference:
src/electrical/matrix_free_mpir_fem/pcb.py).

Each copper layer is a sheet of bilinear (Q1) quadrilaterals with
``K_e = c_x U_x + c_y U_y``, ``c_x = sigma t h_y / h_x`` and
``c_y = sigma t h_x / h_y``; plated vertical connections are conductance
links between layer nodes.  Terminals spread a current over a pad's nodes or
hold a pad at a potential; a current-driven problem must balance and fixes
its gauge with ``reference_node``.  The operator lives on the device
(``pcbhip_sheet_operator_create``) and runs the same MPIR solve, batch and
server paths as the hexahedral operator.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass
from typing import Literal, Sequence

import numpy as np

from pcbhip import _capi, serving

from .grid import check_pitch_axis
from .runtime import make_float32_runtime
from .solver import MPIRConfig, MPIRResult, default_config, solve_mpir
from .two_level import AggregationCoarseCorrection

Node = tuple[int, int, int]
Preconditioner = Literal["two-level", "jacobi"]

COPPER_CONDUCTIVITY_S_PER_M = 1.0 / 1.724e-8


def _node(value) -> Node:
    node = tuple(int(i) for i in value)
    if len(node) != 3:
        raise ValueError("nodes are (layer, row, column) triples")
    return node


@dataclass(frozen=True)
class ViaConnection:
    """A resistive link between two nodes of different layers (pcb.py:39-64)."""

    lower: Node
    upper: Node
    resistance_ohm: float

    def __post_init__(self) -> None:
        lower, upper = _node(self.lower), _node(self.upper)
        if lower == upper:
            raise ValueError("a via must connect two different nodes")
        if lower[0] == upper[0]:
            raise ValueError("a via must connect different layers")
        if not (np.isfinite(self.resistance_ohm) and self.resistance_ohm > 0.0):
            raise ValueError("via resistance must be finite and positive")
        object.__setattr__(self, "lower", lower)
        object.__setattr__(self, "upper", upper)

    @property
    def conductance_s(self) -> float:
        return 1.0 / float(self.resistance_ohm)


@dataclass(frozen=True)
class CurrentTerminal:
    """A total current spread uniformly over a set of nodes (pcb.py:66-82)."""

    nodes: tuple[Node, ...]
    current_a: float
    name: str = "terminal"

    def __post_init__(self) -> None:
        nodes = tuple(_node(node) for node in self.nodes)
        if not nodes:
            raise ValueError("a current terminal needs at least one node")
        if len(set(nodes)) != len(nodes):
            raise ValueError("terminal nodes must be unique")
        if not np.isfinite(self.current_a):
            raise ValueError("terminal current must be finite")
        object.__setattr__(self, "nodes", nodes)


@dataclass(frozen=True)
class VoltageTerminal:
    """A set of nodes held at one potential (pcb.py:86-102)."""

    nodes: tuple[Node, ...]
    voltage_v: float
    name: str = "terminal"

    def __post_init__(self) -> None:
        nodes = tuple(_node(node) for node in self.nodes)
        if not nodes:
            raise ValueError("a voltage terminal needs at least one node")
        if len(set(nodes)) != len(nodes):
            raise ValueError("terminal nodes must be unique")
        if not np.isfinite(self.voltage_v):
            raise ValueError("terminal voltage must be finite")
        object.__setattr__(self, "nodes", nodes)


@dataclass(frozen=True)
class LayeredPCBMesh:
    """Structured Q1 sheets, one per copper layer (pcb.py:105-186).
    ``element_active`` is ``(layers, rows, cols)``; pitches are one value or
    one per column / row; conductivity is a scalar, one per layer or one per
    element."""

    element_active: np.ndarray
    layer_thickness_m: Sequence[float]
    pitch_x_m: float | Sequence[float] | np.ndarray
    pitch_y_m: float | Sequence[float] | np.ndarray
    conductivity_s_per_m: float | Sequence[float] | np.ndarray = COPPER_CONDUCTIVITY_S_PER_M

    def __post_init__(self) -> None:
        active = np.asarray(self.element_active, dtype=bool)
        if active.ndim != 3 or active.shape[1] < 1 or active.shape[2] < 1:
            raise ValueError("element_active must have shape (layers, rows, cols) with positive axes")
        thickness = np.asarray(self.layer_thickness_m, dtype=np.float64)
        if thickness.shape != (active.shape[0],):
            raise ValueError("layer_thickness_m must contain one value per layer")
        if not np.all(np.isfinite(thickness)) or np.any(thickness <= 0.0):
            raise ValueError("layer thicknesses must be finite and positive")
        sigma = np.asarray(self.conductivity_s_per_m, dtype=np.float64)
        if sigma.ndim == 0:
            sigma = np.full(active.shape, float(sigma))
        elif sigma.shape == (active.shape[0],):
            sigma = np.array(np.broadcast_to(sigma[:, None, None], active.shape))
        elif sigma.shape != active.shape:
            raise ValueError("conductivity must be scalar, per-layer, or match element_active")
        if not np.all(np.isfinite(sigma)) or np.any(sigma <= 0.0):
            raise ValueError("conductivity must be finite and positive")
        object.__setattr__(self, "pitch_x_m", check_pitch_axis(self.pitch_x_m, active.shape[2], "pitch_x_m"))
        object.__setattr__(self, "pitch_y_m", check_pitch_axis(self.pitch_y_m, active.shape[1], "pitch_y_m"))
        object.__setattr__(self, "element_active", active.copy())
        object.__setattr__(self, "layer_thickness_m", tuple(float(t) for t in thickness))
        object.__setattr__(self, "conductivity_s_per_m", sigma.copy())

    @property
    def node_shape(self) -> tuple[int, int, int]:
        layers, rows, cols = self.element_active.shape
        return layers, rows + 1, cols + 1

    @property
    def element_shape(self) -> tuple[int, int, int]:
        return tuple(int(v) for v in self.element_active.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.node_shape))

    @property
    def uniform_pitch(self) -> bool:
        return bool(np.all(self.pitch_x_m == self.pitch_x_m[0]) and np.all(self.pitch_y_m == self.pitch_y_m[0]))

    @property
    def cell_area_m2(self) -> np.ndarray:
        return self.pitch_y_m[:, None] * self.pitch_x_m[None, :]

    def element_coefficients(self) -> tuple[np.ndarray, np.ndarray]:
        """``(c_x, c_y)`` per element, zero on inactive elements (pcb.py:173-186)."""

        sheet = self.conductivity_s_per_m * np.asarray(self.layer_thickness_m)[:, None, None] * self.element_active
        hx = self.pitch_x_m[None, None, :]
        hy = self.pitch_y_m[None, :, None]
        return np.ascontiguousarray(sheet * hy / hx), np.ascontiguousarray(sheet * hx / hy)


@dataclass(frozen=True)
class PCBConductionProblem:
    """One conductor role and its drive (pcb.py:189-246): current terminals
    that balance, with ``reference_node`` fixing the gauge, or voltage
    terminals (then ``reference_node`` is None and currents need not
    balance).  A node belongs to at most one terminal."""

    mesh: LayeredPCBMesh
    terminals: tuple[CurrentTerminal, ...] = ()
    reference_node: Node | None = None
    vias: tuple[ViaConnection, ...] = ()
    voltage_terminals: tuple[VoltageTerminal, ...] = ()

    def __post_init__(self) -> None:
        terminals, voltage = tuple(self.terminals), tuple(self.voltage_terminals)
        if voltage:
            if self.reference_node is not None:
                raise ValueError("voltage terminals fix the potential; reference_node must be None")
        else:
            if not terminals:
                raise ValueError("a PCB conduction problem needs current or voltage terminals")
            if self.reference_node is None:
                raise ValueError("a current-driven problem needs a reference_node")
            total = float(sum(t.current_a for t in terminals))
            magnitude = float(sum(abs(t.current_a) for t in terminals))
            if abs(total) > 1.0e-12 * max(1.0, magnitude):
                raise ValueError(f"terminal currents must sum to zero, got {total:.6g} A")
        owner: dict[Node, str] = {}
        for terminal in (*terminals, *voltage):
            for node in terminal.nodes:
                if node in owner:
                    raise ValueError(f"node {node!r} belongs to terminals {owner[node]!r} and {terminal.name!r}")
                owner[node] = terminal.name
        object.__setattr__(self, "terminals", terminals)
        object.__setattr__(self, "voltage_terminals", voltage)
        object.__setattr__(self, "vias", tuple(self.vias))
        if self.reference_node is not None:
            object.__setattr__(self, "reference_node", _node(self.reference_node))

    @property
    def voltage_driven(self) -> bool:
        return bool(self.voltage_terminals)


_S1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
_M1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0


def unit_sheet_matrices() -> tuple[np.ndarray, np.ndarray]:
    """The ``(4, 4)`` unit sheet stiffness matrices ``U_x = M (x) S`` and
    ``U_y = S (x) M``, local order ``2 dy + dx`` (pcb.py:249-256)."""

    return np.kron(_M1, _S1), np.kron(_S1, _M1)


def _flat_index(node: Node, shape: tuple[int, int, int]) -> int:
    if len(node) != 3 or any(i < 0 for i in node):
        raise ValueError(f"invalid node {node!r}")
    try:
        return int(np.ravel_multi_index(node, shape))
    except ValueError as exc:
        raise ValueError(f"node {node!r} lies outside mesh shape {shape}") from exc


def _element_views(grid):
    return grid[:, :-1, :-1], grid[:, :-1, 1:], grid[:, 1:, :-1], grid[:, 1:, 1:]


class MatrixFreePCBOperator:
    """The layered-sheet conductance operator on the device (pcb.py:275-418).

    Fixed nodes are the inactive nodes (no active element and no via), the
    ``reference_node`` and every ``dirichlet_nodes`` entry; the operator is
    the identity on fixed rows and ignores fixed columns, so it stays SPD.
    ``preconditioner="two-level"`` (default) adds the patch-constant coarse
    correction of :mod:`.two_level`, each copper layer a patch layer of its
    own (pcb.py:405-418)."""

    def __init__(self, mesh: LayeredPCBMesh, *, reference_node: Node | None = None,
                 dirichlet_nodes: Sequence[Node] = (), vias: Sequence[ViaConnection] = (), runtime=None,
                 backend: str | None = None, device_id: int = 0, preconditioner: Preconditioner = "two-level",
                 coarse_block_nodes: int | None = None, native: bool | None = None) -> None:
        if runtime is not None and backend is not None:
            raise ValueError("pass either runtime or backend, not both")
        if preconditioner not in ("two-level", "jacobi"):
            raise ValueError("preconditioner must be 'two-level' or 'jacobi'")
        if native is False:
            raise ValueError("native=False asks for the NumPy reference path, which this package does not ship")
        if reference_node is None and not dirichlet_nodes:
            raise ValueError("pass a reference_node or dirichlet_nodes to fix the potential")
        self.high_dtype = np.float64
        self.inner_solver = "pcg"
        self.preconditioner = preconditioner
        self.coarse_correction: AggregationCoarseCorrection | None = None
        self.mesh = mesh
        self.size = mesh.size
        self.runtime = runtime or make_float32_runtime(backend or "hip", device_id=device_id)
        self.vias = tuple(vias)
        self.reference_node = None if reference_node is None else _node(reference_node)
        shape = mesh.node_shape
        active_nodes = np.zeros(shape, dtype=bool)
        for view in _element_views(active_nodes):
            view |= mesh.element_active
        via_a = np.asarray([_flat_index(v.lower, shape) for v in self.vias], dtype=np.int64)
        via_b = np.asarray([_flat_index(v.upper, shape) for v in self.vias], dtype=np.int64)
        via_g = np.asarray([v.conductance_s for v in self.vias], dtype=np.float64)
        active_nodes.flat[via_a] = True
        active_nodes.flat[via_b] = True
        fixed = ~active_nodes
        if self.reference_node is not None:
            index = _flat_index(self.reference_node, shape)
            if not active_nodes.flat[index]:
                raise ValueError("reference_node must lie on active copper or a via endpoint")
            fixed.flat[index] = True
        dirichlet = np.asarray([_flat_index(_node(n), shape) for n in dirichlet_nodes], dtype=np.int64)
        if dirichlet.size and not np.all(active_nodes.flat[dirichlet]):
            raise ValueError("dirichlet_nodes must lie on active copper or via endpoints")
        fixed.flat[dirichlet] = True
        self.active_nodes = active_nodes
        self.free_nodes = ~fixed
        self._via_a, self._via_b, self._via_g = via_a, via_b, via_g
        self._unit = np.stack(unit_sheet_matrices())
        self._coef = np.stack(mesh.element_coefficients())
        try:
            self._device = _capi.Operator.sheet(self.runtime.device_id, mesh.element_shape, self._coef,
                                                self.free_nodes.reshape(-1), via_a, via_b, via_g)
        except ValueError as exc:
            if "positive" in str(exc):
                raise ValueError("every free node must have positive conductivity coupling") from exc
            raise
        self.low_operator_backend = "hip-sheet-q1-fp32"
        self.high_operator_backend = "hip-sheet-q1-fp64"
        if preconditioner == "two-level":
            self.coarse_correction = AggregationCoarseCorrection(
                node_shape=shape, free_nodes=self.free_nodes, diagonal_high=self._device.diagonal(),
                apply_high=self.apply_high, block=coarse_block_nodes, layer_colours=shape[0],
                galerkin=self._device.galerkin_f64, invert=self._device.set_coarse_matrix,
            )

    @property
    def device_operator(self) -> _capi.Operator:
        return self._device

    def _flat(self, values) -> np.ndarray:
        flat = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if flat.size != self.size:
            raise ValueError(f"vector has size {flat.size}, expected {self.size}")
        return flat

    def apply_high(self, vector) -> np.ndarray:
        return self._device.apply_f64(self._flat(vector))

    def apply_low(self, vector):
        """``A x`` in FP32: device vector to device vector, or host to host."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.apply_f32_vec(vector)
        return self._device.apply_f32(np.asarray(vector, dtype=np.float32).reshape(-1))

    def apply_full_high(self, vector) -> np.ndarray:
        """``K v`` with no fixed rows (pcb.py:719-738): entry n is the current
        leaving node n into the copper and vias."""

        return self._device.stiffness_f64(self._flat(vector))

    def diagonal_low(self) -> np.ndarray:
        return self._device.diagonal().astype(np.float32)

    def precondition_low(self, vector) -> np.ndarray:
        """``M^-1 r`` in FP32 on the device (pcb.py:712-717); device vector
        in, device vector out, or host array in and out."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.precondition_f32_vec(vector)
        return self._device.precondition_f32(np.asarray(vector, dtype=np.float32).reshape(-1))

    def native_solve_mpir(self, rhs, config: MPIRConfig, initial_guess=None):
        x, r = serving.solve_mpir_routed(self._device, rhs, initial_guess, config.to_c())
        return (x, r["converged"], r["outer_iterations"], r["inner_iterations"], r["relative_residual"],
                r["high_operator_applications"], r["low_operator_applications"], r["history"])

    def dirichlet_potential(self, voltage_terminals: Sequence[VoltageTerminal] = ()) -> np.ndarray:
        """Each voltage terminal's value on its nodes, zero elsewhere (pcb.py:740-759)."""

        potential = np.zeros(self.size)
        for terminal in voltage_terminals:
            for node in terminal.nodes:
                index = _flat_index(node, self.mesh.node_shape)
                if self.free_nodes.flat[index] or not self.active_nodes.flat[index]:
                    raise ValueError(f"voltage terminal {terminal.name!r} node {node!r} is not a "
                                     "Dirichlet node of this operator")
                potential[index] = float(terminal.voltage_v)
        return potential

    def terminal_currents(self, potential_v, voltage_terminals: Sequence[VoltageTerminal]) -> np.ndarray:
        """Current each voltage terminal drives into the copper (pcb.py:761-780)."""

        injected = self.apply_full_high(np.nan_to_num(np.asarray(potential_v, dtype=np.float64), nan=0.0))
        shape = self.mesh.node_shape
        return np.asarray([float(np.sum(injected[[_flat_index(n, shape) for n in t.nodes]]))
                           for t in voltage_terminals])

    def build_rhs(self, terminals: Sequence[CurrentTerminal],
                  voltage_terminals: Sequence[VoltageTerminal] = ()) -> np.ndarray:
        """Terminal shares on free rows, the gauge equation on fixed rows,
        lifting ``f - K g`` for voltage terminals (pcb.py:782-819)."""

        rhs = np.zeros(self.size)
        shape = self.mesh.node_shape
        for terminal in terminals:
            share = float(terminal.current_a) / len(terminal.nodes)
            for node in terminal.nodes:
                index = _flat_index(node, shape)
                if not self.active_nodes.flat[index]:
                    raise ValueError(f"terminal {terminal.name!r} contains inactive node {node!r}")
                rhs[index] += share
        free = self.free_nodes.reshape(-1)
        rhs[~free] = 0.0
        if voltage_terminals:
            fixed = self.dirichlet_potential(voltage_terminals)
            rhs = np.where(free, rhs - self.apply_full_high(fixed), fixed)
        return rhs

    def element_electric_field(self, potential_v) -> np.ndarray:
        """``E = -grad V`` at element centres, ``(layers, rows, cols, 2)`` (pcb.py:821-833)."""

        v00, v01, v10, v11 = _element_views(np.asarray(potential_v, dtype=np.float64).reshape(self.mesh.node_shape))
        ex = -((v01 + v11) - (v00 + v10)) / (2.0 * self.mesh.pitch_x_m[None, None, :])
        ey = -((v10 + v11) - (v00 + v01)) / (2.0 * self.mesh.pitch_y_m[None, :, None])
        return np.where(self.mesh.element_active[..., None], np.stack((ex, ey), axis=-1), 0.0)

    def via_currents(self, potential_v) -> np.ndarray:
        flat = np.asarray(potential_v, dtype=np.float64).reshape(-1)
        return self._via_g * (flat[self._via_a] - flat[self._via_b])

    def element_joule_loss(self, potential_v) -> np.ndarray:
        """The element quadratic form ``c_x v^T U_x v + c_y v^T U_y v`` (pcb.py:843-860)."""

        v = np.stack(_element_views(np.asarray(potential_v, dtype=np.float64).reshape(self.mesh.node_shape)), axis=-1)
        loss = np.zeros(self.mesh.element_shape)
        for k in range(2):
            loss += self._coef[k] * np.einsum("...i,ij,...j->...", v, self._unit[k], v, optimize=True)
        return loss

    def via_joule_loss(self, potential_v) -> np.ndarray:
        flat = np.asarray(potential_v, dtype=np.float64).reshape(-1)
        drop = flat[self._via_a] - flat[self._via_b]
        return self._via_g * drop * drop

    def joule_loss(self, potential_v) -> float:
        return float(np.sum(self.element_joule_loss(potential_v)) + np.sum(self.via_joule_loss(potential_v)))


@dataclass(frozen=True)
class PCBConductionSolution:
    potential_v: np.ndarray
    current_density_a_per_m2: np.ndarray
    via_current_a: np.ndarray
    joule_loss_w: float
    element_joule_loss_w: np.ndarray
    via_joule_loss_w: np.ndarray
    max_current_density_a_per_m2: float
    solve: MPIRResult
    voltage_terminal_current_a: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))


def solve_pcb_dc(problem: PCBConductionProblem, *, config: MPIRConfig | None = None, runtime=None,
                 backend: str | None = None, device_id: int = 0, initial_potential_v=None,
                 preconditioner: Preconditioner = "two-level", coarse_block_nodes: int | None = None,
                 native: bool | None = None) -> PCBConductionSolution:
    """One layered-PCB DC solve with MPIR on the HIP device (pcb.py:896-978).
    ``initial_potential_v`` warm-starts the refinement (NaN entries count as
    zero)."""

    operator = MatrixFreePCBOperator(
        problem.mesh, reference_node=problem.reference_node,
        dirichlet_nodes=tuple(n for t in problem.voltage_terminals for n in t.nodes), vias=problem.vias,
        runtime=runtime, backend=backend, device_id=device_id, preconditioner=preconditioner,
        coarse_block_nodes=coarse_block_nodes, native=native,
    )
    rhs = operator.build_rhs(problem.terminals, problem.voltage_terminals)
    fixed = operator.dirichlet_potential(problem.voltage_terminals)
    initial = None
    if initial_potential_v is not None:
        initial = np.nan_to_num(np.asarray(initial_potential_v, dtype=np.float64).reshape(-1), nan=0.0)
        if initial.size != operator.size:
            raise ValueError("initial_potential_v must hold one value per node")
        initial = np.where(operator.free_nodes.reshape(-1), initial, fixed)
    elif problem.voltage_terminals:
        initial = fixed
    result = solve_mpir(operator, rhs, config=config or default_config(operator.preconditioner), initial_guess=initial)
    potential = result.solution.reshape(problem.mesh.node_shape)
    field = operator.element_electric_field(potential)
    density = problem.mesh.conductivity_s_per_m[..., None] * field
    magnitude = np.linalg.norm(density, axis=-1)[problem.mesh.element_active]
    return PCBConductionSolution(
        potential_v=np.where(operator.active_nodes, potential, np.nan),
        current_density_a_per_m2=density,
        via_current_a=operator.via_currents(potential),
        joule_loss_w=operator.joule_loss(potential),
        element_joule_loss_w=operator.element_joule_loss(potential),
        via_joule_loss_w=operator.via_joule_loss(potential),
        max_current_density_a_per_m2=float(np.max(magnitude)) if magnitude.size else 0.0,
        solve=result,
        voltage_terminal_current_a=operator.terminal_currents(potential, problem.voltage_terminals),
    )


__all__ = [
    "COPPER_CONDUCTIVITY_S_PER_M",
    "CurrentTerminal",
    "LayeredPCBMesh",
    "MatrixFreePCBOperator",
    "PCBConductionProblem",
    "PCBConductionSolution",
    "ViaConnection",
    "VoltageTerminal",
    "solve_pcb_dc",
    "unit_sheet_matrices",
]
