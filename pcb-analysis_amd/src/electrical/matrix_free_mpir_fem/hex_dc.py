"""DC conduction on the hexahedral Q1 mesh, solved on the HIP device.

The reference's DC solver (src/electrical/matrix_free_mpir_fem/pcb.py) works
on layered 2-D copper sheets joined by via links; a volumetric conductor with
a dielectric (copper planes in a 1e-12 S/m laminate, say) needs the
hexahedral operator of the thermal solver with the electrical conductivity
in place of the thermal one.  This module is that front end.  It keeps the
reference's DC contract: current terminals spread a current uniformly over
their nodes (pcb.py:66-82, 799-811), a current-driven problem must balance
and fixes its gauge with ``reference_node`` whose equation is replaced by
``V = 0`` (pcb.py:367-372, 809-811), voltage terminals hold their nodes at a
potential through the lifting ``f - K g`` (pcb.py:812-819), and the Joule
loss is the exact element quadratic form (pcb.py:843-860).

Many excitations of one conductor (the N-port basis, config-style terminal
pairs) run as one batch of right-hand sides on one device-resident operator
(``solve_hex_dc_many``).
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from pcbhip import _capi, serving
from thermal.matrix_free_mpir_fem.mesh import LayeredThermalMesh, corner_views, flat_index, unit_hexahedron_matrices

from .pcb import CurrentTerminal, VoltageTerminal
from .runtime import make_float32_runtime
from .solver import MPIRConfig, MPIRResult, MPIRStep, jacobi_default_config, solve_mpir

Node = tuple[int, int, int]


def _check_terminals(terminals, voltage_terminals, reference_node) -> None:
    """The PCBConductionProblem contract (pcb.py:205-246)."""

    if voltage_terminals:
        if reference_node is not None:
            raise ValueError("voltage terminals fix the potential; reference_node must be None")
    else:
        if not terminals:
            raise ValueError("a conduction problem needs current or voltage terminals")
        if reference_node is None:
            raise ValueError("a current-driven problem needs a reference_node")
        total = float(sum(t.current_a for t in terminals))
        magnitude = float(sum(abs(t.current_a) for t in terminals))
        if abs(total) > 1.0e-12 * max(1.0, magnitude):
            raise ValueError(f"terminal currents must sum to zero, got {total:.6g} A")
    owner: dict[Node, str] = {}
    for terminal in (*terminals, *voltage_terminals):
        for node in terminal.nodes:
            if node in owner:
                raise ValueError(f"node {node!r} belongs to terminals {owner[node]!r} and {terminal.name!r}")
            owner[node] = terminal.name


@dataclass(frozen=True)
class HexConductionProblem:
    """A volumetric conductor on a hexahedral mesh and its drive.

    ``mesh.conductivity_w_per_m_k`` / ``through_plane_conductivity_w_per_m_k``
    hold the electrical conductivity in S/m (the mesh class is physics
    neutral)."""

    mesh: LayeredThermalMesh
    terminals: tuple[CurrentTerminal, ...] = ()
    reference_node: Node | None = None
    voltage_terminals: tuple[VoltageTerminal, ...] = ()

    def __post_init__(self) -> None:
        terminals, voltage = tuple(self.terminals), tuple(self.voltage_terminals)
        reference = None if self.reference_node is None else tuple(int(v) for v in self.reference_node)
        _check_terminals(terminals, voltage, reference)
        object.__setattr__(self, "terminals", terminals)
        object.__setattr__(self, "voltage_terminals", voltage)
        object.__setattr__(self, "reference_node", reference)


@dataclass(frozen=True)
class HexConductionSolution:
    potential_v: np.ndarray
    element_joule_loss_w: np.ndarray
    joule_loss_w: float
    solve: MPIRResult
    voltage_terminal_current_a: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))


class MatrixFreeHexDCOperator:
    """The hexahedral conduction operator with electrical conductivity, no
    Robin term and fixed gauge / voltage-terminal rows, resident on the device."""

    def __init__(self, mesh: LayeredThermalMesh, *, reference_node: Node | None = None,
                 dirichlet_nodes: Sequence[Node] = (), backend: str | None = None, device_id: int = 0,
                 runtime=None) -> None:
        self.mesh = mesh
        self.size = mesh.size
        self.high_dtype = np.float64
        self.inner_solver = "pcg"
        self.runtime = runtime or make_float32_runtime(backend or "hip", device_id=device_id)
        self.active_nodes = mesh.active_nodes
        fixed = ~self.active_nodes
        for node in tuple(dirichlet_nodes) + ((reference_node,) if reference_node is not None else ()):
            fixed.flat[flat_index(node, mesh.node_shape)] = True
        self.reference_node = reference_node
        self.fixed_nodes = fixed
        self.free_nodes = ~fixed
        self._unit = np.stack(unit_hexahedron_matrices())
        self._coef = np.stack(mesh.element_coefficients())
        self._device = _capi.Operator(self.runtime.device_id, mesh.element_grid_shape, self._coef,
                                      np.zeros(self.size), self.free_nodes.reshape(-1))

    @property
    def device_operator(self) -> _capi.Operator:
        return self._device

    def apply_high(self, vector) -> np.ndarray:
        return self._device.apply_f64(vector)

    def apply_low(self, vector):
        """``A x`` in FP32: device vector to device vector, or host to host."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.apply_f32_vec(vector)
        return self._device.apply_f32(vector)

    def precondition_low(self, vector):
        """``M^-1 r`` in FP32 with the operator's Jacobi diagonal."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.precondition_f32_vec(vector)
        return self._device.precondition_f32(np.asarray(vector, dtype=np.float32).reshape(-1))

    def diagonal_low(self) -> np.ndarray:
        return self._device.diagonal().astype(np.float32)

    def native_solve_mpir(self, rhs, config: MPIRConfig, initial_guess=None):
        x, r = serving.solve_mpir_routed(self._device, rhs, initial_guess, config.to_c())
        return (x, r["converged"], r["outer_iterations"], r["inner_iterations"], r["relative_residual"],
                r["high_operator_applications"], r["low_operator_applications"], r["history"])

    def dirichlet_potential(self, voltage_terminals: Sequence[VoltageTerminal]) -> np.ndarray:
        fixed = np.zeros(self.size)
        for terminal in voltage_terminals:
            for node in terminal.nodes:
                fixed[flat_index(node, self.mesh.node_shape)] = float(terminal.voltage_v)
        return fixed

    def build_rhs(self, terminals: Sequence[CurrentTerminal],
                  voltage_terminals: Sequence[VoltageTerminal] = ()) -> np.ndarray:
        """Terminal shares on free rows, zero on the gauge row, lifting for
        voltage terminals (pcb.py:782-819)."""

        rhs = np.zeros(self.size)
        active = self.active_nodes.reshape(-1)
        for terminal in terminals:
            share = float(terminal.current_a) / len(terminal.nodes)
            for node in terminal.nodes:
                index = flat_index(node, self.mesh.node_shape)
                if not active[index]:
                    raise ValueError(f"terminal {terminal.name!r} contains inactive node {node!r}")
                rhs[index] += share
        free = self.free_nodes.reshape(-1)
        rhs[~free] = 0.0
        if voltage_terminals:
            fixed = self.dirichlet_potential(voltage_terminals)
            lifted = self._device.stiffness_f64(fixed)
            rhs = np.where(free, rhs - lifted, fixed)
        return rhs

    def element_joule_loss(self, potential_v) -> np.ndarray:
        """``a v^T U_x v + b v^T U_y v + c v^T U_z v`` per element (pcb.py:843-860)."""

        v = np.stack(corner_views(np.asarray(potential_v, dtype=np.float64).reshape(self.mesh.node_shape)), axis=-1)
        loss = np.zeros(self.mesh.element_grid_shape)
        for k in range(3):
            loss += self._coef[k] * np.einsum("...i,ij,...j->...", v, self._unit[k], v, optimize=True)
        return loss

    def terminal_currents(self, potential_v, voltage_terminals) -> np.ndarray:
        """Current each voltage terminal drives in: the unconstrained residual
        summed over its nodes (pcb.py:761-780)."""

        flux = self._device.stiffness_f64(np.asarray(potential_v, dtype=np.float64).reshape(-1))
        return np.asarray([float(sum(flux[flat_index(n, self.mesh.node_shape)] for n in t.nodes))
                           for t in voltage_terminals])


def solve_hex_dc(problem: HexConductionProblem, *, config: MPIRConfig | None = None, backend: str | None = None,
                 device_id: int = 0, initial_potential_v=None) -> HexConductionSolution:
    """One DC solve (the reference's ``solve_pcb_dc`` flow, pcb.py:896-978)."""

    operator = MatrixFreeHexDCOperator(
        problem.mesh, reference_node=problem.reference_node,
        dirichlet_nodes=tuple(n for t in problem.voltage_terminals for n in t.nodes),
        backend=backend, device_id=device_id,
    )
    rhs = operator.build_rhs(problem.terminals, problem.voltage_terminals)
    fixed = operator.dirichlet_potential(problem.voltage_terminals)
    initial = None
    if initial_potential_v is not None:
        initial = np.nan_to_num(np.asarray(initial_potential_v, dtype=np.float64).reshape(-1), nan=0.0)
        initial = np.where(operator.free_nodes.reshape(-1), initial, fixed)
    result = solve_mpir(operator, rhs, config=config or jacobi_default_config(16), initial_guess=initial)
    potential = result.solution.reshape(problem.mesh.node_shape)
    losses = operator.element_joule_loss(potential)
    return HexConductionSolution(
        potential_v=potential,
        element_joule_loss_w=losses,
        joule_loss_w=float(np.sum(losses)),
        solve=result,
        voltage_terminal_current_a=operator.terminal_currents(potential, problem.voltage_terminals),
    )


def solve_hex_dc_many(mesh: LayeredThermalMesh, excitations: Sequence[Sequence[CurrentTerminal]],
                      reference_node: Node, *, config: MPIRConfig | None = None, backend: str | None = None,
                      device_id: int = 0) -> list[tuple[np.ndarray, MPIRResult]]:
    """Many balanced current excitations of one conductor, solved together as
    one batch of right-hand sides on one device-resident operator."""

    for terminals in excitations:
        _check_terminals(tuple(terminals), (), reference_node)
    config = config or jacobi_default_config(16)
    operator = MatrixFreeHexDCOperator(mesh, reference_node=reference_node, backend=backend, device_id=device_id)
    batch = _capi.Batch([operator.device_operator] * len(excitations))
    for i, terminals in enumerate(excitations):
        batch.set(i, operator.build_rhs(terminals), None, config.to_c())
    batch.run()
    out = []
    for i in range(len(excitations)):
        x, r = batch.result(i)
        out.append((x.reshape(mesh.node_shape), MPIRResult(
            solution=x, converged=r["converged"], outer_iterations=r["outer_iterations"],
            inner_iterations=r["inner_iterations"], relative_residual=r["relative_residual"],
            high_operator_applications=r["high_operator_applications"],
            low_operator_applications=r["low_operator_applications"], low_runtime="hip-fp32",
            history=tuple(MPIRStep(*h) for h in r["history"]))))
    return out
