"""The copper of one conductor role as an N-port, measured on the HIP device
(reference: src/electrical/matrix_free_mpir_fem/ports.py).

Every port is a pad: a set of nodes held at one potential.  With ``n`` ports,
``n - 1`` unit voltage solves (one driven port at 1 V, every other port at
0 V) give the conductance matrix ``G`` (symmetric, positive semi-definite,
zero row sums) and the unit potential fields.  The potential for port
currents ``I`` is linear in ``I``, so the time-averaged loss of a current
waveform needs only its second moments ``C = <I I^T>``:
``<P_e> = sum_kl C_kl phi_k^T K_e phi_l``, evaluated by diagonalising ``C``
on the driven ports into weighted potential modes.

The unit solves share one operator whose fixed nodes are every pad node.
Here they run as one batch of slots over that device-resident operator
(``pcbhip_batch_*``), all slots iterating together in one captured graph,
where the reference runs them one after another or on a CPU thread pool
(ports.py:358-380).  Each slot's arithmetic is that of its solve alone, so
the basis does not depend on how many ports are batched.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np

from pcbhip import _capi

from .pcb import LayeredPCBMesh, MatrixFreePCBOperator, Node, Preconditioner, ViaConnection, VoltageTerminal
from .solver import MPIRConfig, MPIRResult, MPIRStep, default_config


def reduce_correlation(ports: "PortSet", correlation_a2: np.ndarray) -> tuple[np.ndarray, float]:
    """The driven-port block of a port-current correlation and its scale,
    after the checks every reduction needs (ports.py:103-118): square with
    one row per port, symmetric, and rows summing to zero (KCL)."""

    matrix = np.asarray(correlation_a2, dtype=np.float64)
    n = ports.count
    if matrix.shape != (n, n):
        raise ValueError(f"correlation must be {n}×{n}, one row and column per port")
    if not np.allclose(matrix, matrix.T, rtol=1.0e-10, atol=0.0):
        raise ValueError("correlation must be symmetric")
    scale = max(1.0, float(np.max(np.abs(matrix))))
    if np.any(np.abs(matrix.sum(axis=1)) > 1.0e-9 * scale):
        raise ValueError("correlation rows must sum to zero (port currents satisfy KCL)")
    driven = list(ports.driven)
    return np.ascontiguousarray(matrix[np.ix_(driven, driven)]), scale


@dataclass(frozen=True)
class PortSet:
    """The pads of one conductor role, one node set per port (ports.py:148-203).

    ``reference`` indexes the port whose potential is the zero of every
    reported field.  A node may belong to one port only."""

    pads: tuple[tuple[Node, ...], ...]
    names: tuple[str, ...] = ()
    reference: int = 0

    def __post_init__(self) -> None:
        pads = tuple(tuple(tuple(int(i) for i in node) for node in pad) for pad in self.pads)
        if len(pads) < 2:
            raise ValueError("a port set needs at least two ports")
        if any(len(pad) == 0 for pad in pads):
            raise ValueError("every port needs at least one node")
        owner: dict[Node, int] = {}
        for index, pad in enumerate(pads):
            for node in pad:
                if node in owner:
                    raise ValueError(f"node {node!r} belongs to ports {owner[node]} and {index}")
                owner[node] = index
        names = tuple(self.names) if self.names else tuple(f"port_{i}" for i in range(len(pads)))
        if len(names) != len(pads) or len(set(names)) != len(names):
            raise ValueError("port names must be unique and one per port")
        reference = int(self.reference)
        if not 0 <= reference < len(pads):
            raise ValueError("reference must index a port")
        object.__setattr__(self, "pads", pads)
        object.__setattr__(self, "names", names)
        object.__setattr__(self, "reference", reference)

    @property
    def count(self) -> int:
        return len(self.pads)

    @property
    def driven(self) -> tuple[int, ...]:
        """Every port but the reference, in port order."""

        return tuple(i for i in range(self.count) if i != self.reference)

    def voltage_terminals(self, voltage_v: Sequence[float]) -> tuple[VoltageTerminal, ...]:
        values = np.asarray(voltage_v, dtype=np.float64)
        if values.shape != (self.count,):
            raise ValueError("one voltage per port is required")
        return tuple(VoltageTerminal(pad, float(v), name) for pad, v, name in zip(self.pads, values, self.names))


@dataclass(frozen=True)
class DCPortBasis:
    """The N-port of one conductor and the unit fields that span its
    solutions (ports.py:206-330).

    ``conductance_s[j, k]`` is the current into the copper at port ``j`` with
    port ``k`` at 1 V and the others at 0 V.  ``unit_voltage_potential_v[m]``
    is that field for driven port ``ports.driven[m]``;
    ``unit_current_potential_v[m]`` is the field for 1 A into that port and
    out of the reference port.  ``workers`` is 1: the unit solves run as one
    device batch on one stream."""

    mesh: LayeredPCBMesh
    ports: PortSet
    operator: MatrixFreePCBOperator
    conductance_s: np.ndarray
    unit_voltage_potential_v: np.ndarray
    unit_current_potential_v: np.ndarray
    solves: tuple[MPIRResult, ...]
    workers: int = 1

    @property
    def converged(self) -> bool:
        return all(r.converged for r in self.solves)

    @property
    def inner_iterations(self) -> int:
        return int(sum(r.inner_iterations for r in self.solves))

    @property
    def reduced_conductance_s(self) -> np.ndarray:
        """``G`` on the driven ports, ``(n-1) × (n-1)``, SPD."""

        driven = list(self.ports.driven)
        return self.conductance_s[np.ix_(driven, driven)]

    @property
    def reduced_resistance_ohm(self) -> np.ndarray:
        """Pad-to-pad resistances against the reference port, ``G_rr^-1``."""

        return np.linalg.inv(self.reduced_conductance_s)

    def _driven_currents(self, current_a) -> np.ndarray:
        values = np.asarray(current_a, dtype=np.float64)
        if values.shape != (self.ports.count,):
            raise ValueError("one current per port is required")
        total = float(values.sum())
        if abs(total) > 1.0e-9 * max(1.0, float(np.abs(values).sum())):
            raise ValueError(f"port currents must sum to zero, got {total:.6g} A")
        return values[list(self.ports.driven)]

    def potential_v(self, current_a: Sequence[float]) -> np.ndarray:
        """Potential field (``mesh.node_shape``) for port currents into the copper."""

        return np.tensordot(self._driven_currents(current_a), self.unit_current_potential_v, axes=1)

    def port_voltage_v(self, current_a: Sequence[float]) -> np.ndarray:
        """Every port's potential for the given currents, the reference at 0 V."""

        voltages = np.zeros(self.ports.count)
        voltages[list(self.ports.driven)] = self.reduced_resistance_ohm @ self._driven_currents(current_a)
        return voltages

    def _modes(self, correlation_a2) -> tuple[np.ndarray, np.ndarray]:
        """Weights ``lambda_m`` and potential modes ``psi_m = sum_k u_mk phi_k``
        of the driven block of the correlation."""

        reduced, scale = reduce_correlation(self.ports, correlation_a2)
        weights, vectors = np.linalg.eigh(reduced)
        if float(weights.min()) < -1.0e-10 * scale:
            raise ValueError("correlation must be positive semi-definite")
        keep = weights > 1.0e-14 * scale
        return weights[keep], np.tensordot(vectors[:, keep].T, self.unit_current_potential_v, axes=1)

    def mean_loss_w(self, correlation_a2) -> tuple[np.ndarray, np.ndarray]:
        """Time-averaged Joule loss of every element and every via, in W, for
        port-current second moments ``correlation_a2[k, l] = <I_k I_l>``."""

        weights, modes = self._modes(correlation_a2)
        element = np.zeros(self.mesh.element_active.shape)
        via = np.zeros(len(self.operator.vias))
        for weight, mode in zip(weights, modes):
            element += weight * self.operator.element_joule_loss(mode)
            via += weight * self.operator.via_joule_loss(mode)
        return element, via

    def rms_current_density_a_per_m2(self, correlation_a2) -> np.ndarray:
        """Root-mean-square current-density magnitude of every element."""

        weights, modes = self._modes(correlation_a2)
        sigma = np.asarray(self.mesh.conductivity_s_per_m, dtype=np.float64)
        square = np.zeros(self.mesh.element_active.shape)
        for weight, mode in zip(weights, modes):
            density = sigma[..., None] * self.operator.element_electric_field(mode)
            square += weight * np.sum(density * density, axis=-1)
        return np.sqrt(square)


def dc_port_basis(mesh: LayeredPCBMesh, ports: PortSet, *, vias: Sequence[ViaConnection] = (),
                  config: MPIRConfig | None = None, runtime=None, backend: str | None = None, device_id: int = 0,
                  preconditioner: Preconditioner = "two-level", coarse_block_nodes: int | None = None,
                  native: bool | None = None, initial: DCPortBasis | None = None) -> DCPortBasis:
    """Measure the N-port of ``mesh`` at ``ports`` with ``n - 1`` unit voltage
    solves run as one device batch (ports.py:333-447).

    ``initial`` warm-starts each unit solve on its free nodes from the same
    port's unit field of an earlier basis on the same mesh shape and ports.
    The driven block of ``G`` is symmetrised against solver round-off; the
    reference port's row and column follow from KCL, so the row and column
    sums of ``G`` vanish to rounding."""

    operator = MatrixFreePCBOperator(
        mesh, dirichlet_nodes=tuple(node for pad in ports.pads for node in pad), vias=vias, runtime=runtime,
        backend=backend, device_id=device_id, preconditioner=preconditioner,
        coarse_block_nodes=coarse_block_nodes, native=native,
    )
    if initial is not None and (initial.ports != ports or initial.mesh.node_shape != mesh.node_shape):
        raise ValueError("initial basis must share the ports and the mesh shape")
    config = config or default_config(operator.preconditioner)
    n = ports.count
    driven = list(ports.driven)
    free = operator.free_nodes.reshape(-1)
    batch = _capi.Batch([operator.device_operator] * (n - 1))
    excitations = []
    for column, port in enumerate(driven):
        volts = np.zeros(n)
        volts[port] = 1.0
        terminals = ports.voltage_terminals(volts)
        excitations.append(terminals)
        guess = operator.dirichlet_potential(terminals)
        if initial is not None:
            guess = np.where(free, initial.unit_voltage_potential_v[column].reshape(-1), guess)
        batch.set(column, operator.build_rhs((), terminals), guess, config.to_c())
    batch.run()

    measured = np.zeros((n, n))
    unit_voltage = np.zeros((n - 1,) + mesh.node_shape)
    solves = []
    for column, port in enumerate(driven):
        x, r = batch.result(column)
        unit_voltage[column] = x.reshape(mesh.node_shape)
        measured[:, port] = operator.terminal_currents(x, excitations[column])
        solves.append(MPIRResult(
            solution=x, converged=r["converged"], outer_iterations=r["outer_iterations"],
            inner_iterations=r["inner_iterations"], relative_residual=r["relative_residual"],
            high_operator_applications=r["high_operator_applications"],
            low_operator_applications=r["low_operator_applications"], low_runtime=operator.runtime.name,
            history=tuple(MPIRStep(*h) for h in r["history"])))
    batch.close()

    block = measured[np.ix_(driven, driven)]
    block = 0.5 * (block + block.T)
    conductance = np.zeros((n, n))
    conductance[np.ix_(driven, driven)] = block
    conductance[driven, ports.reference] = -block.sum(axis=1)
    conductance[ports.reference, driven] = -block.sum(axis=0)
    conductance[ports.reference, ports.reference] = float(block.sum())
    # 1 A into driven port k: the voltages R[:, k] on the driven ports, so the
    # field is sum_j R[j, k] times the unit-voltage field of port j.
    resistance = np.linalg.inv(block)
    unit_current = np.tensordot(resistance.T, unit_voltage, axes=1)
    return DCPortBasis(mesh=mesh, ports=ports, operator=operator, conductance_s=conductance,
                       unit_voltage_potential_v=unit_voltage, unit_current_potential_v=unit_current,
                       solves=tuple(solves), workers=1)


__all__ = ["DCPortBasis", "PortSet", "dc_port_basis", "reduce_correlation"]
