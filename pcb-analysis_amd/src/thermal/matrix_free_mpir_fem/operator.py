"""Matrix-free conduction operator resident on the HIP device (reference:
src/thermal/matrix_free_mpir_fem/operator.py:30-525).

The host prepares what the reference's ThermalSystem prepares
(cpp/src/thermal/thermal_system.cpp:252-363): per-element coefficients,
lumped Robin conductances (plus C/dt), the free-node mask, fixed
temperatures and the nodal load.  The operator arrays then live on the
device in both precisions; the Jacobi diagonal, every operator action and
the whole MPIR solve run there (``pcbhip`` C ABI).  Rows of fixed nodes (and
of nodes touched by no active element) are the identity; the Robin term is
on the diagonal, so the operator is symmetric positive definite.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from electrical.matrix_free_mpir_fem.runtime import HipFloat32Runtime, make_float32_runtime
from electrical.matrix_free_mpir_fem.solver import MPIRConfig
from electrical.matrix_free_mpir_fem.two_level import AggregationCoarseCorrection
from pcbhip import _capi, serving

from .mesh import Preconditioner, corner_views, flat_index, unit_hexahedron_matrices
from .problem import ThermalConductionProblem

LOW_KERNEL_NAME = "hip-gfx950-sum-factorised-hex-q1"
HIGH_KERNEL_NAME = "hip-gfx950-sum-factorised-hex-q1-fp64"


class MatrixFreeThermalOperator:
    """Split FP64/FP32 hexahedral conduction operator on one HIP device."""

    def __init__(
        self,
        problem: ThermalConductionProblem,
        *,
        runtime: HipFloat32Runtime | None = None,
        backend: str | None = None,
        device_id: int = 0,
        preconditioner: Preconditioner = "two-level",
        coarse_block_nodes: int | None = None,
        native: bool | None = None,
        capacity_per_s: np.ndarray | None = None,
    ) -> None:
        if runtime is not None and backend is not None:
            raise ValueError("pass either runtime or backend, not both")
        if preconditioner not in ("two-level", "jacobi"):
            raise ValueError("preconditioner must be 'two-level' or 'jacobi'")
        if native is False:
            raise ValueError("native=False asks for the NumPy reference path, which this package does not ship")
        self.preconditioner = preconditioner
        self.coarse_correction: AggregationCoarseCorrection | None = None
        self.high_dtype = np.float64
        self.inner_solver = "pcg"
        self.problem = problem
        self.mesh = mesh = problem.mesh
        self.size = mesh.size
        self.runtime = runtime or make_float32_runtime(backend or "hip", device_id=device_id)

        self._unit_high = np.stack(unit_hexahedron_matrices())
        self._coefficients_high = np.stack(mesh.element_coefficients())
        self.active_nodes = mesh.active_nodes
        self.fixed_nodes = problem.fixed_temperature_mask | ~self.active_nodes
        self.free_nodes = ~self.fixed_nodes
        self.fixed_temperature_k = np.where(
            problem.fixed_temperature_mask & self.active_nodes, problem.fixed_temperature_k, 0.0
        )
        self._robin_weights_high: list[np.ndarray] = []
        self._robin_rhs_weights_high: list[np.ndarray] = []
        self._robin_ambient_k: list[float] = []
        for boundary in problem.convection:
            w, rw = boundary.lumped_nodal_weights(mesh)
            self._robin_weights_high.append(w)
            self._robin_rhs_weights_high.append(rw)
            self._robin_ambient_k.append(boundary.mean_ambient_k())
        n = self.size
        self._robin_only_high = np.sum(self._robin_weights_high, axis=0) if self._robin_weights_high else np.zeros(n)
        if capacity_per_s is None:
            self._capacity_high = np.zeros(n)
        else:
            capacity = np.asarray(capacity_per_s, dtype=np.float64).reshape(-1)
            if capacity.size != n:
                raise ValueError("capacity_per_s must hold one value per node")
            if not np.all(np.isfinite(capacity)) or np.any(capacity < 0.0):
                raise ValueError("capacity_per_s must be finite and non-negative")
            self._capacity_high = capacity
        self._robin_total_high = self._robin_only_high + self._capacity_high
        self._robin_rhs_high = (
            np.sum(self._robin_rhs_weights_high, axis=0) if self._robin_rhs_weights_high else np.zeros(n)
        )
        self._device = _capi.Operator(
            self.runtime.device_id,
            mesh.element_grid_shape,
            self._coefficients_high,
            self._robin_total_high,
            self.free_nodes.reshape(-1),
        )
        self._diagonal_high = self._device.diagonal()
        self._diagonal_low = self._diagonal_high.astype(np.float32)
        if preconditioner == "two-level":
            # The coarse space of the reference's default preconditioner
            # (operator.py:168-178): assembled from device FP64 applies, its
            # inverse resident on the device for every solve of this operator.
            self.coarse_correction = AggregationCoarseCorrection(
                node_shape=mesh.node_shape, free_nodes=self.free_nodes, diagonal_high=self._diagonal_high,
                apply_high=self.apply_high, block=coarse_block_nodes, galerkin=self._device.galerkin_f64,
                invert=self._device.set_coarse_matrix,
            )
        self.low_operator_backend = LOW_KERNEL_NAME
        self.high_operator_backend = HIGH_KERNEL_NAME

    # ------------------------------------------------------------ actions
    def _flat(self, values, name: str = "vector") -> np.ndarray:
        flat = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if flat.size != self.size:
            raise ValueError(f"{name} has size {flat.size}, expected {self.size}")
        return flat

    @property
    def device_operator(self) -> _capi.Operator:
        return self._device

    def apply_high(self, vector: np.ndarray) -> np.ndarray:
        return self._device.apply_f64(self._flat(vector))

    def apply_low(self, vector: Any):
        """``A x`` in FP32: a device vector of the runtime gives a new device
        vector, a host array a host array."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.apply_f32_vec(vector)
        v = np.ascontiguousarray(vector, dtype=np.float32).reshape(-1)
        if v.size != self.size:
            raise ValueError(f"vector has size {v.size}, expected {self.size}")
        return self._device.apply_f32(v)

    def stiffness_high(self, vector: np.ndarray) -> np.ndarray:
        """``K x``: conduction only, no Robin term, every node free."""

        return self._device.stiffness_f64(self._flat(vector))

    def diagonal_low(self) -> np.ndarray:
        return self._diagonal_low

    def precondition_low(self, vector: Any):
        """``M^-1 r`` in FP32 on the device: Jacobi, plus the coarse
        correction with the two-level preconditioner (operator.py:366-371);
        device vector in, device vector out, or host array in and out."""

        if isinstance(vector, _capi.DeviceVector):
            return self._device.precondition_f32_vec(vector)
        v = np.ascontiguousarray(vector, dtype=np.float32).reshape(-1)
        if v.size != self.size:
            raise ValueError(f"vector has size {v.size}, expected {self.size}")
        return self._device.precondition_f32(v)

    def native_solve_mpir(self, rhs_high: np.ndarray, config: MPIRConfig, initial_guess: np.ndarray | None = None):
        """The whole MPIR solve on the device (operator.py:318-347 contract)."""

        x, result = serving.solve_mpir_routed(
            self._device, self._flat(rhs_high, "rhs"),
            None if initial_guess is None else self._flat(initial_guess, "initial_guess"), config.to_c())
        return (
            x,
            result["converged"],
            result["outer_iterations"],
            result["inner_iterations"],
            result["relative_residual"],
            result["high_operator_applications"],
            result["low_operator_applications"],
            result["history"],
        )

    def device_system(self, device: int = 0) -> "DeviceSystem":
        """This operator's solve as the reference's device-system contract
        (``solve_mpir_device``; cpp/bindings/cuda_module.cpp:64-104)."""

        if device != self.runtime.device_id:
            raise ValueError("the operator lives on another device")
        return DeviceSystem(self)

    # --------------------------------------------------------------- loads
    def nodal_load(self) -> np.ndarray:
        """Heat input per node: lumped element heat plus nodal sources (operator.py:374-387)."""

        load = self.problem.nodal_heat_w.copy()
        eighth = self.problem.element_heat_w / 8.0
        for view in corner_views(load):
            view[...] += eighth
        for source in self.problem.heat_sources:
            share = float(source.power_w) / len(source.nodes)
            for node in source.nodes:
                load.flat[flat_index(node, self.mesh.node_shape)] += share
        return load.reshape(-1)

    def default_reference_temperature(self) -> float:
        """Ambient of the first convective face, else the mean fixed value (operator.py:389-397)."""

        if self._robin_ambient_k:
            return float(self._robin_ambient_k[0])
        mask = self.problem.fixed_temperature_mask & self.active_nodes
        return float(np.mean(self.problem.fixed_temperature_k[mask]))

    def _capacity_load(self, previous_temperature_k) -> np.ndarray:
        if previous_temperature_k is None:
            if np.any(self._capacity_high > 0.0):
                raise ValueError("a transient operator needs previous_temperature_k")
            return np.zeros(self.size)
        previous = np.asarray(previous_temperature_k, dtype=np.float64).reshape(-1)
        if previous.size != self.size:
            raise ValueError("previous_temperature_k must hold one value per node")
        return self._capacity_high * np.where(np.isfinite(previous), previous, 0.0)

    def build_rhs(self, reference_temperature_k: float = 0.0, previous_temperature_k=None) -> np.ndarray:
        """Right-hand side for the rise ``T - reference`` (operator.py:411-440):
        fixed temperatures lifted onto the free rows through ``K`` (on the
        device), shifted values on the fixed rows."""

        reference = float(reference_temperature_k)
        fixed = self.fixed_nodes.reshape(-1)
        active = self.active_nodes.reshape(-1)
        shifted = np.where(fixed & active, self.fixed_temperature_k.reshape(-1) - reference, 0.0)
        lifted = self.stiffness_high(shifted)
        robin = self._robin_rhs_high - self._robin_total_high * reference
        rhs = self.nodal_load() + robin + self._capacity_load(previous_temperature_k) - lifted
        rhs[fixed] = shifted[fixed]
        return rhs

    # -------------------------------------------------------- post-process
    def stored_heat_w(self, temperature_k, previous_temperature_k) -> np.ndarray:
        flat = np.asarray(temperature_k, dtype=np.float64).reshape(-1)
        if previous_temperature_k is None:
            return np.zeros(self.size)
        previous = np.asarray(previous_temperature_k, dtype=np.float64).reshape(-1)
        finite = np.isfinite(flat) & np.isfinite(previous)
        return np.where(finite, self._capacity_high * (flat - previous), 0.0)

    def unconstrained_residual(self, temperature_k, previous_temperature_k=None) -> np.ndarray:
        """``K T + R (T - T_amb) + C/dt (T - T_n) - q`` at every node (operator.py:458-482)."""

        flat = self._flat(temperature_k, "temperature_k")
        return (
            self.stiffness_high(flat)
            + self._robin_only_high * flat
            - self._robin_rhs_high
            + self.stored_heat_w(flat, previous_temperature_k)
            - self.nodal_load()
        )

    def convective_heat(self, temperature_k) -> np.ndarray:
        """Heat removed by each convection boundary, in W (operator.py:484-498)."""

        flat = self._flat(temperature_k, "temperature_k")
        return np.asarray(
            [float(np.sum(w * flat) - np.sum(rw)) for w, rw in zip(self._robin_weights_high, self._robin_rhs_weights_high)]
        )

    def element_heat_flux(self, temperature_k) -> np.ndarray:
        """``-k grad T`` at element centres from corner averages (operator.py:500-525)."""

        grid = np.asarray(temperature_k, dtype=np.float64).reshape(self.mesh.node_shape)
        c = corner_views(grid)
        gx = ((c[1] + c[3] + c[5] + c[7]) - (c[0] + c[2] + c[4] + c[6])) / (4.0 * self.mesh.pitch_x_m[None, None, :])
        gy = ((c[2] + c[3] + c[6] + c[7]) - (c[0] + c[1] + c[4] + c[5])) / (4.0 * self.mesh.pitch_y_m[None, :, None])
        hz = np.asarray(self.mesh.slab_thickness_m)[:, None, None]
        gz = ((c[4] + c[5] + c[6] + c[7]) - (c[0] + c[1] + c[2] + c[3])) / (4.0 * hz)
        k_in = np.asarray(self.mesh.conductivity_w_per_m_k)
        k_z = np.asarray(self.mesh.through_plane_conductivity_w_per_m_k)
        return np.stack((-k_in * gx, -k_in * gy, -k_z * gz), axis=-1)


class DeviceSystem:
    """``solve(rhs, x0, rtol, atol, inner_rtol, max_outer, max_inner, threads)
    -> dict``, the reference's DeviceSystem contract, on the HIP device."""

    def __init__(self, operator: MatrixFreeThermalOperator) -> None:
        self.operator = operator

    def solve(self, rhs, x0, rtol, atol, inner_rtol, max_outer, max_inner, threads=1) -> dict:
        config = MPIRConfig(relative_tolerance=rtol, absolute_tolerance=atol, inner_relative_tolerance=inner_rtol,
                            max_outer_iterations=max_outer, max_inner_iterations=max_inner)
        out = self.operator.native_solve_mpir(rhs, config, x0)
        keys = ("solution", "converged", "outer_iterations", "inner_iterations", "relative_residual",
                "high_operator_applications", "low_operator_applications", "history")
        return dict(zip(keys, out))
