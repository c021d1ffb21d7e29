"""Matrix-free MPIR solver contract on the HIP device and the layered-PCB DC
front end (the reference's ``electrical.matrix_free_mpir_fem`` solver,
config and result types, ``pcb`` and ``ports`` modules)."""

from .grid import TensorGrid, graded_edges, refined_grid
from .pcb import (
    COPPER_CONDUCTIVITY_S_PER_M,
    CurrentTerminal,
    LayeredPCBMesh,
    MatrixFreePCBOperator,
    PCBConductionProblem,
    PCBConductionSolution,
    ViaConnection,
    VoltageTerminal,
    solve_pcb_dc,
)
from .ports import DCPortBasis, PortSet, dc_port_basis, reduce_correlation
from .runtime import HipFloat32Runtime, RuntimeBackend, make_float32_runtime
from .solver import MPIRConfig, MPIRResult, MPIRStep, mpir_result_from_core, solve_mpir, solve_mpir_device
from .two_level import DEFAULT_MAX_COARSE_SIZE, AggregationCoarseCorrection, choose_block_size

__all__ = [
    "AggregationCoarseCorrection",
    "COPPER_CONDUCTIVITY_S_PER_M",
    "CurrentTerminal",
    "DCPortBasis",
    "HipFloat32Runtime",
    "LayeredPCBMesh",
    "MatrixFreePCBOperator",
    "PCBConductionProblem",
    "PCBConductionSolution",
    "PortSet",
    "ViaConnection",
    "VoltageTerminal",
    "MPIRConfig",
    "MPIRResult",
    "MPIRStep",
    "RuntimeBackend",
    "TensorGrid",
    "DEFAULT_MAX_COARSE_SIZE",
    "choose_block_size",
    "dc_port_basis",
    "graded_edges",
    "make_float32_runtime",
    "mpir_result_from_core",
    "reduce_correlation",
    "refined_grid",
    "solve_mpir",
    "solve_mpir_device",
    "solve_pcb_dc",
]
