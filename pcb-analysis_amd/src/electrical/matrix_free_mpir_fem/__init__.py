"""Matrix-free MPIR solver contract on the HIP device and the layered-PCB DC
front end (the reference's ``electrical.matrix_free_mpir_fem`` solver,
config and result types and ``pcb`` module)."""

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
from .runtime import HipFloat32Runtime, RuntimeBackend, make_float32_runtime
from .solver import MPIRConfig, MPIRResult, MPIRStep, mpir_result_from_core, solve_mpir, solve_mpir_device

__all__ = [
    "COPPER_CONDUCTIVITY_S_PER_M",
    "CurrentTerminal",
    "HipFloat32Runtime",
    "LayeredPCBMesh",
    "MatrixFreePCBOperator",
    "PCBConductionProblem",
    "PCBConductionSolution",
    "ViaConnection",
    "VoltageTerminal",
    "MPIRConfig",
    "MPIRResult",
    "MPIRStep",
    "RuntimeBackend",
    "make_float32_runtime",
    "mpir_result_from_core",
    "solve_mpir",
    "solve_mpir_device",
    "solve_pcb_dc",
]
