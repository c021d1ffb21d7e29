"""Matrix-free MPIR solver contract on the HIP device (the reference's
``electrical.matrix_free_mpir_fem`` solver, config and result types)."""

from .runtime import HipFloat32Runtime, RuntimeBackend, make_float32_runtime
from .solver import MPIRConfig, MPIRResult, MPIRStep, mpir_result_from_core, solve_mpir, solve_mpir_device

__all__ = [
    "HipFloat32Runtime",
    "MPIRConfig",
    "MPIRResult",
    "MPIRStep",
    "RuntimeBackend",
    "make_float32_runtime",
    "mpir_result_from_core",
    "solve_mpir",
    "solve_mpir_device",
]
