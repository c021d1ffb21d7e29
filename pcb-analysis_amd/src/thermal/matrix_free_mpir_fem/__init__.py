"""Matrix-free hexahedral Q1 steady and transient heat conduction, solved by mixed-precision
iterative refinement on an MI355X (the reference's
``thermal.matrix_free_mpir_fem`` API; see the package modules for the
reference files each one restates)."""

from .boundaries import Convection, ConvectionBoundary, ExposedFaceConvection, HeatSource
from .mesh import (
    COPPER_THERMAL_CONDUCTIVITY_W_PER_M_K,
    FACE_DIRECTIONS,
    FR4_IN_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K,
    FR4_THROUGH_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K,
    LayeredThermalMesh,
    exposed_element_faces,
    unit_hexahedron_matrices,
)
from .operator import MatrixFreeThermalOperator
from .problem import ThermalConductionProblem
from .solve import ThermalConductionSolution, solve_thermal_conduction
from .transient import TimeSchedule, TransientStep, TransientThermalSolution, solve_thermal_transient

__all__ = [
    "COPPER_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "Convection",
    "ConvectionBoundary",
    "ExposedFaceConvection",
    "FACE_DIRECTIONS",
    "FR4_IN_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "FR4_THROUGH_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "HeatSource",
    "LayeredThermalMesh",
    "MatrixFreeThermalOperator",
    "ThermalConductionProblem",
    "ThermalConductionSolution",
    "TimeSchedule",
    "TransientStep",
    "TransientThermalSolution",
    "exposed_element_faces",
    "solve_thermal_conduction",
    "solve_thermal_transient",
    "unit_hexahedron_matrices",
]
