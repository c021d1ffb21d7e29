"""Matrix-free hexahedral Q1 steady and transient heat conduction, solved by mixed-precision
iterative refinement on an MI355X (the reference's
``thermal.matrix_free_mpir_fem`` API; see the package modules for the
reference files each one restates)."""

from electrical.matrix_free_mpir_fem.two_level import AggregationCoarseCorrection, choose_block_size

from .boundaries import Convection, ConvectionBoundary, ExposedFaceConvection, HeatSource
from .coupling import element_heat_from_losses, element_joule_heat_w, via_heat_sources, via_joule_heat_sources
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
from .radiation import (
    STEFAN_BOLTZMANN_W_PER_M2_K4,
    ExposedFaceRadiation,
    Radiation,
    RadiationBoundary,
    element_temperature_k,
    face_temperature_k,
    newton_linearisation,
)
from .solve import ThermalConductionSolution, solve_thermal_conduction
from .transient import TimeSchedule, TransientStep, TransientThermalSolution, solve_thermal_transient

__all__ = [
    "AggregationCoarseCorrection",
    "COPPER_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "Convection",
    "ConvectionBoundary",
    "ExposedFaceConvection",
    "ExposedFaceRadiation",
    "FACE_DIRECTIONS",
    "FR4_IN_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "FR4_THROUGH_PLANE_THERMAL_CONDUCTIVITY_W_PER_M_K",
    "HeatSource",
    "LayeredThermalMesh",
    "MatrixFreeThermalOperator",
    "Radiation",
    "RadiationBoundary",
    "STEFAN_BOLTZMANN_W_PER_M2_K4",
    "ThermalConductionProblem",
    "ThermalConductionSolution",
    "TimeSchedule",
    "TransientStep",
    "TransientThermalSolution",
    "choose_block_size",
    "element_heat_from_losses",
    "element_joule_heat_w",
    "element_temperature_k",
    "exposed_element_faces",
    "face_temperature_k",
    "newton_linearisation",
    "solve_thermal_conduction",
    "solve_thermal_transient",
    "unit_hexahedron_matrices",
    "via_heat_sources",
    "via_joule_heat_sources",
]
