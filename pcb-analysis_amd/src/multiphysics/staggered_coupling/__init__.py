"""Staggered electro-thermal coupling on the HIP device (the reference's
``multiphysics.staggered_coupling`` electro-thermal loop)."""

from .electro_thermal import (
    COPPER_TEMPERATURE_COEFFICIENT_PER_K,
    CouplingConfig,
    CouplingStep,
    ElectroThermalResult,
    ElectroThermalScenario,
    TemperatureFixedPoint,
    conductivity_at_temperature,
    electrical_layer_temperature_k,
    heated_electrical_problem,
    run_electro_thermal,
    run_electro_thermal_many,
    slab_element_temperature_k,
    thermal_problem_with_joule_heat,
    via_node_temperature_k,
    via_temperature_k,
)

__all__ = [
    "COPPER_TEMPERATURE_COEFFICIENT_PER_K",
    "CouplingConfig",
    "CouplingStep",
    "ElectroThermalResult",
    "ElectroThermalScenario",
    "TemperatureFixedPoint",
    "conductivity_at_temperature",
    "electrical_layer_temperature_k",
    "heated_electrical_problem",
    "run_electro_thermal",
    "run_electro_thermal_many",
    "slab_element_temperature_k",
    "thermal_problem_with_joule_heat",
    "via_node_temperature_k",
    "via_temperature_k",
]
