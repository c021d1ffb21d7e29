"""Thermal side of the conduction solver (a drop-in for the reference's
``thermal.matrix_free_mpir_fem`` steady-conduction API on the HIP device)."""
