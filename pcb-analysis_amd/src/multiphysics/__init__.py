"""Coupled analyses driving the HIP conduction solves (the reference's
``multiphysics`` package; only the electro-thermal staggered loop)."""
