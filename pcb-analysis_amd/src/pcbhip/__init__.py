"""MI355X (gfx950) hexahedral Q1 conduction solver: the C-ABI binding, the
seeded workload generators and the batch/serving drivers.

The solver front ends that mirror the reference's Python API live in the
``thermal.matrix_free_mpir_fem`` and ``electrical.matrix_free_mpir_fem``
packages beside this one.
"""

from . import _capi

__all__ = ["_capi"]
