"""Electrical side of the conduction solver: the shared MPIR driver and the
hexahedral DC-conduction front end (a drop-in for the reference's
``electrical.matrix_free_mpir_fem`` solver contract on the HIP device)."""
