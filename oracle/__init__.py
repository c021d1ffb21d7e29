"""ORACLE -- test infrastructure only (the checker, never the product).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It restates, on the CPU, the
reference's hexahedral Q1 conduction path:

* ``oracle.problem``   -- problem arrays, element coefficients, unit tensors,
  Robin lumping, nodal loads, diagonal, right-hand side and heat budget in
  FP64 NumPy (reference ``src/thermal/matrix_free_mpir_fem/{mesh,boundaries,
  operator,solve}.py``);
* ``oracle.assembled`` -- the operator assembled element by element into a
  SciPy sparse matrix, the independent check of every matrix-free action;
* ``oracle.native``    -- ctypes bindings of ``q1hex_oracle.c``: the line
  gather, the SPMD FP32 Jacobi PCG and the FP64 MPIR loop of the reference's
  ``native/hex_q1.cpp``, built by ``oracle/Makefile``;
* ``oracle.sheet``     -- the layered-sheet DC operator assembled into a
  sparse matrix, its direct solve, port conductances and element losses
  (reference ``src/electrical/matrix_free_mpir_fem/{pcb,ports}.py``);
* ``oracle.two_level`` -- the two-level preconditioner from an explicit
  prolongation: ``P^T A P``, its FP64 action and an FP64 PCG
  (reference ``src/electrical/matrix_free_mpir_fem/two_level.py``).

Each module's docstring names what pins it; the pinning tests are
``tests/test_oracle.py``, ``tests/test_oracle_sheet.py`` and
``tests/test_two_level.py``.
"""
