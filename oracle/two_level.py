"""ORACLE -- test infrastructure only.  The two-level preconditioner of the
reference (src/electrical/matrix_free_mpir_fem/two_level.py:54-229; native
hex_q1.cpp:388-432, 703-835) written from its definition with an explicit
sparse prolongation, the check of the product's coloured assembly and of its
device action.

``P`` is the N x n_c patch indicator on free nodes: node (l, y, x) belongs to
patch (l, y // block, x // block).  The coarse matrix is ``P^T A P`` with a
unit diagonal on patches that hold no free node (two_level.py:101-107, the
reference's own test tests/test_thermal_conduction.py:259-303 builds its
expectation the same way) and the preconditioner is
``M^-1 r = r / diag + P C^-1 P^T r``.

Pinned by: the Galerkin product against a brute-force double sum over node
pairs on tiny grids, SPD of ``M^-1``, and the reduction to Jacobi when every
patch is empty (tests/test_two_level.py).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def coarse_shape(node_shape, block):
    layers, rows, cols = node_shape
    return (layers, -(-rows // block), -(-cols // block))


def prolongation(node_shape, free, block) -> sp.csr_matrix:
    """The patch indicator ``P`` (N x n_c), zero rows on fixed nodes."""

    shape = coarse_shape(node_shape, block)
    layer, row, col = np.indices(node_shape)
    patch = np.ravel_multi_index((layer, row // block, col // block), shape).reshape(-1)
    keep = np.asarray(free, dtype=bool).reshape(-1)
    n = keep.size
    nodes = np.arange(n)[keep]
    return sp.csr_matrix((np.ones(nodes.size), (nodes, patch[keep])), shape=(n, int(np.prod(shape))))


def galerkin(a: sp.spmatrix, node_shape, free, block) -> np.ndarray:
    """``P^T A P`` with a unit diagonal on empty patches."""

    p = prolongation(node_shape, free, block)
    coarse = np.asarray((p.T @ (a @ p)).todense())
    empty = np.asarray(p.sum(axis=0)).reshape(-1) == 0.0
    coarse[empty, :] = 0.0
    coarse[:, empty] = 0.0
    coarse[empty, empty] = 1.0
    return coarse


def galerkin_brute(a_dense: np.ndarray, node_shape, free, block) -> np.ndarray:
    """The same matrix as a double loop over node pairs (tiny grids only)."""

    shape = coarse_shape(node_shape, block)
    nc = int(np.prod(shape))
    free = np.asarray(free, dtype=bool).reshape(-1)
    layer, row, col = np.indices(node_shape)
    patch = np.ravel_multi_index((layer, row // block, col // block), shape).reshape(-1)
    coarse = np.zeros((nc, nc))
    count = np.zeros(nc)
    n = free.size
    for i in range(n):
        if not free[i]:
            continue
        count[patch[i]] += 1
        for j in range(n):
            if free[j]:
                coarse[patch[i], patch[j]] += a_dense[i, j]
    for k in range(nc):
        if count[k] == 0:
            coarse[k, :] = 0.0
            coarse[:, k] = 0.0
            coarse[k, k] = 1.0
    return coarse


def preconditioner(a: sp.spmatrix, diag, node_shape, free, block):
    """The FP64 action ``r -> r / diag + P C^-1 P^T r``."""

    p = prolongation(node_shape, free, block)
    inverse = np.linalg.inv(galerkin(a, node_shape, free, block))
    diag = np.asarray(diag, dtype=np.float64).reshape(-1)

    def apply(r):
        r = np.asarray(r, dtype=np.float64).reshape(-1)
        return r / diag + p @ (inverse @ (p.T @ r))

    return apply


def pcg(a: sp.spmatrix, b, apply_m, tol: float, max_iterations: int):
    """FP64 preconditioned CG from zero to ``||r|| <= tol ||b||``; returns
    (x, iterations)."""

    b = np.asarray(b, dtype=np.float64).reshape(-1)
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = float(r @ z)
    norm_b = float(np.linalg.norm(b))
    for k in range(1, max_iterations + 1):
        q = a @ p
        alpha = rz / float(p @ q)
        x += alpha * p
        r -= alpha * q
        if float(np.linalg.norm(r)) <= tol * norm_b:
            return x, k
        z = apply_m(r)
        rz_next = float(r @ z)
        p = z + (rz_next / rz) * p
        rz = rz_next
    return x, max_iterations
