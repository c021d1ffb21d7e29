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
pairs on tiny grids, SPD of ``M^-1``, the reduction to Jacobi when every
patch is empty, the coloured assembly against the sparse product, and the C
port's two-level solve of the reference's 4-slab fixture against the
iteration counts the reference publishes for it (tests/test_two_level.py).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def block_size(node_shape, max_coarse_size: int = 2048, minimum: int = 4) -> int:
    """The default patch width (two_level.py:37-51): the smallest width from
    ``minimum`` whose coarse space holds at most ``max_coarse_size`` patches,
    or the width of the wider node side."""

    layers, rows, cols = (int(v) for v in node_shape)
    block = max(1, int(minimum))
    while layers * -(-rows // block) * -(-cols // block) > max_coarse_size and block < max(rows, cols):
        block += 1
    return block


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


def coloured_galerkin(apply_f64, node_shape, free, block) -> np.ndarray:
    """``P^T A P`` as the native path assembles it (hex_q1.cpp:703-835): 27
    FP64 applications of ``A`` to the indicator of the free nodes in patches
    of one colour (node layer mod 3, patch row mod 3, patch column mod 3),
    each restricted to patch sums in node order; patches of one colour are
    three apart, so a patch's sum is its entry against the one neighbour of
    that colour.  Empty patches get a unit diagonal."""

    shape = coarse_shape(node_shape, block)
    free = np.asarray(free, dtype=bool).reshape(node_shape)
    layer, row, col = np.indices(node_shape)
    patch = np.ravel_multi_index((layer, row // block, col // block), shape)
    colour = (layer % 3) * 9 + ((row // block) % 3) * 3 + (col // block) % 3
    size = int(np.prod(shape))
    free_patches = patch[free]
    coarse_layer, coarse_row, coarse_col = np.indices(shape)
    coarse_colour = (coarse_layer % 3) * 9 + (coarse_row % 3) * 3 + coarse_col % 3
    matrix = np.zeros((size, size))
    for c in range(27):
        probe = free & (colour == c)
        if not probe.any():
            continue
        action = np.asarray(apply_f64(probe.astype(np.float64).reshape(-1)), dtype=np.float64).reshape(node_shape)
        restricted = np.bincount(free_patches, weights=action[free], minlength=size)
        for dl in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nl, ny, nx = coarse_layer + dl, coarse_row + dy, coarse_col + dx
                    ok = ((nl >= 0) & (nl < shape[0]) & (ny >= 0) & (ny < shape[1]) & (nx >= 0) & (nx < shape[2]))
                    i = np.ravel_multi_index((coarse_layer[ok], coarse_row[ok], coarse_col[ok]), shape)
                    j = np.ravel_multi_index((nl[ok], ny[ok], nx[ok]), shape)
                    hit = coarse_colour.reshape(-1)[j] == c
                    matrix[i[hit], j[hit]] = restricted[i[hit]]
    empty = np.bincount(free_patches, minlength=size) == 0
    matrix[empty, :] = 0.0
    matrix[:, empty] = 0.0
    matrix[empty, empty] = 1.0
    return matrix


def spd_inverse(matrix: np.ndarray) -> np.ndarray:
    """Inverse through the Cholesky factor, symmetrised (the native path's
    LLT solve of the identity, hex_q1.cpp:818-823)."""

    factor = np.linalg.cholesky(matrix)
    lower_inverse = np.linalg.solve(factor, np.eye(matrix.shape[0]))
    inverse = lower_inverse.T @ lower_inverse
    return 0.5 * (inverse + inverse.T)


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
