"""ORACLE -- test infrastructure only.  The hexahedral Q1 operator assembled
element by element into a SciPy sparse matrix.

This is the independent FP64 check of every matrix-free action: the element
stiffness ``K_e = a_x U_x + a_y U_y + a_z U_z`` (mesh.py:276-293, 346-359) is
scattered to its corner nodes ``4 dz + 2 dy + dx`` (mesh.py:28-37) and summed,
Robin conductance goes on the diagonal and the rows and columns of fixed
nodes are replaced by the identity (operator.py:271-287).  Memory grows as
64 entries per element, so it serves problems of up to a few hundred
thousand elements.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def element_nodes(element_shape):
    """(E, 8) flat node index of every element corner."""

    s, r, c = element_shape
    nr, nc = r + 1, c + 1
    ez, ey, ex = np.meshgrid(np.arange(s), np.arange(r), np.arange(c), indexing="ij")
    base = ((ez * nr + ey) * nc + ex).reshape(-1)
    offs = np.array([(k >> 2) * nr * nc + ((k >> 1) & 1) * nc + (k & 1) for k in range(8)])
    return base[:, None] + offs[None, :]


def stiffness_matrix(coef, unit, element_shape):
    """Unconstrained conduction matrix ``K`` (every node, no Robin)."""

    s, r, c = element_shape
    n = (s + 1) * (r + 1) * (c + 1)
    nodes = element_nodes(element_shape)
    a = coef.reshape(3, -1)
    ke = a[0][:, None, None] * unit[0] + a[1][:, None, None] * unit[1] + a[2][:, None, None] * unit[2]
    rows = np.repeat(nodes, 8, axis=1).reshape(-1)
    cols = np.tile(nodes, (1, 8)).reshape(-1)
    return sp.csr_matrix((ke.reshape(-1), (rows, cols)), shape=(n, n))


def system_matrix(coef, unit, robin_total, free, element_shape):
    """``A``: free block of ``K`` plus Robin, identity on fixed rows."""

    k = stiffness_matrix(coef, unit, element_shape)
    f = sp.diags(free.astype(float))
    fixed = sp.diags((~free).astype(float))
    return (f @ k @ f + sp.diags(robin_total * free) + fixed).tocsr()


def solve_direct(a, b):
    """FP64 sparse direct solve of ``A x = b``."""

    return spla.spsolve(a.tocsc(), b)
