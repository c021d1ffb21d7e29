"""Test infrastructure only (the oracle): an element-assembled FP64 restatement
of the reference's layered-sheet DC operator, for parity checks of the device
kernels.  Nothing in the product imports this module.

Follows src/electrical/matrix_free_mpir_fem/pcb.py:
  * element coefficients c_x = sigma t h_y / h_x, c_y = sigma t h_x / h_y,
    zero on inactive elements (pcb.py:173-186);
  * unit sheet matrices U_x = M (x) S, U_y = S (x) M, local order 2 dy + dx
    (pcb.py:249-256);
  * the action: element products on the free-masked vector, via links
    g (w_a - w_b) added at both ends, identity on fixed rows (pcb.py:575-604);
  * fixed nodes: inactive nodes (no active element and no via), the
    reference node and the Dirichlet nodes (pcb.py:356-395).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_S = np.array([[1.0, -1.0], [-1.0, 1.0]])
_M = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0


def unit_matrices() -> np.ndarray:
    return np.stack((np.kron(_M, _S), np.kron(_S, _M)))


def coefficients(active, thickness, pitch_x, pitch_y, sigma) -> np.ndarray:
    """(2, layers, rows, cols) sheet coefficients."""

    active = np.asarray(active, dtype=bool)
    sheet = np.asarray(sigma, dtype=float) * np.asarray(thickness, dtype=float)[:, None, None] * active
    hx = np.broadcast_to(np.asarray(pitch_x, dtype=float), (active.shape[2],))[None, None, :]
    hy = np.broadcast_to(np.asarray(pitch_y, dtype=float), (active.shape[1],))[None, :, None]
    return np.stack((sheet * hy / hx, sheet * hx / hy))


def element_nodes(element_shape) -> np.ndarray:
    """(E, 4) flat node indices of every element, corners 2 dy + dx."""

    layers, rows, cols = element_shape
    node_shape = (layers, rows + 1, cols + 1)
    l, y, x = np.meshgrid(np.arange(layers), np.arange(rows), np.arange(cols), indexing="ij")
    corners = [np.ravel_multi_index((l, y + dy, x + dx), node_shape).reshape(-1) for dy in (0, 1) for dx in (0, 1)]
    return np.stack(corners, axis=1)


def stiffness(coef, element_shape, via_a=(), via_b=(), via_g=()) -> sp.csr_matrix:
    """The unconstrained conductance matrix K (every node free)."""

    layers, rows, cols = element_shape
    n = layers * (rows + 1) * (cols + 1)
    nodes = element_nodes(element_shape)
    unit = unit_matrices()
    c = coef.reshape(2, -1)
    local = c[0][:, None, None] * unit[0] + c[1][:, None, None] * unit[1]  # (E, 4, 4)
    rows_i = np.repeat(nodes, 4, axis=1).reshape(-1)
    cols_i = np.tile(nodes, (1, 4)).reshape(-1)
    k = sp.coo_matrix((local.reshape(-1), (rows_i, cols_i)), shape=(n, n)).tocsr()
    a = np.asarray(via_a, dtype=np.int64)
    b = np.asarray(via_b, dtype=np.int64)
    g = np.asarray(via_g, dtype=float)
    if a.size:
        v = sp.coo_matrix((np.concatenate((g, g, -g, -g)),
                           (np.concatenate((a, b, a, b)), np.concatenate((a, b, b, a)))), shape=(n, n)).tocsr()
        k = k + v
    return k.tocsr()


def active_nodes(active, via_a=(), via_b=()) -> np.ndarray:
    active = np.asarray(active, dtype=bool)
    layers, rows, cols = active.shape
    out = np.zeros((layers, rows + 1, cols + 1), dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            out[:, dy:dy + rows, dx:dx + cols] |= active
    out.flat[np.asarray(via_a, dtype=np.int64)] = True
    out.flat[np.asarray(via_b, dtype=np.int64)] = True
    return out


def system_matrix(k: sp.csr_matrix, free: np.ndarray) -> sp.csr_matrix:
    """Free block of K with the identity on fixed rows and columns."""

    f = sp.diags(free.reshape(-1).astype(float))
    fixed = sp.diags((~free.reshape(-1)).astype(float))
    return (f @ k @ f + fixed).tocsr()


def solve(k: sp.csr_matrix, free: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    return spla.spsolve(system_matrix(k, free).tocsc(), rhs)


def element_view_apply(coef, element_shape, x, free, via_a=(), via_b=(), via_g=()) -> np.ndarray:
    """The action by element views (the reference's loop form, pcb.py:575-604),
    an independent restatement used to pin the assembly."""

    layers, rows, cols = element_shape
    grid = np.where(free, np.asarray(x, dtype=float).reshape(layers, rows + 1, cols + 1), 0.0)
    views = (grid[:, :-1, :-1], grid[:, :-1, 1:], grid[:, 1:, :-1], grid[:, 1:, 1:])
    out = np.zeros_like(grid)
    targets = (out[:, :-1, :-1], out[:, :-1, 1:], out[:, 1:, :-1], out[:, 1:, 1:])
    unit = unit_matrices()
    for i in range(4):
        acc = np.zeros_like(coef[0])
        for j in range(4):
            acc = acc + (coef[0] * unit[0, i, j] + coef[1] * unit[1, i, j]) * views[j]
        targets[i][...] += acc
    flat_in = grid.reshape(-1)
    flat = out.reshape(-1)
    a = np.asarray(via_a, dtype=np.int64)
    if a.size:
        b = np.asarray(via_b, dtype=np.int64)
        delta = np.asarray(via_g, dtype=float) * (flat_in[a] - flat_in[b])
        np.add.at(flat, a, delta)
        np.add.at(flat, b, -delta)
    return np.where(free.reshape(-1), flat, np.asarray(x, dtype=float).reshape(-1))


def port_conductance(k: sp.csr_matrix, active: np.ndarray, pads) -> tuple[np.ndarray, np.ndarray]:
    """Conductance matrix of the N-port whose ports are the node sets ``pads``
    (flat indices), by one sparse LU of the free block (the Schur complement
    ``K_pp - K_pf K_ff^-1 K_fp`` summed over each pad), and the potential for
    each port at 1 V with the others at 0 V.  Follows the definition of
    src/electrical/matrix_free_mpir_fem/ports.py:206-215 (G[j, k] is the
    current into port j for port k at 1 V); free nodes are the active nodes
    outside every pad."""

    n = k.shape[0]
    on_pad = np.zeros(n, dtype=bool)
    for pad in pads:
        on_pad[np.asarray(pad, dtype=np.int64)] = True
    free = np.asarray(active, dtype=bool).reshape(-1) & ~on_pad
    fi = np.flatnonzero(free)
    lu = spla.splu(k[fi][:, fi].tocsc())
    count = len(pads)
    g = np.zeros((count, count))
    fields = np.zeros((count, n))
    for col, pad in enumerate(pads):
        x = np.zeros(n)
        x[np.asarray(pad, dtype=np.int64)] = 1.0
        x[fi] = lu.solve(-(k[fi] @ x))
        injected = k @ x
        fields[col] = x
        for row, other in enumerate(pads):
            g[row, col] = injected[np.asarray(other, dtype=np.int64)].sum()
    return g, fields


def element_loss(coef, element_shape, x) -> np.ndarray:
    """Per-element ``c_x v^T U_x v + c_y v^T U_y v`` of the field ``x``."""

    v = np.asarray(x, dtype=float).reshape(-1)[element_nodes(element_shape)]
    unit = unit_matrices()
    c = coef.reshape(2, -1)
    loss = c[0] * np.einsum("ei,ij,ej->e", v, unit[0], v) + c[1] * np.einsum("ei,ij,ej->e", v, unit[1], v)
    return loss.reshape(element_shape)
