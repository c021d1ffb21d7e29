"""Two-level preconditioner of a layered node grid: Jacobi plus a coarse
correction in the space of functions constant over square node patches of
one node layer (reference: src/electrical/matrix_free_mpir_fem/two_level.py;
native construction hex_q1.cpp:703-835, application hex_q1.cpp:388-432).

``M^-1 r = D^-1 r + Z (Z^T A Z)^-1 Z^T r`` with ``Z`` the patch indicator on
free nodes.  The coarse matrix is built here, on the host, from FP64
operator applications on the device: the patches are coloured so that two
patches of one colour never touch a common element (nor, on a sheet
operator, a common via), so one application per colour gives every
coarse-matrix entry between a patch and its neighbour of that colour.  Its
Cholesky factor proves it SPD and gives the dense inverse (one
factorisation at a time per process), which moves to
the device once (``pcbhip_operator_set_coarse``).  Every inner PCG iteration
then runs the restriction, one dense row per coarse unknown and the
prolongation on the device, inside the batch's captured iterations.
"""

from __future__ import annotations

import threading
from typing import Callable

import numpy as np

DEFAULT_MAX_COARSE_SIZE = 2048

# One dense factorisation at a time in the process: each already uses every
# BLAS thread, and host threads that build operators together (a served
# electro-thermal loop per thread) would otherwise multiply the BLAS threads
# far past the cores.
_FACTOR_LOCK = threading.Lock()


def _spd_inverse(matrix: np.ndarray) -> np.ndarray:
    """Inverse of an SPD matrix, ``L^-T L^-1`` from its Cholesky factor,
    symmetrised against rounding; raises ValueError when the factorisation
    fails.  NumPy's own LAPACK only (a second BLAS library would start a
    second thread pool beside NumPy's)."""

    with _FACTOR_LOCK:
        try:
            factor = np.linalg.cholesky(matrix)
        except np.linalg.LinAlgError as exc:
            raise ValueError(
                "the coarse operator is not positive definite; the fine operator must be SPD for the two-level "
                "preconditioner"
            ) from exc
        lower_inverse = np.linalg.solve(factor, np.eye(matrix.shape[0]))
        inverse = lower_inverse.T @ lower_inverse
    return 0.5 * (inverse + inverse.T)


def _cells(extent: int, width: int) -> int:
    return -(-extent // width)


def choose_block_size(node_shape: tuple[int, int, int], *, max_coarse_size: int = DEFAULT_MAX_COARSE_SIZE,
                      minimum: int = 4) -> int:
    """The narrowest patch width, at least ``minimum``, whose coarse space has
    at most ``max_coarse_size`` unknowns; it stops growing once one patch
    spans the wider side of a layer (two_level.py:37-51)."""

    layers, rows, cols = (int(v) for v in node_shape)
    width = max(1, int(minimum))
    widest = max(rows, cols)
    while width < widest and layers * _cells(rows, width) * _cells(cols, width) > max_coarse_size:
        width += 1
    return width


class AggregationCoarseCorrection:
    """The coarse space of one operator and the dense inverse of its Galerkin
    matrix (two_level.py:54-229).

    ``node_shape`` is (node layers, node rows, node columns); ``apply_high``
    is the operator's FP64 action on a flat vector.  ``layer_colours`` is 3
    for a hexahedral grid (neighbouring node planes share elements) and the
    layer count for a sheet operator, whose layers touch only through vias
    (every layer its own colour, so a via between any two layers is seen);
    in-plane, a via's two ends must lie within one patch width of each
    other.  Patches made only of fixed nodes get a unit diagonal, as in the
    reference, so the matrix stays SPD; their restricted values are zero.
    ``galerkin(block, vectors)``, when given, returns ``Z^T A Z v`` for each
    row of ``vectors`` in one call (the device's FP64 transfers and action,
    ``pcbhip_operator_galerkin_f64``); otherwise each product is formed on
    the host around ``apply_high``.  ``invert(block, matrix)``, when given,
    returns the inverse of the finished coarse matrix (the device's
    factorisation, ``pcbhip_operator_set_coarse_matrix``, which also installs
    it); otherwise it is formed on the host."""

    def __init__(self, *, node_shape: tuple[int, int, int], free_nodes: np.ndarray, diagonal_high: np.ndarray,
                 apply_high: Callable[[np.ndarray], np.ndarray], block: int | None = None,
                 max_coarse_size: int = DEFAULT_MAX_COARSE_SIZE, layer_colours: int = 3,
                 galerkin: Callable[[int, np.ndarray], np.ndarray] | None = None,
                 invert: Callable[[int, np.ndarray], np.ndarray] | None = None) -> None:
        layers, rows, cols = (int(v) for v in node_shape)
        self.node_shape = (layers, rows, cols)
        if block is None:
            block = choose_block_size(self.node_shape, max_coarse_size=max_coarse_size)
        if int(block) < 1:
            raise ValueError("block must be a positive node count")
        self.block = int(block)
        self.coarse_shape = (layers, _cells(rows, self.block), _cells(cols, self.block))
        self.coarse_size = int(np.prod(self.coarse_shape))
        free = np.asarray(free_nodes, dtype=bool).reshape(self.node_shape)
        self._free = free
        diagonal = np.asarray(diagonal_high, dtype=np.float64).reshape(-1)
        if diagonal.size != free.size:
            raise ValueError("diagonal_high must hold one value per node")
        self._diagonal = diagonal
        if galerkin is None:
            galerkin = self._host_galerkin(apply_high)
        matrix = self._galerkin(galerkin, max(1, min(int(layer_colours), layers)))
        empty = self.restrict(free.astype(np.float64)).reshape(-1) <= 0.0
        matrix[empty, :] = 0.0
        matrix[:, empty] = 0.0
        matrix[empty, empty] = 1.0
        self.coarse_matrix = matrix
        self.coarse_inverse = _spd_inverse(matrix) if invert is None else np.asarray(invert(self.block, matrix))

    # ---------------------------------------------------------- transfers
    def restrict(self, grid: np.ndarray) -> np.ndarray:
        """``Z^T`` without the free mask: the sum of a node grid over each patch."""

        layers, rows, cols = self.node_shape
        _, crow, ccol = self.coarse_shape
        padded = np.zeros((layers, crow * self.block, ccol * self.block), dtype=np.float64)
        padded[:, :rows, :cols] = np.asarray(grid, dtype=np.float64).reshape(self.node_shape)
        return padded.reshape(layers, crow, self.block, ccol, self.block).sum(axis=(2, 4))

    def prolong(self, coarse: np.ndarray) -> np.ndarray:
        """``Z`` without the free mask: each patch value on all of its nodes."""

        _, rows, cols = self.node_shape
        values = np.asarray(coarse, dtype=np.float64).reshape(self.coarse_shape)
        spread = np.repeat(np.repeat(values, self.block, axis=1), self.block, axis=2)
        return spread[:, :rows, :cols]

    # ------------------------------------------------------- construction
    def _host_galerkin(self, apply_high: Callable[[np.ndarray], np.ndarray]):
        """``Z^T A Z v`` for rows ``v``, around a host FP64 action."""

        free = self._free

        def products(block: int, vectors: np.ndarray) -> np.ndarray:
            out = []
            for v in np.asarray(vectors, dtype=np.float64):
                fine = np.where(free, self.prolong(v), 0.0)
                action = np.asarray(apply_high(fine.reshape(-1))).reshape(self.node_shape)
                out.append(self.restrict(np.where(free, action, 0.0)).reshape(-1))
            return np.asarray(out)

        return products

    def _galerkin(self, galerkin: Callable[[int, np.ndarray], np.ndarray], layer_colours: int) -> np.ndarray:
        """``Z^T A Z`` from one FP64 application per patch colour.

        A patch's colour is (layer mod ``layer_colours``, row mod 3, column
        mod 3).  Patches of one colour are three patches apart in-plane (and
        three layers apart, or in different layers of a sheet operator), so
        nothing couples two of them, and a patch I meets at most one patch of
        each colour: the sum over I of A applied to that colour's indicator
        is the single entry (I, J) for that patch J."""

        index = np.indices(self.coarse_shape)
        flat = np.arange(self.coarse_size).reshape(self.coarse_shape)
        periods = (layer_colours, 3, 3)
        colours, indicators = [], []
        for colour in np.ndindex(*periods):
            chosen = np.ones(self.coarse_shape, dtype=bool)
            for axis in range(3):
                chosen &= index[axis] % periods[axis] == colour[axis]
            if chosen.any():
                colours.append(colour)
                indicators.append(chosen.reshape(-1).astype(np.float64))
        # The last product checks the whole matrix: Z^T A Z v for a fixed
        # pseudo-random v.  An entry assigned to the wrong patch pair (a
        # coupling the colouring does not separate) shows up there.
        v = np.random.default_rng(12345).uniform(0.5, 1.5, self.coarse_size)
        products = np.asarray(galerkin(self.block, np.stack(indicators + [v])))
        matrix = np.zeros((self.coarse_size, self.coarse_size))
        for colour, sums in zip(colours, products[:-1]):
            sums = sums.reshape(self.coarse_shape)
            # The colour's patch next to (or at) each patch: in-plane, the
            # offset in {-1, 0, 1} that lands on the colour's residue; across
            # layers the same with period 3, or the colour's own layer when
            # every layer is a colour of its own.
            partner = []
            valid = np.ones(self.coarse_shape, dtype=bool)
            for axis in range(3):
                if axis == 0 and periods[0] == self.coarse_shape[0]:
                    partner.append(np.full(self.coarse_shape, colour[0]))
                    continue
                step = (colour[axis] - index[axis]) % 3
                target = index[axis] + np.where(step == 2, -1, step)
                valid &= (target >= 0) & (target < self.coarse_shape[axis])
                partner.append(np.clip(target, 0, self.coarse_shape[axis] - 1))
            matrix[flat[valid], flat[tuple(p[valid] for p in partner)]] = sums[valid]
        direct = products[-1]
        if float(np.max(np.abs(matrix @ v - direct))) > 1e-9 * max(float(np.max(np.abs(matrix) @ np.abs(v))), 1e-300):
            raise ValueError("the coloured coarse assembly does not reproduce Z^T A Z: some coupling spans more "
                             "patches than the colouring separates; use preconditioner='jacobi'")
        return matrix

    # ------------------------------------------------------------ actions
    def apply_high(self, vector: np.ndarray) -> np.ndarray:
        """FP64 host action of the whole preconditioner, for verification."""

        flat = np.asarray(vector, dtype=np.float64).reshape(-1)
        masked = np.where(self._free, flat.reshape(self.node_shape), 0.0)
        coarse = self.coarse_inverse @ self.restrict(masked).reshape(-1)
        correction = np.where(self._free, self.prolong(coarse), 0.0)
        return flat / self._diagonal + correction.reshape(-1)


__all__ = ["AggregationCoarseCorrection", "DEFAULT_MAX_COARSE_SIZE", "choose_block_size"]
