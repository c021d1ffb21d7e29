"""Seeded synthetic boards: the workloads the tests and the benchmark run.

Every board is a 0.25 mm-pitch hexahedral stack, 1.6 mm thick split evenly
into ``S`` slabs, FR-4 (0.3 W/m K, isotropic) with copper (385 W/m K) on the
bottom and top slabs under a random occupancy mask of 4 x 4-cell blocks,
the bottom node plane held at 298.15 K and the top face cooled by h = 10
W/m^2 K to 298.15 K.  Heat enters as element heat spread uniformly over
square patches of top-slab cells.  Draw order is fixed (bottom mask, top
mask, then sources), so a seed names a board.

* ``board_64x64x8``       -- 8 x 64 x 64 elements, 40 % copper, 1 W on a 4 x 4 patch (seed 1).
* ``board_512x512x16``    -- 16 x 512 x 512, 30-60 % copper per plane, 8 sources of
                             0.1-2 W (log-uniform) on 3..8-cell patches (seed 2).
* ``uneven_boards``       -- 200 boards, sides log-uniform 32..384, depth 4..16,
                             20-70 % copper, 1..8 sources of 0.1-3 W (seed 4).
* ``dc_board``, ``serving_pool`` -- the DC copper/dielectric board and the serving mix.
* ``coupling_scenarios``  -- 16 voltage-driven electro-thermal current loops (seed 6).
* ``four_slab_fixture``   -- the reference's own thermal benchmark stack (no randomness).
* ``general_board``       -- the 512 x 512 x 16 geometry with every element's conductivity
                             jittered and every face cooled (seed 7): no node plane is
                             free of Robin terms and no element layer is uniform.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from thermal.matrix_free_mpir_fem import ConvectionBoundary, LayeredThermalMesh, ThermalConductionProblem

PITCH_M = 0.25e-3
STACK_M = 1.6e-3
FR4_K = 0.3
COPPER_K = 385.0
AMBIENT_K = 298.15
FILM_W_PER_M2_K = 10.0


def copper_mask(rng: np.random.Generator, rows: int, cols: int, fill: float) -> np.ndarray:
    """Occupancy of 4 x 4-cell blocks, cropped to (rows, cols)."""

    blocks = rng.random((-(-rows // 4), -(-cols // 4))) < fill
    return np.kron(blocks, np.ones((4, 4), dtype=bool))[:rows, :cols]


@dataclass(frozen=True)
class Source:
    row: int
    col: int
    side: int
    power_w: float


def layered_board(slabs: int, rows: int, cols: int, masks: tuple[np.ndarray, np.ndarray],
                  sources: list[Source]) -> ThermalConductionProblem:
    k = np.full((slabs, rows, cols), FR4_K)
    k[0] = np.where(masks[0], COPPER_K, FR4_K)
    k[-1] = np.where(masks[1], COPPER_K, FR4_K)
    mesh = LayeredThermalMesh(
        slab_thickness_m=(STACK_M / slabs,) * slabs,
        pitch_x_m=PITCH_M,
        pitch_y_m=PITCH_M,
        conductivity_w_per_m_k=k,
    )
    heat = np.zeros(mesh.element_grid_shape)
    for s in sources:
        heat[-1, s.row : s.row + s.side, s.col : s.col + s.side] += s.power_w / (s.side * s.side)
    fixed = np.zeros(mesh.node_shape, dtype=bool)
    fixed[0] = True
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", FILM_W_PER_M2_K, AMBIENT_K),),
        fixed_temperature_mask=fixed,
        fixed_temperature_k=AMBIENT_K,
        element_heat_w=heat,
    )


def board_64x64x8(seed: int = 1) -> ThermalConductionProblem:
    rng = np.random.default_rng(seed)
    masks = (copper_mask(rng, 64, 64, 0.40), copper_mask(rng, 64, 64, 0.40))
    row, col = (int(v) for v in rng.integers(0, 61, 2))
    return layered_board(8, 64, 64, masks, [Source(row, col, 4, 1.0)])


def random_sources(rng: np.random.Generator, rows: int, cols: int, count: int, p_lo: float, p_hi: float,
                   side_lo: int = 3, side_hi: int = 8) -> list[Source]:
    out = []
    for _ in range(count):
        power = float(np.exp(rng.uniform(np.log(p_lo), np.log(p_hi))))
        side = int(rng.integers(side_lo, side_hi + 1))
        side = min(side, rows, cols)
        row = int(rng.integers(0, rows - side + 1))
        col = int(rng.integers(0, cols - side + 1))
        out.append(Source(row, col, side, power))
    return out


def four_slab_fixture(elements: int) -> ThermalConductionProblem:
    """The stack of the reference's native thermal benchmark
    (experiments/thermal_native_benchmark.py:43-61): 35 um Cu / 0.7 mm FR-4
    (0.8 in-plane, 0.3 through) / 0.7 mm FR-4 / 35 um Cu over 50 x 50 mm in
    ``elements`` x ``elements`` cells, h = 10 W/m^2 K on both faces, 1 W on a
    trace across the middle half of the top slab.  50, 100 and 200 elements
    give the 13,005, 51,005 and 202,005 nodes of its published timings."""

    mesh = LayeredThermalMesh(
        slab_thickness_m=(35.0e-6, 0.7e-3, 0.7e-3, 35.0e-6),
        pitch_x_m=50.0e-3 / elements,
        pitch_y_m=50.0e-3 / elements,
        conductivity_w_per_m_k=(COPPER_K, 0.8, 0.8, COPPER_K),
        through_plane_conductivity_w_per_m_k=(COPPER_K, 0.3, 0.3, COPPER_K),
        element_shape=(elements, elements),
    )
    heat = np.zeros(mesh.element_grid_shape)
    heat[3, elements // 2, elements // 4: 3 * elements // 4] = 1.0 / (elements // 2)
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", FILM_W_PER_M2_K, AMBIENT_K),
                    ConvectionBoundary("bottom", FILM_W_PER_M2_K, AMBIENT_K)),
        element_heat_w=heat,
    )


def board_512x512x16(seed: int = 2, size: int = 512, slabs: int = 16) -> ThermalConductionProblem:
    rng = np.random.default_rng(seed)
    fills = rng.uniform(0.3, 0.6, 2)
    masks = (copper_mask(rng, size, size, fills[0]), copper_mask(rng, size, size, fills[1]))
    return layered_board(slabs, size, size, masks, random_sources(rng, size, size, 8, 0.1, 2.0))


def general_board(seed: int = 7, size: int = 512, slabs: int = 16) -> ThermalConductionProblem:
    """A board on which no stencil shortcut applies: the copper masks and
    sources of ``board_512x512x16``'s kind, every element's in-plane and
    through-plane conductivity scaled by its own U(0.8, 1.25) draw (so no
    element layer is uniform), h = 10 W/m^2 K on the top and bottom faces and
    on the four sides (so every node plane holds Robin terms), and no fixed
    node."""

    from thermal.matrix_free_mpir_fem import ExposedFaceConvection

    rng = np.random.default_rng(seed)
    fills = rng.uniform(0.3, 0.6, 2)
    masks = (copper_mask(rng, size, size, fills[0]), copper_mask(rng, size, size, fills[1]))
    sources = random_sources(rng, size, size, 8, 0.1, 2.0)
    k = np.full((slabs, size, size), FR4_K)
    k[0] = np.where(masks[0], COPPER_K, FR4_K)
    k[-1] = np.where(masks[1], COPPER_K, FR4_K)
    k_in = k * rng.uniform(0.8, 1.25, k.shape)
    k_z = k * rng.uniform(0.8, 1.25, k.shape)
    mesh = LayeredThermalMesh(
        slab_thickness_m=(STACK_M / slabs,) * slabs,
        pitch_x_m=PITCH_M,
        pitch_y_m=PITCH_M,
        conductivity_w_per_m_k=k_in,
        through_plane_conductivity_w_per_m_k=k_z,
    )
    heat = np.zeros(mesh.element_grid_shape)
    for s in sources:
        heat[-1, s.row : s.row + s.side, s.col : s.col + s.side] += s.power_w / (s.side * s.side)
    return ThermalConductionProblem(
        mesh,
        convection=(ConvectionBoundary("top", FILM_W_PER_M2_K, AMBIENT_K),
                    ConvectionBoundary("bottom", FILM_W_PER_M2_K, AMBIENT_K),
                    ExposedFaceConvection(FILM_W_PER_M2_K, AMBIENT_K, ("-x", "+x", "-y", "+y"))),
        element_heat_w=heat,
    )


def uneven_board_shapes(count: int = 200, seed: int = 4) -> list[tuple[int, int, int]]:
    """(slabs, rows, cols) of the uneven batch, without building the boards."""

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        rows, cols = (int(round(float(np.exp(v)))) for v in rng.uniform(np.log(32), np.log(384), 2))
        slabs = int(rng.integers(4, 17))
        out.append((slabs, rows, cols))
    return out


def uneven_boards(count: int = 200, seed: int = 4) -> list[ThermalConductionProblem]:
    shapes = uneven_board_shapes(count, seed)
    rng = np.random.default_rng(seed + 1000)
    boards = []
    for slabs, rows, cols in shapes:
        fills = rng.uniform(0.2, 0.7, 2)
        masks = (copper_mask(rng, rows, cols, fills[0]), copper_mask(rng, rows, cols, fills[1]))
        n_src = int(rng.integers(1, 9))
        boards.append(layered_board(slabs, rows, cols, masks, random_sources(rng, rows, cols, n_src, 0.1, 3.0)))
    return boards


COPPER_SIGMA_S_PER_M = 5.8e7
DIELECTRIC_SIGMA_S_PER_M = 1.0e-12


@dataclass(frozen=True)
class DCBoard:
    """A volumetric DC-conduction board: mesh (conductivity in S/m), the
    terminal-pair excitations, the gauge node and the top-slab copper cluster
    the terminals sit in (element mask)."""

    mesh: LayeredThermalMesh
    excitations: tuple
    reference_node: tuple[int, int, int]
    cluster: np.ndarray


def dc_board(seed: int = 3, size: int = 256, slabs: int = 12, pairs: int = 64, rows: int | None = None) -> DCBoard:
    """Copper (5.8e7 S/m) on the bottom and top slabs at U(0.3, 0.6) block
    fill, 1e-12 S/m elsewhere; ``pairs`` excitations of +1 A / -1 A on two
    3 x 3-cell fully-copper patches of the top slab, both inside the largest
    connected copper cluster of that slab (copper islands that float would
    leave the potential undetermined), 1/16 A on each of a patch's 16
    top-face nodes; the gauge is the cluster's first top-face node."""

    from scipy import ndimage

    from electrical.matrix_free_mpir_fem.hex_dc import CurrentTerminal

    rows = size if rows is None else rows
    rng = np.random.default_rng(seed)
    fills = rng.uniform(0.3, 0.6, 2)
    masks = (copper_mask(rng, rows, size, fills[0]), copper_mask(rng, rows, size, fills[1]))
    sigma = np.full((slabs, rows, size), DIELECTRIC_SIGMA_S_PER_M)
    sigma[0] = np.where(masks[0], COPPER_SIGMA_S_PER_M, DIELECTRIC_SIGMA_S_PER_M)
    sigma[-1] = np.where(masks[1], COPPER_SIGMA_S_PER_M, DIELECTRIC_SIGMA_S_PER_M)
    mesh = LayeredThermalMesh(slab_thickness_m=(STACK_M / slabs,) * slabs, pitch_x_m=PITCH_M, pitch_y_m=PITCH_M,
                              conductivity_w_per_m_k=sigma)
    # Elements sharing a node conduct into each other: 8-connectivity.
    labels, count = ndimage.label(masks[1], structure=np.ones((3, 3), dtype=int))
    sizes = ndimage.sum(np.ones_like(labels), labels, index=np.arange(1, count + 1))
    cluster = labels == (1 + int(np.argmax(sizes)))
    # 3 x 3 windows made only of cluster elements.
    inside = ndimage.minimum_filter(cluster.astype(np.uint8), size=3, mode="constant", cval=0)[1:-1, 1:-1] > 0
    rows_ok, cols_ok = np.nonzero(inside)  # window top-left corners (row, col)
    top = slabs
    cells = np.argwhere(cluster)
    ref = (top, int(cells[0][0]), int(cells[0][1]))
    excitations = []
    for _ in range(pairs):
        while True:
            a, b = rng.choice(rows_ok.size, 2, replace=False)
            pa = (int(rows_ok[a]), int(cols_ok[a]))
            pb = (int(rows_ok[b]), int(cols_ok[b]))
            nodes_a = tuple((top, pa[0] + i, pa[1] + j) for i in range(4) for j in range(4))
            nodes_b = tuple((top, pb[0] + i, pb[1] + j) for i in range(4) for j in range(4))
            if ref in nodes_a or ref in nodes_b or set(nodes_a) & set(nodes_b):
                continue
            break
        excitations.append((CurrentTerminal(nodes_a, 1.0, "source"), CurrentTerminal(nodes_b, -1.0, "sink")))
    return DCBoard(mesh, tuple(excitations), ref, cluster)


@dataclass
class PoolBoard:
    """One board of the serving pool: its resident device operator, up to
    ``max_rhs`` right-hand sides and the front-end operator that built them."""

    kind: str
    operator: object  # _capi.Operator
    rhs: list
    front_end: object


def element_heat_load(mesh: LayeredThermalMesh, element_heat_w: np.ndarray) -> np.ndarray:
    """Element heat lumped an eighth onto each corner (the nodal load of
    element heat, reference operator.py:374-387)."""

    from thermal.matrix_free_mpir_fem.mesh import corner_views

    load = np.zeros(mesh.node_shape)
    eighth = np.asarray(element_heat_w, dtype=np.float64) / 8.0
    for view in corner_views(load):
        view[...] += eighth
    return load.reshape(-1)


def serving_pool(seed: int = 5, thermal_boards: int = 16, dc_boards: int = 8, max_rhs: int = 16,
                 max_side: int = 384) -> list[PoolBoard]:
    """Boards for the serving stream: thermal boards of the uneven-batch size
    and material distribution whose right-hand sides are independent draws of
    1..8 heat sources, and DC boards of the copper/dielectric materials on
    the same size distribution whose right-hand sides are terminal pairs in
    the largest copper cluster.  Operators are built on device 0."""

    from electrical.matrix_free_mpir_fem.hex_dc import MatrixFreeHexDCOperator
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    rng = np.random.default_rng(seed)
    pool: list[PoolBoard] = []
    kinds = ["thermal"] * thermal_boards + ["dc"] * dc_boards
    for kind in kinds:
        rows, cols = (min(max_side, int(round(float(np.exp(v))))) for v in rng.uniform(np.log(32), np.log(384), 2))
        slabs = int(rng.integers(4, 17))
        if kind == "thermal":
            fills = rng.uniform(0.2, 0.7, 2)
            masks = (copper_mask(rng, rows, cols, fills[0]), copper_mask(rng, rows, cols, fills[1]))
            problem = layered_board(slabs, rows, cols, masks, random_sources(rng, rows, cols, 1, 0.1, 3.0))
            op = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
            reference = op.default_reference_temperature()
            base = op.build_rhs(reference) - op.nodal_load()
            free = op.free_nodes.reshape(-1)
            rhs = []
            for _ in range(max_rhs):
                heat = np.zeros(problem.mesh.element_grid_shape)
                for s in random_sources(rng, rows, cols, int(rng.integers(1, 9)), 0.1, 3.0):
                    heat[-1, s.row:s.row + s.side, s.col:s.col + s.side] += s.power_w / (s.side * s.side)
                rhs.append(np.where(free, base + element_heat_load(problem.mesh, heat), base))
            pool.append(PoolBoard(kind, op.device_operator, rhs, op))
        else:
            board = dc_board(seed=int(rng.integers(1 << 30)), size=cols, rows=rows, slabs=slabs, pairs=max_rhs)
            op = MatrixFreeHexDCOperator(board.mesh, reference_node=board.reference_node)
            pool.append(PoolBoard(kind, op.device_operator, [op.build_rhs(t) for t in board.excitations], op))
    return pool


def coupling_scenarios(count: int = 16, seed: int = 6):
    """Electro-thermal scenarios of the served coupling workload (seed 6).

    Each board is a current loop on two 35 um copper layers at 0.5 mm pitch,
    ``rows`` x ``cols`` elements (rows 8..48, cols 32..256, log-uniform),
    copper in every row but the outer two, fed by two voltage pads at x = 0
    (top layer source, bottom layer return) and closed by vias in the last
    two node columns; the voltage is set for a cold current of 1..4 A
    (log-uniform) from the strip resistance.  The thermal mesh is copper /
    1.5 mm laminate / copper with films of 10 W/m^2 K on both faces."""

    from electrical.matrix_free_mpir_fem import LayeredPCBMesh, PCBConductionProblem, ViaConnection, VoltageTerminal
    from multiphysics.staggered_coupling import ElectroThermalScenario

    rng = np.random.default_rng(seed)
    pitch, thickness, sigma, via_r = 0.5e-3, 35e-6, 1.0 / 1.724e-8, 1.5e-3
    out = []
    for _ in range(count):
        rows = int(round(np.exp(rng.uniform(np.log(8), np.log(48)))))
        cols = int(round(np.exp(rng.uniform(np.log(32), np.log(256)))))
        current = float(np.exp(rng.uniform(np.log(1.0), np.log(4.0))))
        active = np.zeros((2, rows, cols), dtype=bool)
        active[:, 1:rows - 1, :] = True
        mesh = LayeredPCBMesh(active, (thickness, thickness), pitch, pitch, conductivity_s_per_m=sigma)
        pad = tuple(range(1, rows))
        vias = tuple(ViaConnection((0, r, c), (1, r, c), via_r) for r in pad[1:-1] for c in (cols - 1, cols))
        strip = cols / (sigma * thickness * (rows - 2))
        voltage = current * (2.0 * strip + via_r / len(vias))
        problem = PCBConductionProblem(mesh, voltage_terminals=(
            VoltageTerminal(tuple((1, r, 0) for r in pad), voltage, "source"),
            VoltageTerminal(tuple((0, r, 0) for r in pad), 0.0, "return")), vias=vias)
        k = np.full((3, rows, cols), 0.8)
        k[0] = np.where(active[0], 385.0, 0.8)
        k[2] = np.where(active[1], 385.0, 0.8)
        kz = k.copy()
        kz[1] = 0.3
        thermal = LayeredThermalMesh((thickness, 1.5e-3, thickness), pitch, pitch, k,
                                     through_plane_conductivity_w_per_m_k=kz)
        films = (ConvectionBoundary("top", FILM_W_PER_M2_K, AMBIENT_K),
                 ConvectionBoundary("bottom", FILM_W_PER_M2_K, AMBIENT_K))
        out.append(ElectroThermalScenario(problem, thermal, (0, 2), convection=films))
    return out
