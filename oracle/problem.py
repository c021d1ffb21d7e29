"""ORACLE -- test infrastructure only.  FP64 NumPy restatement of the set-up
and post-processing of the reference's hexahedral Q1 conduction solve.

A problem is a plain record of arrays (``OracleProblem``); nothing here
imports the product package.  Each function cites the reference code it
restates (paths relative to the reference repository root).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

FACE_DIRECTIONS = ("-x", "+x", "-y", "+y", "-z", "+z")
# Local corners ``4 dz + 2 dy + dx`` of each face (mesh.py:45-53).
FACE_CORNERS = {
    "-z": (0, 1, 2, 3),
    "+z": (4, 5, 6, 7),
    "-y": (0, 1, 4, 5),
    "+y": (2, 3, 6, 7),
    "-x": (0, 2, 4, 6),
    "+x": (1, 3, 5, 7),
}


@dataclass
class Convection:
    """One convection boundary: ``kind`` is "top", "bottom" or "exposed".

    ``coefficient`` / ``ambient`` are scalars, ``(rows, cols)`` arrays for a
    top/bottom face, or ``(slabs, rows, cols)`` for exposed faces."""

    kind: str
    coefficient: object
    ambient: object
    directions: tuple = FACE_DIRECTIONS


@dataclass
class OracleProblem:
    thickness: np.ndarray  # (S,)
    pitch_x: np.ndarray  # (C,)
    pitch_y: np.ndarray  # (R,)
    k_in: np.ndarray  # (S, R, C), zero on void
    k_z: np.ndarray  # (S, R, C)
    active: np.ndarray  # (S, R, C) bool
    fixed_mask: np.ndarray  # node shape bool
    fixed_values: np.ndarray  # node shape
    convection: list = field(default_factory=list)
    element_heat: np.ndarray | None = None  # (S, R, C)
    nodal_heat: np.ndarray | None = None  # node shape
    sources: list = field(default_factory=list)  # [(flat node indices, power)]

    @property
    def element_shape(self):
        return tuple(int(v) for v in self.k_in.shape)

    @property
    def node_shape(self):
        s, r, c = self.element_shape
        return (s + 1, r + 1, c + 1)

    @property
    def size(self):
        return int(np.prod(self.node_shape))


def corner_views(grid):
    """Eight corner views ordered ``4 dz + 2 dy + dx`` (mesh.py:28-37)."""

    s, r, c = (a - 1 for a in grid.shape)
    return [grid[dz : dz + s, dy : dy + r, dx : dx + c] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]


def unit_matrices():
    """``U_x = M(z) M(y) S(x)``, ``U_y = M S M``, ``U_z = S M M`` (mesh.py:342-359)."""

    stiff = np.array([[1.0, -1.0], [-1.0, 1.0]])
    mass = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    unit = np.zeros((3, 8, 8))
    for i in range(8):
        zi, yi, xi = i >> 2, (i >> 1) & 1, i & 1
        for j in range(8):
            zj, yj, xj = j >> 2, (j >> 1) & 1, j & 1
            unit[0, i, j] = mass[zi, zj] * (mass[yi, yj] * stiff[xi, xj])
            unit[1, i, j] = mass[zi, zj] * (stiff[yi, yj] * mass[xi, xj])
            unit[2, i, j] = stiff[zi, zj] * (mass[yi, yj] * mass[xi, xj])
    return unit


def element_coefficients(p: OracleProblem):
    """``a_x = k_in hy hz / hx``, ``a_y = k_in hx hz / hy``, ``a_z = k_z hx hy / hz``
    (mesh.py:276-293; thermal_system.cpp:264-277)."""

    hz = np.asarray(p.thickness, float)[:, None, None]
    hx = np.asarray(p.pitch_x, float)[None, None, :]
    hy = np.asarray(p.pitch_y, float)[None, :, None]
    k_in = np.where(p.active, p.k_in, 0.0)
    k_z = np.where(p.active, p.k_z, 0.0)
    return np.stack(((k_in * hy) * hz / hx, (k_in * hx) * hz / hy, (k_z * hx) * hy / hz))


def active_nodes(p: OracleProblem):
    nodes = np.zeros(p.node_shape, dtype=bool)
    for view in corner_views(nodes):
        view |= p.active
    return nodes


def exposed_faces(active):
    """Active element faces bordering a void element or the grid edge (mesh.py:53-77)."""

    out = {}
    for direction in FACE_DIRECTIONS:
        axis = {"x": 2, "y": 1, "z": 0}[direction[1]]
        neighbour = np.zeros_like(active)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        if direction[0] == "+":
            dst[axis], src[axis] = slice(None, -1), slice(1, None)
        else:
            dst[axis], src[axis] = slice(1, None), slice(None, -1)
        neighbour[tuple(dst)] = active[tuple(src)]
        out[direction] = active & ~neighbour
    return out


def face_area(p: OracleProblem, direction):
    """Area of every element face facing ``direction`` (mesh.py:95-106)."""

    shape = p.element_shape
    hz = np.asarray(p.thickness, float)[:, None, None]
    hx = np.asarray(p.pitch_x, float)[None, None, :]
    hy = np.asarray(p.pitch_y, float)[None, :, None]
    if direction[1] == "z":
        return np.broadcast_to(hx * hy, shape).copy()
    if direction[1] == "y":
        return np.broadcast_to(hx * hz, shape).copy()
    return np.broadcast_to(hy * hz, shape).copy()


def lump_faces(face_weight, direction, node_shape):
    """A quarter of each face weight onto each of its four nodes (mesh.py:80-92)."""

    weights = np.zeros(node_shape)
    views = corner_views(weights)
    share = face_weight / 4.0
    for corner in FACE_CORNERS[direction]:
        views[corner] += share
    return weights


def convection_weights(p: OracleProblem, b: Convection):
    """Nodal conductance ``R`` and load ``R T_amb`` of one boundary, flat
    (boundaries.py:82-102 for top/bottom, :170-182 for exposed faces)."""

    shape = p.element_shape
    if b.kind in ("top", "bottom"):
        slab = shape[0] - 1 if b.kind == "top" else 0
        direction = "+z" if b.kind == "top" else "-z"
        coef = np.broadcast_to(np.asarray(b.coefficient, float), shape[1:])
        amb = np.broadcast_to(np.asarray(b.ambient, float), shape[1:])
        face = np.zeros(shape)
        face[slab] = coef * p.active[slab]
        face *= face_area(p, direction)
        load = face.copy()
        load[slab] *= amb
        return (
            lump_faces(face, direction, p.node_shape).reshape(-1),
            lump_faces(load, direction, p.node_shape).reshape(-1),
        )
    faces = exposed_faces(p.active)
    coef = np.broadcast_to(np.asarray(b.coefficient, float), shape)
    amb = np.broadcast_to(np.asarray(b.ambient, float), shape)
    w = np.zeros(p.node_shape)
    rhs = np.zeros(p.node_shape)
    for direction in b.directions:
        face = coef * face_area(p, direction) * faces[direction]
        w += lump_faces(face, direction, p.node_shape)
        rhs += lump_faces(face * amb, direction, p.node_shape)
    return w.reshape(-1), rhs.reshape(-1)


def nodal_load(p: OracleProblem):
    """Element heat / 8 onto the corners, nodal heat, and each source's power
    shared over its nodes (operator.py:374-387)."""

    load = np.zeros(p.node_shape) if p.nodal_heat is None else np.array(p.nodal_heat, float).reshape(p.node_shape)
    if p.element_heat is not None:
        share = np.asarray(p.element_heat, float) / 8.0
        for view in corner_views(load):
            view += share
    flat = load.reshape(-1)
    for nodes, power in p.sources:
        nodes = np.asarray(nodes, dtype=np.int64).reshape(-1)
        per = float(power) / len(nodes)
        for node in nodes:
            flat[node] += per
    return flat


def operator_arrays(p: OracleProblem, capacity=None):
    """Everything the matrix-free operator holds (operator.py:84-146;
    thermal_system.cpp:252-363), FP64."""

    coef = element_coefficients(p)
    unit = unit_matrices()
    act = active_nodes(p)
    fixed = (np.asarray(p.fixed_mask, bool) | ~act).reshape(-1)
    free = ~fixed
    fixed_t = np.where((np.asarray(p.fixed_mask, bool) & act).reshape(-1), np.asarray(p.fixed_values, float).reshape(-1), 0.0)
    weights, rhs_weights, ambients = [], [], []
    for b in p.convection:
        w, rw = convection_weights(p, b)
        weights.append(w)
        rhs_weights.append(rw)
        ambients.append(float(np.mean(b.ambient)))
    robin_only = np.sum(weights, axis=0) if weights else np.zeros(p.size)
    cap = np.zeros(p.size) if capacity is None else np.asarray(capacity, float).reshape(-1)
    robin_total = robin_only + cap
    robin_rhs = np.sum(rhs_weights, axis=0) if rhs_weights else np.zeros(p.size)
    diag = diagonal(coef, unit, robin_total, free, p.node_shape)
    return dict(
        coef=coef,
        unit=unit,
        robin_only=robin_only,
        robin_total=robin_total,
        robin_rhs=robin_rhs,
        robin_weights=weights,
        robin_rhs_weights=rhs_weights,
        ambients=ambients,
        capacity=cap,
        free=free,
        fixed=fixed,
        active_nodes=act.reshape(-1),
        fixed_temperature=fixed_t,
        diag=diag,
        load=nodal_load(p),
        node_shape=p.node_shape,
        element_shape=p.element_shape,
    )


def diagonal(coef, unit, robin_total, free, node_shape):
    """Corner-major sum of the element diagonals plus Robin; fixed rows 1
    (operator.py:289-306; thermal_system.cpp:329-349)."""

    d = np.zeros(node_shape)
    views = corner_views(d)
    for k in range(8):
        views[k] += (coef[0] * unit[0, k, k] + coef[1] * unit[1, k, k]) + coef[2] * unit[2, k, k]
    flat = d.reshape(-1) + robin_total
    if np.any(flat[free] <= 0.0):
        raise ValueError("every free node must have positive conductance")
    return np.where(free, flat, 1.0)


def default_reference(ops, p: OracleProblem):
    """Ambient of the first convection boundary, else the mean fixed value
    (operator.py:389-397)."""

    if ops["ambients"]:
        return ops["ambients"][0]
    mask = (np.asarray(p.fixed_mask, bool).reshape(-1)) & ops["active_nodes"]
    return float(np.mean(np.asarray(p.fixed_values, float).reshape(-1)[mask]))


def build_rhs(ops, reference, stiffness, previous=None):
    """Right-hand side for the rise above ``reference`` (operator.py:411-440;
    thermal_system.cpp:411-436).  ``stiffness(x)`` is ``K x`` (no Robin,
    every node free)."""

    fixed = ops["fixed"]
    shifted = np.where(fixed & ops["active_nodes"], ops["fixed_temperature"] - reference, 0.0)
    boundary = stiffness(shifted)
    stored = 0.0 if previous is None else ops["capacity"] * np.where(np.isfinite(previous), previous, 0.0)
    robin = ops["robin_rhs"] - ops["robin_total"] * reference
    rhs = ops["load"] + robin + stored - boundary
    rhs[fixed] = shifted[fixed]
    return rhs


def initial_rise(ops, guess, reference):
    """The warm start of ``_solve_linear`` (solve.py:206-222)."""

    g = np.asarray(guess, float)
    x = np.full(ops["free"].size, float(g)) if g.ndim == 0 else g.reshape(-1).copy()
    x = np.where(ops["fixed"], ops["fixed_temperature"] + reference * ~ops["active_nodes"], x)
    return x - reference


def heat_budget(ops, p: OracleProblem, temperature, stiffness):
    """Fixed-node heat, convective heat per boundary, total input and the
    balance error (solve.py:255-280; thermal_system.cpp:444-511)."""

    t = np.asarray(temperature, float).reshape(-1)
    residual = stiffness(t) + ops["robin_only"] * t - ops["robin_rhs"] - ops["load"]
    fixed_heat = -float(np.sum(residual[ops["fixed"]]))
    conv = np.array([float(np.sum(w * t) - np.sum(rw)) for w, rw in zip(ops["robin_weights"], ops["robin_rhs_weights"])])
    total = float(np.sum(ops["load"]))
    return dict(
        fixed=fixed_heat,
        convective=conv,
        total=total,
        balance=total - float(np.sum(conv)) - fixed_heat,
        residual=residual,
    )


def element_heat_flux(p: OracleProblem, temperature):
    """``-k grad T`` at element centres from the corner averages (operator.py:500-525)."""

    g = np.asarray(temperature, float).reshape(p.node_shape)
    c = corner_views(g)
    xp_, xm = c[1] + c[3] + c[5] + c[7], c[0] + c[2] + c[4] + c[6]
    yp, ym = c[2] + c[3] + c[6] + c[7], c[0] + c[1] + c[4] + c[5]
    zp, zm = c[4] + c[5] + c[6] + c[7], c[0] + c[1] + c[2] + c[3]
    gx = (xp_ - xm) / (4.0 * np.asarray(p.pitch_x)[None, None, :])
    gy = (yp - ym) / (4.0 * np.asarray(p.pitch_y)[None, :, None])
    gz = (zp - zm) / (4.0 * np.asarray(p.thickness)[:, None, None])
    k_in = np.where(p.active, p.k_in, 0.0)
    k_z = np.where(p.active, p.k_z, 0.0)
    return np.stack((-k_in * gx, -k_in * gy, -k_z * gz), axis=-1)


def from_problem(problem) -> OracleProblem:
    """Plain arrays of a reference-API ``ThermalConductionProblem`` (read by
    attribute, so the oracle imports nothing from the package it checks)."""

    mesh = problem.mesh
    conv = []
    for b in problem.convection:
        if hasattr(b, "side"):
            conv.append(Convection(b.side, b.coefficient_w_per_m2_k, b.ambient_temperature_k))
        else:
            conv.append(Convection("exposed", b.coefficient_w_per_m2_k, b.ambient_temperature_k, tuple(b.directions)))
    shape = mesh.node_shape
    sources = [([int(np.ravel_multi_index(n, shape)) for n in s.nodes], s.power_w) for s in problem.heat_sources]
    return OracleProblem(
        thickness=np.asarray(mesh.slab_thickness_m),
        pitch_x=np.asarray(mesh.pitch_x_m),
        pitch_y=np.asarray(mesh.pitch_y_m),
        k_in=np.asarray(mesh.conductivity_w_per_m_k),
        k_z=np.asarray(mesh.through_plane_conductivity_w_per_m_k),
        active=np.asarray(mesh.active),
        fixed_mask=np.asarray(problem.fixed_temperature_mask),
        fixed_values=np.asarray(problem.fixed_temperature_k),
        convection=conv,
        element_heat=problem.element_heat_w,
        nodal_heat=problem.nodal_heat_w,
        sources=sources,
    )
