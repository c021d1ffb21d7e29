"""Device operator parity: the gfx950 sum-factorised Q1 action, its FP64
twin, the unconstrained stiffness and the Jacobi diagonal against the
oracle's element-assembled FP64 matrix (tests/test_native_thermal.py:55-96
is the reference's version of these checks: FP32 within a few eps32 of the
FP64 action, FP64 within 64 eps64, fixed rows exactly the identity)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import assembled as A
from oracle import problem as P
from pcbhip import _capi
from tests.helpers import DOUBLE_EPS, SINGLE_EPS, random_body, stack, to_oracle

pytestmark = pytest.mark.gpu

# (slabs, rows, cols): single elements, strips narrower and wider than one
# wave strip (62 output columns), several workgroups in y.
SHAPES = [(1, 1, 1), (2, 1, 4), (3, 2, 3), (1, 1, 61), (1, 2, 62), (2, 3, 63), (3, 12, 20), (2, 40, 130),
          (5, 70, 63), (4, 33, 125)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _device_and_matrix(problem):
    ops = P.operator_arrays(to_oracle(problem))
    dev = _capi.Operator(0, problem.mesh.element_grid_shape, ops["coef"], ops["robin_total"], ops["free"])
    a = A.system_matrix(ops["coef"], ops["unit"], ops["robin_total"], ops["free"], problem.mesh.element_grid_shape)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], problem.mesh.element_grid_shape)
    return dev, ops, a, k


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("void", [False, True])
def test_device_actions_match_the_assembled_operator(shape, void):
    print(f"apply parity shape={shape} void={void}", flush=True)
    problem = random_body(shape, seed=sum(shape), void=void)
    dev, ops, a, k = _device_and_matrix(problem)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(problem.mesh.size)
    scale = float((abs(a) @ np.abs(x)).max())

    y64 = dev.apply_f64(x)
    np.testing.assert_allclose(y64, a @ x, rtol=0, atol=64 * DOUBLE_EPS * scale)

    x32 = x.astype(np.float32)
    y32 = dev.apply_f32(x32)
    assert y32.dtype == np.float32
    exact32 = a @ x32.astype(np.float64)
    assert np.max(np.abs(y32 - exact32)) <= 32 * SINGLE_EPS * scale
    fixed = ops["fixed"]
    np.testing.assert_array_equal(y32[fixed], x32[fixed])
    np.testing.assert_array_equal(y64[fixed], x[fixed])

    ks = dev.stiffness_f64(x)
    np.testing.assert_allclose(ks, k @ x, rtol=0, atol=64 * DOUBLE_EPS * float((abs(k) @ np.abs(x)).max()))

    np.testing.assert_allclose(dev.diagonal(), a.diagonal(), rtol=4 * DOUBLE_EPS, atol=0)
    # Bitwise repeatable: every node is written by one lane in a fixed order.
    assert dev.apply_f32(x32).tobytes() == y32.tobytes()
    assert dev.apply_f64(x).tobytes() == y64.tobytes()


@pytest.mark.parametrize("shape", [(1, 1), (1, 4), (3, 2), (12, 20)])
@pytest.mark.parametrize("fixed", [False, True])
def test_reference_stack_fixture(shape, fixed):
    print(f"stack fixture shape={shape} fixed={fixed}", flush=True)
    problem = stack(*shape, fixed=fixed)
    dev, ops, a, _ = _device_and_matrix(problem)
    x = np.random.default_rng(3).standard_normal(problem.mesh.size).astype(np.float32)
    y = dev.apply_f32(x)
    expected = a @ x.astype(np.float64)
    assert np.linalg.norm(y - expected) <= 8.0 * SINGLE_EPS * np.linalg.norm(expected) + 8.0 * SINGLE_EPS * float(
        (abs(a) @ np.abs(x.astype(np.float64))).max())


def test_pure_conduction_annihilates_constants_on_the_device():
    print("constant kernel", flush=True)
    problem = random_body((3, 9, 70), seed=3)
    ops = P.operator_arrays(to_oracle(problem))
    dev = _capi.Operator(0, problem.mesh.element_grid_shape, ops["coef"], ops["robin_total"], ops["free"])
    ones = np.ones(problem.mesh.size)
    k = A.stiffness_matrix(ops["coef"], ops["unit"], problem.mesh.element_grid_shape)
    # The sum-factorised action works on edge differences, so a constant
    # field gives exactly zero.
    assert np.all(dev.stiffness_f64(ones) == 0.0)
    assert np.abs(k @ ones).max() < 1e-10 * abs(k).max()


def test_a_free_node_without_conductance_is_refused():
    print("invalid operator", flush=True)
    coef = np.zeros(3)
    robin = np.zeros(8)
    free = np.ones(8, dtype=np.uint8)
    with pytest.raises(ValueError, match="positive conductance"):
        _capi.Operator(0, (1, 1, 1), coef, robin, free)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("void", [False, True])
@pytest.mark.parametrize("terms", [2, 3])
def test_matrix_core_action_matches_the_assembled_operator(shape, void, terms):
    print(f"mfma apply parity shape={shape} void={void} terms={terms}", flush=True)
    problem = random_body(shape, seed=sum(shape) + 1, void=void)
    dev, ops, a, _ = _device_and_matrix(problem)
    dev.set_engine(terms)
    x32 = np.random.default_rng(8).standard_normal(problem.mesh.size).astype(np.float32)
    exact = a @ x32.astype(np.float64)
    scale = float((abs(a) @ np.abs(x32.astype(np.float64))).max())
    y = dev.apply_f32(x32)
    # Three bf16 terms hold every FP32 value exactly and the products are
    # exact, so that action has the FP32 action's error bound.  Two terms
    # (hi + lo, rounded to nearest) keep 16 significant bits: every corner
    # value is off by at most 2^-16 of itself.
    bound = 32 * SINGLE_EPS * scale if terms == 3 else 2.0**-15 * scale
    assert np.max(np.abs(y - exact)) <= bound
    np.testing.assert_array_equal(y[ops["fixed"]], x32[ops["fixed"]])
    assert dev.apply_f32(x32).tobytes() == y.tobytes()
    # Every element's product is one matrix-core column: the bits do not
    # depend on the tile shape.
    for rows, zb in ((8, 2), (4, 2), (8, 1 << 20)):
        dev.set_shape(rows, zb)
        assert dev.apply_f32(x32).tobytes() == y.tobytes(), (rows, zb)


@pytest.mark.parametrize("shape", [(9, 21, 70), (4, 5, 70), (2, 12, 20), (1, 3, 7)],
                         ids=lambda s: "x".join(map(str, s)))
def test_vector_action_bits_do_not_depend_on_the_tile_shape(shape):
    print(f"vector apply tile-shape independence shape={shape}", flush=True)
    # Neighbouring z blocks sweep in opposite directions, and every node adds
    # the partial of the element layer below it to that of the layer above in
    # a fixed order.  So z blocks of 2 planes (both directions; with 4 slabs
    # the last block is the top plane alone, swept down), of 4 and 6, and one
    # block going up give the same bits, in FP32 and FP64, and the answer is
    # the assembled operator's.
    problem = random_body(shape, seed=31 + sum(shape), void=True)
    dev, ops, a, _ = _device_and_matrix(problem)
    x = np.random.default_rng(9).standard_normal(problem.mesh.size)
    x32 = x.astype(np.float32)
    dev.set_shape(4, 1 << 20, 4, 1 << 20)
    y32 = dev.apply_f32(x32)
    y64 = dev.apply_f64(x)
    scale = float((abs(a) @ np.abs(x)).max())
    np.testing.assert_allclose(y64, a @ x, rtol=0, atol=64 * DOUBLE_EPS * scale)
    assert np.max(np.abs(y32 - a @ x32.astype(np.float64))) <= 32 * SINGLE_EPS * scale
    for rows32, zb, rows64 in ((4, 2, 2), (8, 2, 4), (4, 4, 2), (8, 6, 4), (8, 1 << 20, 2)):
        dev.set_shape(rows32, zb, rows64, zb)
        assert dev.apply_f32(x32).tobytes() == y32.tobytes(), (rows32, zb)
        assert dev.apply_f64(x).tobytes() == y64.tobytes(), (rows64, zb)


def test_plain_planes_and_uniform_layers_match_the_assembled_operator():
    print("apply on a board with plain planes and uniform layers", flush=True)
    # The 64 x 64 x 8 board is fixed on its bottom plane and cooled on its top
    # one; its 7 inner planes are free with no Robin term, and there the
    # stencils take robin = +0 without loading it.  Copper is on its outer
    # slabs only, so the 6 FR-4 slabs between are uniform element layers,
    # whose coefficients come from a per-layer table.
    from pcbhip import workloads

    problem = workloads.board_64x64x8()
    dev, ops, a, k = _device_and_matrix(problem)
    s, r, c = problem.mesh.element_grid_shape
    coef = np.asarray(ops["coef"]).reshape(3, s, r, c)
    uniform = [ez for ez in range(s) if all(np.all(coef[d, ez] == coef[d, ez].flat[0]) for d in range(3))]
    assert uniform == list(range(1, s - 1))
    robin = ops["robin_total"].reshape(s + 1, r + 1, c + 1)
    free = ops["free"].reshape(s + 1, r + 1, c + 1)
    plain = [z for z in range(s + 1) if np.all(robin[z] == 0) and np.all(free[z] == 1)]
    assert plain == list(range(1, s))
    x = np.random.default_rng(10).standard_normal(problem.mesh.size)
    x32 = x.astype(np.float32)
    scale = float((abs(a) @ np.abs(x)).max())
    y64 = dev.apply_f64(x)
    np.testing.assert_allclose(y64, a @ x, rtol=0, atol=64 * DOUBLE_EPS * scale)
    y32 = dev.apply_f32(x32)
    assert np.max(np.abs(y32 - a @ x32.astype(np.float64))) <= 32 * SINGLE_EPS * scale
    np.testing.assert_array_equal(y32[ops["fixed"]], x32[ops["fixed"]])
    ks = dev.stiffness_f64(x)
    np.testing.assert_allclose(ks, k @ x, rtol=0, atol=64 * DOUBLE_EPS * float((abs(k) @ np.abs(x)).max()))
