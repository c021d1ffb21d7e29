"""The allocation-free preconditioner entry point against the allocating one.

``pcbhip_operator_precondition_f32_noalloc`` runs the same kernels as
``pcbhip_operator_precondition_f32`` with its coarse vectors in a static
device array, so its output must equal the existing entry point's bit for
bit.  Two two-level operators are preconditioned in a row, the second
smaller (fewer nodes and coarse unknowns) than the first, so the second call
runs over scratch the first left dirty; a Jacobi-only operator follows."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from pcbhip import _capi
from tests.helpers import copper_stack
from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _capi.require_device(0)


def _hip():
    """The HIP runtime the library itself loaded (same soname), for the
    test's own device buffers."""

    _capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return hip


def _noalloc(dev, r):
    hip = _hip()
    lib = _capi.load()
    nbytes = r.nbytes
    r_dev, z_dev = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(r_dev), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(z_dev), nbytes) == 0
    try:
        assert hip.hipMemcpy(r_dev, r.ctypes.data, nbytes, 1) == 0  # host to device
        assert hip.hipMemset(z_dev, 0xFF, nbytes) == 0  # NaN bits: every output must be written
        assert hip.hipDeviceSynchronize() == 0
        # NULL stream: the operator's own stream.
        _capi.check(lib.pcbhip_operator_precondition_f32_noalloc(dev.handle, r_dev, z_dev, None))
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty_like(r)
        assert hip.hipMemcpy(out.ctypes.data, z_dev, nbytes, 2) == 0  # device to host
        return out
    finally:
        hip.hipFree(r_dev)
        hip.hipFree(z_dev)


def test_noalloc_preconditioner_equals_the_allocating_entry_point():
    cases = [("two-level", (40, 90), 4), ("two-level", (12, 20), 2), ("jacobi", (16, 30), None)]
    for k, (pre, shape, block) in enumerate(cases):
        print(f"noalloc precondition case {k}: {pre} {shape} block={block}", flush=True)
        op = MatrixFreeThermalOperator(copper_stack(*shape), preconditioner=pre, coarse_block_nodes=block)
        dev = op.device_operator
        r = np.random.default_rng(k).standard_normal(op.size).astype(np.float32)
        want = dev.precondition_f32(r)
        got = _noalloc(dev, r)
        assert got.tobytes() == want.tobytes(), (pre, shape)
        assert np.all(np.isfinite(got))


def test_noalloc_preconditioner_rejects_null_arguments():
    lib = _capi.load()
    code = lib.pcbhip_operator_precondition_f32_noalloc(None, None, None, None)
    assert code == _capi.INVALID_INPUT
    with pytest.raises(ValueError, match="null argument"):
        _capi.check(code)
