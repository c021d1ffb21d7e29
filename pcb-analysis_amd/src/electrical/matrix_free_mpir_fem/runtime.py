"""The FP32 inner-solve runtime of the HIP path.

The reference interface is the ``LowPrecisionRuntime`` contract of
src/electrical/matrix_free_mpir_fem/runtime.py:20-44 with the device runtime
of :121-170: ``from_host``, ``to_host``, ``zeros_like``, ``copy``, ``dot``,
``norm``, ``axpy``, ``divide`` and ``synchronize``, on arrays that live on the
device.  Here the arrays are ``pcbhip._capi.DeviceVector`` (FP32 on one HIP
device) and every operation is a kernel of ``libpcbhip.so`` that returns when
it is done, so a caller that drives a PCG loop from the host -- the reference's
``_pcg_low`` shape, with an operator's ``apply_low`` / ``precondition_low`` on
the same vectors -- runs it on the device.  ``project`` and ``combine`` add
the basis products of GMRES's Gram-Schmidt passes (the GEMV calls of the
reference's device GMRES, solver.py:400-577).  The product's own solves do
not go through here: they run the whole iteration in fused device kernels.

Differences from the reference's device runtime, both deliberate: ``dot`` and
``norm`` accumulate in FP64 in an order fixed by the length (the same bits
on every call) where CuPy accumulates in FP32; and there is no array
library behind the vectors, so ``namespace`` is None.  ``backend`` "hip",
"cuda" (the reference's name for its GPU path) and "auto" select the HIP
device; "cpu" is refused: this package has no host fallback.
"""

from __future__ import annotations

import ctypes
import math
from typing import Literal

import numpy as np

RuntimeBackend = Literal["hip", "cuda", "auto", "cpu"]


class HipFloat32Runtime:
    """FP32 vectors and their operations on one HIP device."""

    name = "hip-fp32"
    namespace = None
    dtype = np.float32
    is_cuda = True  # the reference's flag for "device runtime"
    is_hip = True

    def __init__(self, *, device_id: int = 0) -> None:
        from pcbhip import _capi

        _capi.require_device(int(device_id))
        self.device_id = int(device_id)
        self._capi = _capi
        self._lib = _capi.load()

    # -- arrays ----------------------------------------------------------------

    def _device(self, value):
        """``value`` as a device vector of this device (host arrays are uploaded)."""

        if isinstance(value, self._capi.DeviceVector):
            if value.device != self.device_id:
                raise ValueError(f"vector lives on device {value.device}, the runtime on {self.device_id}")
            return value
        return self._capi.DeviceVector.from_host(self.device_id, value)

    @staticmethod
    def _same_size(a, b) -> int:
        if a.size != b.size:
            raise ValueError(f"vectors of sizes {a.size} and {b.size}")
        return a.size

    def from_host(self, value: np.ndarray):
        return self._capi.DeviceVector.from_host(self.device_id, value)

    def to_host(self, value) -> np.ndarray:
        if isinstance(value, self._capi.DeviceVector):
            return value.to_host().astype(np.float64)
        return np.asarray(value, dtype=np.float64)

    def zeros_like(self, value):
        size = value.size if isinstance(value, self._capi.DeviceVector) else int(np.size(value))
        return self._capi.DeviceVector(self.device_id, size)

    def copy(self, value):
        src = self._device(value)
        out = self._capi.DeviceVector(self.device_id, src.size)
        self._capi.check(self._lib.pcbhip_vec_copy(self.device_id, src.ptr, out.ptr, src.size))
        return out

    # -- reductions --------------------------------------------------------------

    def dot(self, left, right) -> float:
        a, b = self._device(left), self._device(right)
        n = self._same_size(a, b)
        out = ctypes.c_double()
        self._capi.check(self._lib.pcbhip_vec_dot(self.device_id, a.ptr, b.ptr, n, ctypes.byref(out)))
        return float(out.value)

    def norm(self, value) -> float:
        v = self._device(value)
        return math.sqrt(max(self.dot(v, v), 0.0))

    # -- element-wise --------------------------------------------------------------

    def axpy(self, alpha: float, x, y):
        """A new vector ``y + alpha x`` in FP32."""

        a, b = self._device(x), self._device(y)
        n = self._same_size(a, b)
        out = self._capi.DeviceVector(self.device_id, n)
        self._capi.check(self._lib.pcbhip_vec_axpy(self.device_id, float(np.float32(alpha)), a.ptr, b.ptr, out.ptr, n))
        return out

    def divide(self, numerator, denominator):
        """A new vector ``numerator / denominator`` in FP32 (a host
        denominator, such as an operator's FP32 diagonal, is uploaded)."""

        a, b = self._device(numerator), self._device(denominator)
        n = self._same_size(a, b)
        out = self._capi.DeviceVector(self.device_id, n)
        self._capi.check(self._lib.pcbhip_vec_divide(self.device_id, a.ptr, b.ptr, out.ptr, n))
        return out

    # -- Gram-Schmidt products of a GMRES basis ------------------------------------

    # Rows one library call takes; longer bases are taken in chunks.
    basis_chunk = 64

    def project(self, rows, vector) -> np.ndarray:
        """``[row . vector for row in rows]`` as FP64 values, each with the
        bits :meth:`dot` gives for the same pair, in one pass over
        ``vector`` per chunk of rows."""

        w = self._device(vector)
        vectors = [self._device(row) for row in rows]
        for row in vectors:
            self._same_size(row, w)
        out = np.zeros(len(vectors), dtype=np.float64)
        for lo in range(0, len(vectors), self.basis_chunk):
            chunk = vectors[lo : lo + self.basis_chunk]
            table = (ctypes.c_void_p * len(chunk))(*[row.ptr for row in chunk])
            part = np.zeros(len(chunk), dtype=np.float64)
            self._capi.check(self._lib.pcbhip_vec_project(
                self.device_id, table, len(chunk), w.ptr, w.size,
                part.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            out[lo : lo + len(chunk)] = part
        return out

    def combine(self, coefficients, rows, base=None):
        """A new vector ``base + sum_k coefficients[k] rows[k]`` (``base``
        None: zero), accumulated in FP64 per chunk of rows and rounded to
        FP32 once per chunk."""

        coefficients = np.ascontiguousarray(coefficients, dtype=np.float64).reshape(-1)
        vectors = [self._device(row) for row in rows]
        if coefficients.size != len(vectors) or not vectors:
            raise ValueError("combine needs one coefficient per row and at least one row")
        n = vectors[0].size
        for row in vectors:
            if row.size != n:
                raise ValueError(f"vectors of sizes {row.size} and {n}")
        out = self._capi.DeviceVector(self.device_id, n)
        previous = None if base is None else self._device(base)
        if previous is not None:
            self._same_size(previous, out)
        for lo in range(0, len(vectors), self.basis_chunk):
            chunk = vectors[lo : lo + self.basis_chunk]
            table = (ctypes.c_void_p * len(chunk))(*[row.ptr for row in chunk])
            coef = np.ascontiguousarray(coefficients[lo : lo + len(chunk)])
            self._capi.check(self._lib.pcbhip_vec_combine(
                self.device_id, table, coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(chunk),
                None if previous is None else previous.ptr, out.ptr, n))
            previous = out
        return out

    def synchronize(self) -> None:
        # Every operation has finished when it returns.
        return None


def make_float32_runtime(backend: RuntimeBackend = "hip", *, device_id: int = 0) -> HipFloat32Runtime:
    if backend in ("hip", "cuda", "auto"):
        return HipFloat32Runtime(device_id=device_id)
    if backend == "cpu":
        raise ValueError("this package runs the solve on the HIP device only; use backend='hip'")
    raise ValueError("backend must be 'hip', 'cuda', or 'auto'")
