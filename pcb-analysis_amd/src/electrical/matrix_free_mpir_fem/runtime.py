"""The low-precision runtime of the HIP path (reference contract:
src/electrical/matrix_free_mpir_fem/runtime.py:20-44, 277-294).

On MI355X the whole inner solve runs in fused device kernels, so the runtime
is a descriptor of the device the operator lives on rather than a vector
library.  ``backend`` "hip", "cuda" (the reference's name for its GPU path)
and "auto" select the HIP device; "cpu" is refused: this package has no host
fallback.
"""

from __future__ import annotations

from typing import Literal

import numpy as np

RuntimeBackend = Literal["hip", "cuda", "auto", "cpu"]


class HipFloat32Runtime:
    """FP32 inner solves on one HIP device."""

    name = "hip-fp32"
    dtype = np.float32
    is_cuda = True  # the reference's flag for "device runtime"
    is_hip = True

    def __init__(self, *, device_id: int = 0) -> None:
        from pcbhip import _capi

        _capi.require_device(int(device_id))
        self.device_id = int(device_id)

    def from_host(self, value: np.ndarray) -> np.ndarray:
        return np.asarray(value, dtype=self.dtype)

    def to_host(self, value) -> np.ndarray:
        return np.asarray(value, dtype=np.float64)

    def synchronize(self) -> None:
        return None


def make_float32_runtime(backend: RuntimeBackend = "hip", *, device_id: int = 0) -> HipFloat32Runtime:
    if backend in ("hip", "cuda", "auto"):
        return HipFloat32Runtime(device_id=device_id)
    if backend == "cpu":
        raise ValueError("this package runs the solve on the HIP device only; use backend='hip'")
    raise ValueError("backend must be 'hip', 'cuda', or 'auto'")
