"""ctypes binding of ``libpcbhip.so``, the C ABI declared in ``include/pcbhip.h``.

This is the thin layer between the Python solvers and the gfx950 kernels
(the counterpart of the reference's pybind modules ``cpp/bindings/thermal.cpp``
and ``cpp/bindings/cuda_module.cpp``).  There is no CPU fallback: when the
library is missing, or no HIP device is visible, every call raises.

Status codes map onto the Python exceptions of the reference's translator
(``cpp/include/pcbcore/errors.hpp:1-31``, ``cpp/bindings/module.cpp:571-583``):
invalid input -> ValueError, runtime failure -> RuntimeError, non-finite ->
FloatingPointError, allocation failure -> MemoryError.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpcbhip.so"

OK, INVALID_INPUT, RUNTIME, NON_FINITE, NO_MEMORY, FULL = 0, 1, 2, 3, 4, 5


class ServerFull(RuntimeError):
    """Every slot of a server is in use (the caller keeps the request queued)."""


_ERRORS = {INVALID_INPUT: ValueError, RUNTIME: RuntimeError, NON_FINITE: FloatingPointError, NO_MEMORY: MemoryError,
           FULL: ServerFull}

# Every symbol include/pcbhip.h declares.
EXPORTS = (
    "pcbhip_last_error",
    "pcbhip_device_count",
    "pcbhip_operator_create",
    "pcbhip_sheet_operator_create",
    "pcbhip_operator_destroy",
    "pcbhip_operator_size",
    "pcbhip_operator_diagonal",
    "pcbhip_operator_set_coarse",
    "pcbhip_operator_set_coarse_matrix",
    "pcbhip_operator_precondition_f32",
    "pcbhip_operator_precondition_f32_noalloc",
    "pcbhip_operator_galerkin_f64",
    "pcbhip_apply_f32",
    "pcbhip_apply_f64",
    "pcbhip_stiffness_f64",
    "pcbhip_apply_f32_dev",
    "pcbhip_apply_f64_dev",
    "pcbhip_stiffness_f64_dev",
    "pcbhip_solve_mpir",
    "pcbhip_batch_create",
    "pcbhip_batch_destroy",
    "pcbhip_batch_set",
    "pcbhip_batch_set_shape",
    "pcbhip_batch_set_schedule",
    "pcbhip_operator_set_shape",
    "pcbhip_operator_set_engine",
    "pcbhip_batch_set_block",
    "pcbhip_batch_rhs_dev",
    "pcbhip_batch_run",
    "pcbhip_batch_result",
    "pcbhip_batch_time_pass",
    "pcbhip_operator_bytes",
    "pcbhip_server_create",
    "pcbhip_server_destroy",
    "pcbhip_server_set_shape",
    "pcbhip_server_slot_bytes",
    "pcbhip_server_submit",
    "pcbhip_server_step",
    "pcbhip_server_finished",
    "pcbhip_server_collect",
    "pcbhip_server_stats",
    "pcbhip_server_trim",
    "pcbhip_copy_dev",
    "pcbhip_streams_dev",
    "pcbhip_vec_create",
    "pcbhip_vec_destroy",
    "pcbhip_vec_write",
    "pcbhip_vec_read",
    "pcbhip_vec_copy",
    "pcbhip_vec_dot",
    "pcbhip_vec_axpy",
    "pcbhip_vec_divide",
    "pcbhip_vec_project",
    "pcbhip_vec_combine",
    "pcbhip_operator_precondition_f32_dev",
    "pcbhip_operator_synchronize",
)


class MpirConfigC(ctypes.Structure):
    _fields_ = [
        ("relative_tolerance", ctypes.c_double),
        ("absolute_tolerance", ctypes.c_double),
        ("inner_relative_tolerance", ctypes.c_double),
        ("max_outer_iterations", ctypes.c_int),
        ("max_inner_iterations", ctypes.c_int),
    ]


class ServerInfoC(ctypes.Structure):
    _fields_ = [
        ("queued", ctypes.c_int),
        ("running", ctypes.c_int),
        ("finished", ctypes.c_int),
        ("free_slots", ctypes.c_int),
        ("bytes_in_use", ctypes.c_int64),
        ("bytes_pooled", ctypes.c_int64),
        ("replays", ctypes.c_int64),
        ("graph_builds", ctypes.c_int64),
        ("steps_per_replay", ctypes.c_int),
        ("wait_ms", ctypes.c_double),
        ("replay_ms", ctypes.c_double),
    ]


class MpirResultC(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("inner_iterations", ctypes.c_int),
        ("high_operator_applications", ctypes.c_int),
        ("low_operator_applications", ctypes.c_int),
        ("not_spd", ctypes.c_int),
        ("relative_residual", ctypes.c_double),
    ]


_lib = None
_vp = ctypes.c_void_p
_int = ctypes.c_int


def load() -> ctypes.CDLL:
    """The library, loaded once; raises ImportError when it is not built."""

    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"the HIP library {LIB_PATH} is not built; run python -c 'import __graft_entry__ as g; g.build()'"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.pcbhip_last_error.restype = ctypes.c_char_p
    lib.pcbhip_device_count.argtypes = [ctypes.POINTER(_int)]
    lib.pcbhip_operator_create.argtypes = [_int, _int, _int, _int, _vp, _vp, _vp, ctypes.POINTER(_vp)]
    lib.pcbhip_sheet_operator_create.argtypes = [_int, _int, _int, _int, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                                 ctypes.POINTER(_vp)]
    lib.pcbhip_operator_destroy.argtypes = [_vp]
    lib.pcbhip_operator_size.argtypes = [_vp]
    lib.pcbhip_operator_size.restype = ctypes.c_int64
    lib.pcbhip_operator_diagonal.argtypes = [_vp, _vp]
    lib.pcbhip_operator_set_coarse.argtypes = [_vp, _int, _int, _vp]
    lib.pcbhip_operator_set_coarse_matrix.argtypes = [_vp, _int, _int, _vp, _vp]
    lib.pcbhip_operator_precondition_f32.argtypes = [_vp, _vp, _vp]
    lib.pcbhip_operator_precondition_f32_noalloc.argtypes = [_vp, _vp, _vp, _vp]
    lib.pcbhip_operator_galerkin_f64.argtypes = [_vp, _int, _int, _vp, _vp]
    for name in ("pcbhip_apply_f32", "pcbhip_apply_f64", "pcbhip_stiffness_f64"):
        getattr(lib, name).argtypes = [_vp, _vp, _vp]
    for name in ("pcbhip_apply_f32_dev", "pcbhip_apply_f64_dev", "pcbhip_stiffness_f64_dev"):
        getattr(lib, name).argtypes = [_vp, _vp, _vp, _vp]
    lib.pcbhip_solve_mpir.argtypes = [_vp, _vp, _vp, ctypes.POINTER(MpirConfigC), _int, ctypes.POINTER(MpirResultC), _vp]
    lib.pcbhip_batch_create.argtypes = [_int, ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp)]
    lib.pcbhip_batch_destroy.argtypes = [_vp]
    lib.pcbhip_batch_set.argtypes = [_vp, _int, _vp, _vp, ctypes.POINTER(MpirConfigC)]
    lib.pcbhip_batch_rhs_dev.argtypes = [_vp, _int, ctypes.POINTER(_vp)]
    lib.pcbhip_batch_set_shape.argtypes = [_vp, _int, _int]
    lib.pcbhip_batch_set_schedule.argtypes = [_vp, _int, _int, _int]
    lib.pcbhip_operator_set_shape.argtypes = [_vp, _int, _int, _int, _int]
    lib.pcbhip_operator_set_engine.argtypes = [_vp, _int]
    lib.pcbhip_batch_set_block.argtypes = [_vp, _int, _int]
    lib.pcbhip_batch_run.argtypes = [_vp]
    lib.pcbhip_batch_result.argtypes = [_vp, _int, _vp, ctypes.POINTER(MpirResultC), _vp]
    lib.pcbhip_batch_time_pass.argtypes = [_vp, _int, ctypes.POINTER(ctypes.c_double)]
    lib.pcbhip_copy_dev.argtypes = [_vp, _vp, ctypes.c_int64, _vp]
    lib.pcbhip_streams_dev.argtypes = [ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp), _int, ctypes.c_int64, _vp]
    _i64p = ctypes.POINTER(ctypes.c_int64)
    lib.pcbhip_operator_bytes.argtypes = [_vp, _i64p]
    lib.pcbhip_server_create.argtypes = [_int, _int, _int, _int, ctypes.POINTER(_vp)]
    lib.pcbhip_server_destroy.argtypes = [_vp]
    lib.pcbhip_server_set_shape.argtypes = [_vp, _int, _int]
    lib.pcbhip_server_slot_bytes.argtypes = [_vp, _vp, _int, _i64p]
    lib.pcbhip_server_submit.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(MpirConfigC), ctypes.c_int64,
                                         ctypes.POINTER(_int)]
    lib.pcbhip_server_step.argtypes = [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]
    lib.pcbhip_server_finished.argtypes = [_vp, _vp, _vp, _int, ctypes.POINTER(_int)]
    lib.pcbhip_server_collect.argtypes = [_vp, _int, _vp, ctypes.POINTER(MpirResultC), _vp]
    lib.pcbhip_server_stats.argtypes = [_vp, ctypes.POINTER(ServerInfoC)]
    lib.pcbhip_server_trim.argtypes = [_vp]
    _i64 = ctypes.c_int64
    lib.pcbhip_vec_create.argtypes = [_int, _i64, ctypes.POINTER(_vp)]
    lib.pcbhip_vec_destroy.argtypes = [_int, _vp]
    for name in ("pcbhip_vec_write", "pcbhip_vec_read", "pcbhip_vec_copy"):
        getattr(lib, name).argtypes = [_int, _vp, _vp, _i64]
    lib.pcbhip_vec_dot.argtypes = [_int, _vp, _vp, _i64, ctypes.POINTER(ctypes.c_double)]
    lib.pcbhip_vec_axpy.argtypes = [_int, ctypes.c_float, _vp, _vp, _vp, _i64]
    lib.pcbhip_vec_divide.argtypes = [_int, _vp, _vp, _vp, _i64]
    _pp = ctypes.POINTER(_vp)
    _dp = ctypes.POINTER(ctypes.c_double)
    lib.pcbhip_vec_project.argtypes = [_int, _pp, _int, _vp, _i64, _dp]
    lib.pcbhip_vec_combine.argtypes = [_int, _pp, _dp, _int, _vp, _vp, _i64]
    lib.pcbhip_operator_precondition_f32_dev.argtypes = [_vp, _vp, _vp]
    lib.pcbhip_operator_synchronize.argtypes = [_vp]
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != OK:
        message = load().pcbhip_last_error().decode() or f"pcbhip error {code}"
        raise _ERRORS.get(code, RuntimeError)(message)


def device_count() -> int:
    n = _int(0)
    check(load().pcbhip_device_count(ctypes.byref(n)))
    return int(n.value)


def require_device(device: int = 0) -> None:
    count = device_count()
    if count < 1:
        raise RuntimeError("HIP backend requested, but no HIP device is visible")
    if not 0 <= device < count:
        raise ValueError(f"HIP device_id {device} is outside the visible range [0, {count - 1}]")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def config_c(relative_tolerance, absolute_tolerance, inner_relative_tolerance, max_outer, max_inner) -> MpirConfigC:
    return MpirConfigC(float(relative_tolerance), float(absolute_tolerance), float(inner_relative_tolerance),
                       int(max_outer), int(max_inner))


class DeviceVector:
    """An FP32 vector on one HIP device, owned by this object (released by
    ``close`` or when collected): the array type of the FP32 inner-solve
    runtime.  ``np.asarray(v)`` copies it to the host."""

    dtype = np.dtype(np.float32)

    def __init__(self, device: int, size: int):
        lib = load()
        self._lib = lib
        self.device = int(device)
        self.size = int(size)
        self.ptr = None
        p = _vp()
        check(lib.pcbhip_vec_create(self.device, self.size, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, device: int, values) -> "DeviceVector":
        host = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        v = cls(device, host.size)
        check(v._lib.pcbhip_vec_write(v.device, v.ptr, _ptr(host), host.size))
        return v

    @property
    def shape(self) -> tuple[int]:
        return (self.size,)

    def __len__(self) -> int:
        return self.size

    def to_host(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.float32)
        check(self._lib.pcbhip_vec_read(self.device, self.ptr, _ptr(out), self.size))
        return out

    def __array__(self, dtype=None, copy=None):
        host = self.to_host()
        return host if dtype is None else host.astype(dtype)

    def close(self) -> None:
        if self.ptr is not None:
            ptr, self.ptr = self.ptr, None
            check(self._lib.pcbhip_vec_destroy(self.device, ptr))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Owning handle of one device-resident operator (``pcbhip_operator``)."""

    def __init__(self, device: int, element_shape, coef, robin, free):
        s, r, c = (int(v) for v in element_shape)
        n = (s + 1) * (r + 1) * (c + 1)
        coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        robin = np.ascontiguousarray(robin, dtype=np.float64).reshape(-1)
        free = np.ascontiguousarray(free, dtype=np.uint8).reshape(-1)
        if coef.size != 3 * s * r * c or robin.size != n or free.size != n:
            raise ValueError("operator arrays have the wrong size")
        lib = load()
        self._lib = lib
        self.handle = None
        self.device = int(device)
        self.element_shape = (s, r, c)
        self.node_layers = s + 1
        self.size = n
        self.coarse_size = 0
        h = _vp()
        check(lib.pcbhip_operator_create(self.device, s, r, c, _ptr(coef), _ptr(robin), _ptr(free), ctypes.byref(h)))
        self.handle = h

    @classmethod
    def sheet(cls, device: int, element_shape, coef, free, via_a=(), via_b=(), via_g=()) -> "Operator":
        """The layered-sheet DC operator (``pcbhip_sheet_operator_create``):
        ``element_shape`` = (layers, rows, cols), ``coef`` = (2, layers, rows,
        cols) sheet coefficients, vias as flat node index pairs."""

        layers, r, c = (int(v) for v in element_shape)
        n = layers * (r + 1) * (c + 1)
        coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        free = np.ascontiguousarray(free, dtype=np.uint8).reshape(-1)
        a = np.ascontiguousarray(via_a, dtype=np.int64).reshape(-1)
        b = np.ascontiguousarray(via_b, dtype=np.int64).reshape(-1)
        g = np.ascontiguousarray(via_g, dtype=np.float64).reshape(-1)
        if coef.size != 2 * layers * r * c or free.size != n or not (a.size == b.size == g.size):
            raise ValueError("sheet operator arrays have the wrong size")
        self = cls.__new__(cls)
        self._lib = load()
        self.handle = None
        self.device = int(device)
        self.element_shape = (layers, r, c)
        self.node_layers = layers
        self.size = n
        self.coarse_size = 0
        h = _vp()
        check(self._lib.pcbhip_sheet_operator_create(self.device, layers, r, c, _ptr(coef), _ptr(free), int(a.size),
                                                     _ptr(a) if a.size else None, _ptr(b) if b.size else None,
                                                     _ptr(g) if g.size else None, ctypes.byref(h)))
        self.handle = h
        return self

    def close(self) -> None:
        if self.handle is not None:
            self._lib.pcbhip_operator_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def diagonal(self) -> np.ndarray:
        out = np.empty(self.size)
        check(self._lib.pcbhip_operator_diagonal(self.handle, _ptr(out)))
        return out

    def _vec(self, x, dtype):
        v = np.ascontiguousarray(x, dtype=dtype).reshape(-1)
        if v.size != self.size:
            raise ValueError(f"vector has size {v.size}, expected {self.size}")
        return v

    def apply_f32(self, x) -> np.ndarray:
        v = self._vec(x, np.float32)
        out = np.empty(self.size, dtype=np.float32)
        check(self._lib.pcbhip_apply_f32(self.handle, _ptr(v), _ptr(out)))
        return out

    def apply_f64(self, x) -> np.ndarray:
        v = self._vec(x, np.float64)
        out = np.empty(self.size)
        check(self._lib.pcbhip_apply_f64(self.handle, _ptr(v), _ptr(out)))
        return out

    def stiffness_f64(self, x) -> np.ndarray:
        v = self._vec(x, np.float64)
        out = np.empty(self.size)
        check(self._lib.pcbhip_stiffness_f64(self.handle, _ptr(v), _ptr(out)))
        return out

    def set_coarse(self, block: int, inverse: np.ndarray | None) -> None:
        """Two-level preconditioning from now on: ``inverse`` is the dense
        inverse of the coarse Galerkin matrix of ``block``-wide patches
        (``None`` returns to Jacobi)."""

        if inverse is None:
            check(self._lib.pcbhip_operator_set_coarse(self.handle, 0, 0, None))
            self.coarse_size = 0
            return
        inv = np.ascontiguousarray(inverse, dtype=np.float64)
        if inv.ndim != 2 or inv.shape[0] != inv.shape[1]:
            raise ValueError("the coarse inverse must be a square matrix")
        check(self._lib.pcbhip_operator_set_coarse(self.handle, int(block), int(inv.shape[0]), _ptr(inv)))
        self.coarse_size = int(inv.shape[0])

    def galerkin_f64(self, block: int, vectors: np.ndarray) -> np.ndarray:
        """``Z^T A Z v`` for each row ``v`` of ``vectors`` (count x n_c), FP64 on
        the device, ``Z`` the free-node indicator of ``block``-wide patches."""

        v = np.ascontiguousarray(vectors, dtype=np.float64)
        _, rows, cols = self.element_shape
        nc = self.node_layers * -(-(rows + 1) // int(block)) * -(-(cols + 1) // int(block))
        if v.ndim != 2 or v.shape[1] != nc or v.shape[0] < 1:
            raise ValueError(f"vectors must be a (count, {nc}) array for block {block}")
        out = np.empty_like(v)
        check(self._lib.pcbhip_operator_galerkin_f64(self.handle, int(block), int(v.shape[0]), _ptr(v), _ptr(out)))
        return out

    def set_coarse_matrix(self, block: int, matrix: np.ndarray) -> np.ndarray:
        """Two-level preconditioning from now on, from the coarse Galerkin
        matrix of ``block``-wide patches: factored and inverted on the
        device.  Returns the FP64 inverse."""

        m = np.ascontiguousarray(matrix, dtype=np.float64)
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise ValueError("the coarse matrix must be square")
        inverse = np.empty_like(m)
        check(self._lib.pcbhip_operator_set_coarse_matrix(self.handle, int(block), int(m.shape[0]), _ptr(m),
                                                          _ptr(inverse)))
        self.coarse_size = int(m.shape[0])
        return inverse

    def precondition_f32(self, r) -> np.ndarray:
        """``M^-1 r`` in FP32 with the operator's preconditioner."""

        v = self._vec(r, np.float32)
        out = np.empty(self.size, dtype=np.float32)
        check(self._lib.pcbhip_operator_precondition_f32(self.handle, _ptr(v), _ptr(out)))
        return out

    def set_shape(self, rows32: int, zb32: int, rows64: int = 4, zb64: int = 1 << 20) -> None:
        check(self._lib.pcbhip_operator_set_shape(self.handle, rows32, zb32, rows64, zb64))

    def set_engine(self, engine: int) -> None:
        """FP32 element action: 0 vector units, 2 or 3 matrix cores with a
        two- or three-term bf16 split."""

        check(self._lib.pcbhip_operator_set_engine(self.handle, int(engine)))

    def _vector_of_size(self, v: "DeviceVector") -> None:
        if v.size != self.size or v.device != self.device:
            raise ValueError(f"device vector of size {v.size} on device {v.device}, expected size {self.size} "
                             f"on device {self.device}")

    def apply_f32_vec(self, x: "DeviceVector") -> "DeviceVector":
        """``A x`` in FP32 from one device vector into a new one."""

        self._vector_of_size(x)
        y = DeviceVector(self.device, self.size)
        check(self._lib.pcbhip_apply_f32_dev(self.handle, x.ptr, y.ptr, None))
        check(self._lib.pcbhip_operator_synchronize(self.handle))
        return y

    def precondition_f32_vec(self, r: "DeviceVector") -> "DeviceVector":
        """``M^-1 r`` in FP32 (Jacobi or two-level) from one device vector into a new one."""

        self._vector_of_size(r)
        z = DeviceVector(self.device, self.size)
        check(self._lib.pcbhip_operator_precondition_f32_dev(self.handle, r.ptr, z.ptr))
        return z

    def apply_f32_dev(self, x_ptr: int, y_ptr: int, stream: int | None = None) -> None:
        check(self._lib.pcbhip_apply_f32_dev(self.handle, x_ptr, y_ptr, stream))

    def apply_f64_dev(self, x_ptr: int, y_ptr: int, stream: int | None = None) -> None:
        check(self._lib.pcbhip_apply_f64_dev(self.handle, x_ptr, y_ptr, stream))

    def device_bytes(self) -> int:
        n = ctypes.c_int64(0)
        check(self._lib.pcbhip_operator_bytes(self.handle, ctypes.byref(n)))
        return int(n.value)


def result_dict(res: MpirResultC, history: np.ndarray) -> dict:
    rows = history.reshape(-1, 4)[: res.outer_iterations]
    return dict(
        converged=bool(res.converged),
        outer_iterations=int(res.outer_iterations),
        inner_iterations=int(res.inner_iterations),
        relative_residual=float(res.relative_residual),
        high_operator_applications=int(res.high_operator_applications),
        low_operator_applications=int(res.low_operator_applications),
        history=[(int(a), float(b), int(c), float(d)) for a, b, c, d in rows],
    )


class Batch:
    """Owning handle of a set of right-hand sides solved together (``pcbhip_batch``)."""

    def __init__(self, operators, inner_fp64: bool = False):
        ops = list(operators)
        lib = load()
        self._lib = lib
        self.handle = None
        self.operators = ops  # keeps the operator handles alive
        arr = (_vp * len(ops))(*[op.handle for op in ops])
        h = _vp()
        check(lib.pcbhip_batch_create(len(ops), arr, int(bool(inner_fp64)), ctypes.byref(h)))
        self.handle = h
        self.max_outer = [8] * len(ops)

    def close(self) -> None:
        if self.handle is not None:
            self._lib.pcbhip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __len__(self):
        return len(self.operators)

    def set(self, slot: int, rhs, x0=None, config: MpirConfigC | None = None) -> None:
        op = self.operators[slot]
        b = None if rhs is None else op._vec(rhs, np.float64)
        x = None if x0 is None else op._vec(x0, np.float64)
        check(self._lib.pcbhip_batch_set(self.handle, slot, None if b is None else _ptr(b),
                                         None if x is None else _ptr(x),
                                         None if config is None else ctypes.byref(config)))
        if config is not None:
            self.max_outer[slot] = int(config.max_outer_iterations)

    def set_shape(self, rows_inner: int, target_groups: int) -> None:
        check(self._lib.pcbhip_batch_set_shape(self.handle, int(rows_inner), int(target_groups)))

    def set_schedule(self, replay_steps: int = 64, outer_every: int = 32, compact_percent: int = 25) -> None:
        """Steps per captured replay, inner iterations between outer steps and
        the finished share (percent) that triggers compaction."""

        check(self._lib.pcbhip_batch_set_schedule(self.handle, int(replay_steps), int(outer_every),
                                                   int(compact_percent)))

    def set_block(self, group_size: int, engine: int = 3) -> None:
        """Apply slots that share an operator in groups of ``group_size`` with
        element action ``engine`` (0 vector, 2 or 3 matrix-core bf16 terms)."""

        check(self._lib.pcbhip_batch_set_block(self.handle, int(group_size), int(engine)))

    def rhs_device_pointer(self, slot: int) -> int:
        p = _vp()
        check(self._lib.pcbhip_batch_rhs_dev(self.handle, slot, ctypes.byref(p)))
        return int(p.value)

    def run(self) -> None:
        check(self._lib.pcbhip_batch_run(self.handle))

    def time_pass(self, launches: int = 100) -> float:
        """Average device microseconds of one fused PCG pass over every slot
        (``pcbhip_batch_time_pass``); call after ``run``."""

        us = ctypes.c_double()
        check(self._lib.pcbhip_batch_time_pass(self.handle, int(launches), ctypes.byref(us)))
        return us.value

    def result(self, slot: int, *, solution: bool = True):
        op = self.operators[slot]
        x = np.empty(op.size) if solution else None
        res = MpirResultC()
        history = np.zeros(4 * max(1, self.max_outer[slot]))
        check(self._lib.pcbhip_batch_result(self.handle, slot, None if x is None else _ptr(x), ctypes.byref(res),
                                            _ptr(history)))
        return x, result_dict(res, history)


def solve_mpir(op: Operator, rhs, x0, config: MpirConfigC, inner_fp64: bool = False):
    """One solve through ``pcbhip_solve_mpir``; returns (solution, result dict)."""

    b = op._vec(rhs, np.float64)
    x = np.zeros(op.size) if x0 is None else np.array(op._vec(x0, np.float64))
    res = MpirResultC()
    history = np.zeros(4 * max(1, config.max_outer_iterations))
    code = load().pcbhip_solve_mpir(op.handle, _ptr(b), _ptr(x), ctypes.byref(config), int(bool(inner_fp64)),
                                    ctypes.byref(res), _ptr(history))
    check(code)
    return x, result_dict(res, history)


class Server:
    """Owning handle of a continuous-batching server (``pcbhip_server``).

    ``submit`` queues one right-hand side of a resident operator, ``step``
    is one replay boundary (retire, admit, launch), ``finished`` lists ended
    slots and ``collect`` copies a result out and frees its slot."""

    def __init__(self, device: int = 0, max_slots: int = 256, inner_fp64: bool = False, steps_per_replay: int = 16):
        lib = load()
        self._lib = lib
        self.handle = None
        h = _vp()
        check(lib.pcbhip_server_create(int(device), int(max_slots), int(bool(inner_fp64)), int(steps_per_replay),
                                       ctypes.byref(h)))
        self.handle = h
        self.max_slots = int(max_slots)
        self._slots: dict[int, tuple[Operator, int]] = {}  # slot -> (operator kept alive, max_outer)
        self._ids = np.zeros(self.max_slots, dtype=np.int32)
        self._tags = np.zeros(self.max_slots, dtype=np.int64)

    def close(self) -> None:
        if self.handle is not None:
            self._lib.pcbhip_server_destroy(self.handle)
            self.handle = None
            self._slots.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def set_shape(self, rows_inner: int, target_groups: int) -> None:
        check(self._lib.pcbhip_server_set_shape(self.handle, int(rows_inner), int(target_groups)))

    def slot_bytes(self, op: Operator, max_outer: int) -> int:
        n = ctypes.c_int64(0)
        check(self._lib.pcbhip_server_slot_bytes(self.handle, op.handle, int(max_outer), ctypes.byref(n)))
        return int(n.value)

    def submit(self, op: Operator, rhs, x0=None, config: MpirConfigC | None = None, tag: int = 0) -> int:
        b = op._vec(rhs, np.float64)
        x = None if x0 is None else op._vec(x0, np.float64)
        cfg = config if config is not None else config_c(1e-10, 0.0, 2e-3, 8, 200)
        slot = _int(-1)
        check(self._lib.pcbhip_server_submit(self.handle, op.handle, _ptr(b), None if x is None else _ptr(x),
                                             ctypes.byref(cfg), int(tag), ctypes.byref(slot)))
        self._slots[int(slot.value)] = (op, int(cfg.max_outer_iterations))
        return int(slot.value)

    def step(self) -> tuple[int, int]:
        running, finished = _int(0), _int(0)
        check(self._lib.pcbhip_server_step(self.handle, ctypes.byref(running), ctypes.byref(finished)))
        return int(running.value), int(finished.value)

    def finished(self) -> list[tuple[int, int]]:
        n = _int(0)
        check(self._lib.pcbhip_server_finished(self.handle, _ptr(self._ids), _ptr(self._tags), self.max_slots,
                                               ctypes.byref(n)))
        return [(int(self._ids[i]), int(self._tags[i])) for i in range(n.value)]

    def collect(self, slot: int, *, solution: bool = True):
        op, max_outer = self._slots[slot]
        x = np.empty(op.size) if solution else None
        res = MpirResultC()
        history = np.zeros(4 * max(1, max_outer))
        code = self._lib.pcbhip_server_collect(self.handle, int(slot), None if x is None else _ptr(x),
                                               ctypes.byref(res), _ptr(history))
        if code in (OK, RUNTIME):  # collected (a failed solve frees its slot too)
            self._slots.pop(slot, None)
        check(code)
        return x, result_dict(res, history)

    def info(self) -> dict:
        st = ServerInfoC()
        check(self._lib.pcbhip_server_stats(self.handle, ctypes.byref(st)))
        return {name: (float if kind is ctypes.c_double else int)(getattr(st, name))
                for name, kind in ServerInfoC._fields_}

    def trim(self) -> None:
        check(self._lib.pcbhip_server_trim(self.handle))
