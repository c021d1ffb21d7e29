"""ORACLE -- test infrastructure only.  ctypes bindings of ``q1hex_oracle.c``
(the C restatement of the reference's ``native/hex_q1.cpp``), built by
``oracle/Makefile`` into ``oracle/_build/libq1hex_oracle.so``."""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "libq1hex_oracle.so"
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_int = ctypes.c_int
_dbl = ctypes.c_double


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def use_native_build() -> str:
    """Build the port for this host's ISA (``make native``, -march=native as
    the reference's default build) and load that build; returns which build
    is in use.  Falls back to the portable build when the host cannot
    compile it.  Call before the first use of the library."""

    global LIB_PATH
    if _lib is not None:
        return "-march=native" if LIB_PATH.parent.name == "native" else "portable x86-64-v3 (already loaded)"
    try:
        subprocess.run(["make", "-s", "-C", str(_HERE), "native"], check=True, capture_output=True, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return "portable x86-64-v3 (the native build failed on this host)"
    LIB_PATH = _HERE / "_build" / "native" / "libq1hex_oracle.so"
    return "-march=native"


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        for suf, fp in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(L, f"oracle_apply_{suf}")
            f.argtypes = [fp, fp, fp, _u8p, fp, _int, _int, _int, fp, fp, _int]
            f.restype = None
            g = getattr(L, f"oracle_pcg_{suf}")
            g.argtypes = [fp, fp, fp, _u8p, fp, _int, _int, _int, _f64p, fp, _dbl, _int, fp,
                          ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_dbl), _int]
            g.restype = _int
        L.oracle_solve_mpir.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                        _u8p, _int, _int, _int, _f64p, _f64p, _dbl, _dbl, _dbl, _int, _int,
                                        _int, _int, _i32p, _f64p]
        L.oracle_solve_mpir.restype = _dbl
        L.oracle_pcg_two_level_f32.argtypes = [_f32p, _f32p, _f32p, _u8p, _f32p, _int, _int, _int, _f64p, _f32p,
                                               _int, _f32p, _dbl, _int, _f32p, ctypes.POINTER(_int),
                                               ctypes.POINTER(_int), ctypes.POINTER(_dbl), _int]
        L.oracle_pcg_two_level_f32.restype = _int
        L.oracle_solve_mpir_two_level.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p, _f64p,
                                                  _f64p, _u8p, _int, _int, _int, _int, _f32p, _f64p, _f64p, _dbl,
                                                  _dbl, _dbl, _int, _int, _int, _i32p, _f64p]
        L.oracle_solve_mpir_two_level.restype = _dbl
        L.oracle_lane_sum.argtypes = [_f64p, ctypes.c_int64]
        L.oracle_lane_sum.restype = _dbl
        L.oracle_openmp.restype = _int
        _lib = L
    return _lib


def threads_default() -> int:
    return max(1, len(os.sched_getaffinity(0)))


class OracleOperator:
    """Both precisions of one prepared operator, as the reference's
    ``ThermalSystem`` holds them (thermal_system.cpp:354-362)."""

    def __init__(self, ops: dict, *, stiffness: bool = False):
        s, r, c = ops["element_shape"]
        self.dims = (s, r, c)
        self.size = (s + 1) * (r + 1) * (c + 1)
        coef = np.ascontiguousarray(ops["coef"], dtype=np.float64).reshape(-1)
        unit = np.ascontiguousarray(ops["unit"], dtype=np.float64).reshape(-1)
        if stiffness:
            robin = np.zeros(self.size)
            free = np.ones(self.size, dtype=bool)
        else:
            robin = np.asarray(ops["robin_total"], dtype=np.float64)
            free = np.asarray(ops["free"], dtype=bool)
        self.free_u8 = np.ascontiguousarray(free, dtype=np.uint8)
        self.coef64, self.unit64 = coef, unit
        self.robin64 = np.ascontiguousarray(robin)
        self.mask64 = np.ascontiguousarray(free, dtype=np.float64)
        self.coef32 = coef.astype(np.float32)
        self.unit32 = unit.astype(np.float32)
        self.robin32 = self.robin64.astype(np.float32)
        self.mask32 = self.mask64.astype(np.float32)
        if "diag" in ops and not stiffness:
            self.diag64 = np.ascontiguousarray(ops["diag"], dtype=np.float64)
            self.diag32 = self.diag64.astype(np.float32)

    def apply_f64(self, x, threads=None):
        y = np.empty(self.size)
        lib().oracle_apply_f64(self.coef64, self.unit64, self.robin64, self.free_u8, self.mask64, *self.dims,
                               np.ascontiguousarray(x, dtype=np.float64), y, threads or threads_default())
        return y

    def apply_f32(self, x, threads=None):
        y = np.empty(self.size, dtype=np.float32)
        lib().oracle_apply_f32(self.coef32, self.unit32, self.robin32, self.free_u8, self.mask32, *self.dims,
                               np.ascontiguousarray(x, dtype=np.float32), y, threads or threads_default())
        return y

    @property
    def node_shape(self):
        s, r, c = self.dims
        return (s + 1, r + 1, c + 1)

    def two_level(self, block=None, threads=None) -> dict:
        """The two-level coarse stage of this operator: patch width (the
        default rule unless given), Galerkin matrix from 27 FP64 applies and
        its dense inverse, rounded to FP32 for the inner solve."""

        from . import two_level as tl

        block = tl.block_size(self.node_shape) if block is None else int(block)
        matrix = tl.coloured_galerkin(lambda v: self.apply_f64(v, threads), self.node_shape,
                                      self.free_u8.astype(bool), block)
        inverse = tl.spd_inverse(matrix)
        return dict(block=block, matrix=matrix, inverse32=np.ascontiguousarray(inverse, dtype=np.float32))

    def pcg_f32(self, rhs, tol, max_its, threads=None, coarse=None):
        x = np.empty(self.size, dtype=np.float32)
        its, apps, rel = _int(), _int(), _dbl()
        args = (self.coef32, self.unit32, self.robin32, self.free_u8, self.mask32, *self.dims,
                np.ascontiguousarray(rhs, dtype=np.float64), self.diag32)
        tail = (tol, max_its, x, ctypes.byref(its), ctypes.byref(apps), ctypes.byref(rel),
                threads or threads_default())
        if coarse is None:
            bad = lib().oracle_pcg_f32(*args, *tail)
        else:
            bad = lib().oracle_pcg_two_level_f32(*args, coarse["block"], coarse["inverse32"], *tail)
        if bad:
            raise RuntimeError("inner PCG requires a finite symmetric positive-definite operator")
        return x, its.value, rel.value, apps.value

    def solve_mpir(self, rhs, x0=None, *, rtol=1e-10, atol=0.0, inner_tol=2e-3, max_outer=8, max_inner=200,
                   inner_f64=False, threads=None, coarse=None):
        """Returns (solution, dict of the MPIR result fields); ``coarse``
        (from ``two_level``) makes the FP32 inner PCG two-level."""

        x = np.zeros(self.size) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
        stats = np.zeros(6, dtype=np.int32)
        history = np.zeros(4 * max(1, max_outer))
        ops = (self.coef32, self.unit32, self.robin32, self.mask32, self.diag32, self.coef64, self.unit64,
               self.robin64, self.mask64, self.diag64, self.free_u8, *self.dims)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        team = threads or threads_default()
        if coarse is None:
            rel = lib().oracle_solve_mpir(*ops, rhs, x, rtol, atol, inner_tol, max_outer, max_inner, int(inner_f64),
                                          team, stats, history)
        else:
            if inner_f64:
                raise ValueError("the two-level inner solve is FP32")
            rel = lib().oracle_solve_mpir_two_level(*ops, coarse["block"], coarse["inverse32"], rhs, x, rtol, atol,
                                                    inner_tol, max_outer, max_inner, team, stats, history)
        if stats[5]:
            raise RuntimeError("inner PCG requires a finite SPD operator")
        rows = history.reshape(-1, 4)[: int(stats[1])] if not stats[0] else history.reshape(-1, 4)[: int(stats[1])]
        return x, dict(
            converged=bool(stats[0]),
            outer_iterations=int(stats[1]),
            inner_iterations=int(stats[2]),
            high_operator_applications=int(stats[3]),
            low_operator_applications=int(stats[4]),
            relative_residual=float(rel),
            history=[(int(a), float(b), int(c), float(d)) for a, b, c, d in rows],
        )
