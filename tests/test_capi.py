"""The C-ABI library loads and exports every symbol include/pcbhip.h declares
(CPU only: nothing here launches a kernel)."""

from __future__ import annotations

import re
from pathlib import Path

import pytest

from pcbhip import _capi

HEADER = Path(__file__).resolve().parents[1] / "include" / "pcbhip.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    return set(re.findall(r"\b(pcbhip_[a-z0-9_]+)\s*\(", text))


def test_header_and_binding_agree():
    assert declared_functions() == set(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_device_count_is_callable_without_a_device():
    assert _capi.device_count() >= 0


def test_cpu_backend_is_refused():
    from electrical.matrix_free_mpir_fem import make_float32_runtime

    with pytest.raises(ValueError, match="HIP device only"):
        make_float32_runtime("cpu")


def test_errors_map_to_python_exceptions():
    lib = _capi.load()
    code = lib.pcbhip_operator_create(0, 0, 1, 1, None, None, None, None)
    assert code == _capi.INVALID_INPUT
    with pytest.raises(ValueError, match="null argument"):
        _capi.check(code)


def test_basis_products_check_their_arguments_before_any_device_work():
    import ctypes

    lib = _capi.load()
    out = (ctypes.c_double * 1)()
    coef = (ctypes.c_double * 1)(1.0)
    rows = (ctypes.c_void_p * 1)(None)
    assert lib.pcbhip_vec_project(0, None, 1, None, 4, out) == _capi.INVALID_INPUT
    assert lib.pcbhip_vec_project(0, rows, 0, None, 4, out) == _capi.INVALID_INPUT
    assert lib.pcbhip_vec_project(0, rows, 65, None, 4, out) == _capi.INVALID_INPUT
    assert lib.pcbhip_vec_combine(0, rows, coef, 65, None, None, 4) == _capi.INVALID_INPUT
    assert lib.pcbhip_vec_combine(0, rows, None, 1, None, None, 4) == _capi.INVALID_INPUT
    # A null row of a non-empty vector is refused.
    assert lib.pcbhip_vec_combine(0, rows, coef, 1, None, None, 4) == _capi.INVALID_INPUT
    with pytest.raises(ValueError, match="1..64 rows"):
        _capi.check(lib.pcbhip_vec_project(0, rows, 65, None, 4, out))
