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
