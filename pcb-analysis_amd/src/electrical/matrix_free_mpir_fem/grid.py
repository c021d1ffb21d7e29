"""Per-axis pitch validation shared by the structured meshes
(reference: src/electrical/matrix_free_mpir_fem/grid.py, check_pitch_axis)."""

from __future__ import annotations

import numpy as np


def check_pitch_axis(pitch, count: int, name: str) -> np.ndarray:
    """One positive finite width per cell: a scalar is repeated ``count`` times."""

    values = np.asarray(pitch, dtype=np.float64)
    if values.ndim == 0:
        values = np.full(int(count), float(values))
    elif values.shape != (int(count),):
        raise ValueError(f"{name} must be a scalar or hold one value per cell ({count})")
    if not np.all(np.isfinite(values)) or np.any(values <= 0.0):
        raise ValueError(f"{name} must be finite and positive")
    return values.copy()
