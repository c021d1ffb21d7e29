"""Test configuration: import paths and the ``gpu`` marker.

Tests marked ``gpu`` need an MI355X and call the HIP library through its C
ABI; everything else runs on the CPU (the oracle, host logic, and that the
library loads and exports its symbols).
"""

from __future__ import annotations

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for path in (ROOT / "pcb-analysis_amd" / "src", ROOT):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
