"""Build the gfx950 HIP library ``libpcbhip.so`` in-tree with hipcc.

The library is written next to the Python package that loads it
(``pcb-analysis_amd/src/pcbhip/libpcbhip.so``).  Every HIP source is
compiled for gfx950 only, with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["q1hex_kernels.hip", "pcbhip_runtime.hip"]
OUTPUT = HERE.parent / "src" / "pcbhip" / "libpcbhip.so"
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # No FMA contraction: every rounding is written in the source, so the
    # same element gives the same bits whatever tile shape or template
    # instantiation computes it (the reference builds its kernels the same
    # way: -ffp-contract=off, --fmad=false).
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


# rocSOLVER (and its rocBLAS) factor and invert the two-level coarse
# matrices on the device.
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(output: Path = OUTPUT) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = output.with_suffix(".so.tmp")
    cmd = [hipcc, *FLAGS, "-o", str(tmp), *[str(HERE / s) for s in SOURCES], *LIBS]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, output)
    return output


if __name__ == "__main__":
    print(build())
