"""Build recipe for the CPU FP64 oracle (test infrastructure only).

The oracle is compiled with the CPU-baseline flags: -O3 -fno-fast-math
-ffp-contract=off, so its arithmetic is IEEE double with no contraction.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ORACLE = Path(__file__).resolve().parent
REPO = ORACLE.parent
LIBRARY = ORACLE / "libnumi_cloth_oracle.so"
SOURCES = ["cloth_oracle.cpp", "cloth_adjoint.cpp"]


def build(verbose: bool = False) -> Path:
    temporary = LIBRARY.with_suffix(".so.tmp")
    command = ["g++", "-O3", "-fno-fast-math", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared",
               "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{REPO / 'include'}",
               *[str(ORACLE / name) for name in SOURCES], "-o", str(temporary)]
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(temporary, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
