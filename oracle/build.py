"""Build recipe for the CPU FP64 oracle (test infrastructure only).

The oracle is compiled with the CPU-baseline flags: -O3 -fno-fast-math
-ffp-contract=off, so its arithmetic is IEEE double with no contraction.

build_reference() compiles the reference's own calibration fitter,
tools/cloth_calibration.cpp (one source file, standard headers only), from
where it lies in the reference checkout into _ref/, when that checkout is
present. Tests use it only to check that the observation CSV this library
writes is read and fitted by the reference's fitter.
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


REFERENCE = Path("/root/reference")
FITTER_SOURCE = REFERENCE / "tools" / "cloth_calibration.cpp"
FITTER = ORACLE / "_ref" / "numi-solver-cloth-calibration"


def build_reference(verbose: bool = False) -> Path | None:
    """The reference fitter into _ref/ (None when the reference checkout is absent)."""
    if not FITTER_SOURCE.exists():
        return None
    FITTER.parent.mkdir(exist_ok=True)
    temporary = FITTER.with_name(FITTER.name + ".tmp")
    command = ["g++", "-std=c++20", "-O2", str(FITTER_SOURCE), "-o", str(temporary)]
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("reference fitter build failed")
    os.replace(temporary, FITTER)
    if verbose:
        print(f"built {FITTER}")
    return FITTER


if __name__ == "__main__":
    build(verbose=True)
    build_reference(verbose=True)
