"""Build libnsfem_hip.so in-tree for gfx950 (hipcc cross-compiles without a
GPU).  ``python -m navier_stokes_fem_amd.build`` or ``build()``."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnsfem_hip.so")


def build(jobs: int = 8, force: bool = True) -> str:
    """Compile every HIP/C++ source of the library.  ``force`` recompiles all
    translation units (the build check must compile every time)."""
    if force:
        subprocess.run(["make", "-C", CSRC, "-s", "clean"], check=True)
    subprocess.run(["make", "-C", CSRC, f"-j{int(jobs)}"], check=True)
    if not os.path.exists(LIB):
        raise RuntimeError("build finished without producing " + LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--incremental" not in sys.argv))
