"""Builds libgga.so and gga_run in-tree (hipcc for gfx950 + gcc), via csrc/Makefile."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> None:
    """Compiles every HIP kernel for gfx950 and the C host layer (always rebuilds when force)."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


if __name__ == "__main__":
    build()
    print(os.path.join(_HERE, "lib", "libgga.so"))
