"""Builds librigidfibers.so (gfx950) in-tree with the Makefile in csrc/."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force: bool = True, jobs: int = 8) -> str:
    cmd = ["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    path = os.path.join(HERE, "librigidfibers.so")
    if not os.path.exists(path):
        raise RuntimeError(f"build did not produce {path}")
    return path


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
