"""Build librmesh_hip.so in-tree with hipcc for gfx950 (rmeshvk_amd/csrc/Makefile)."""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 4, force: bool = True) -> str:
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", csrc, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "librmesh_hip.so")


if __name__ == "__main__":
    print(build())
