"""Build libhrrt_osem.so in-tree with hipcc for gfx950 (csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force: bool = True) -> str:
    args = ["make", "-s", "-C", os.path.join(HERE, "csrc")]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "libhrrt_osem.so")


if __name__ == "__main__":
    print(build())
