"""Build liblepic_hip.so for gfx950 with hipcc (csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 5, arch: str = "gfx950") -> str:
    env = dict(os.environ)
    subprocess.run(["make", "-s", "-B", f"-j{jobs}", "-C", CSRC, f"ARCH={arch}"], check=True, env=env)
    lib = os.path.join(HERE, "_lib", "liblepic_hip.so")
    if not os.path.exists(lib):
        raise RuntimeError("build finished without producing liblepic_hip.so")
    return lib


if __name__ == "__main__":
    print(build())
