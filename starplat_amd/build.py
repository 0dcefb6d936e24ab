"""In-tree build of the native libraries (gfx950 HIP library, host data layer)
and, for the test suite, the CPU oracle. Uses the committed Makefiles."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "starplat_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


def _make(d, jobs=8, force=False):
    if force:
        subprocess.run(["make", "-C", d, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", d, f"-j{jobs}"], check=True)


def build(force=True, oracle=True, jobs=None):
    jobs = jobs or min(16, os.cpu_count() or 8)
    _make(CSRC, jobs, force)
    if oracle:
        _make(ORACLE, jobs, force)


if __name__ == "__main__":
    build()
