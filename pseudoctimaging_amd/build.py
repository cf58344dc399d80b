"""Builds the in-tree native code: libpct.so + pct_cli (HIP for gfx950 and
C++ host code), and the test oracle under oracle/."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def _make(directory: str, *targets: str) -> None:
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-C", directory, "-j", jobs, *targets], check=True)


def build_library(force: bool = True) -> None:
    """Compiles libpct.so and pct_cli (always recompiles when force)."""
    if force:
        _make(os.path.join(PKG, "csrc"), "clean")
    _make(os.path.join(PKG, "csrc"), "all")


def build_oracle() -> None:
    """Compiles the C oracle and, when the reference tree is present, the
    reference CPU path into oracle/_ref (test infrastructure)."""
    _make(os.path.join(ROOT, "oracle"), "-B", "all")


def build() -> None:
    build_library(force=True)
    build_oracle()


if __name__ == "__main__":
    build()
