"""Packaging hook: the wheel carries liboptlab_mi.so, so building one compiles the HIP engine first (build.py, hipcc
for gfx950). Metadata, the entry point included, is in pyproject.toml."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

from setuptools import setup
from setuptools.command.build_py import build_py


class BuildWithEngine(build_py):
    def run(self):
        subprocess.run([sys.executable, str(Path(__file__).resolve().parent / "build.py")], check=True)
        super().run()


setup(cmdclass={"build_py": BuildWithEngine})
