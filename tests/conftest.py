"""Test configuration: the in-tree package on sys.path, the `gpu` marker, and the built libraries.

`-m "not gpu"` runs everywhere (the CPU oracle against the reference's golden vectors, the host plan layer, the C
ABI's exports and request checks); `-m gpu` runs the HIP engine against the oracle on a device."""

from __future__ import annotations

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "option-pricing-lab_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP engine")
    config.addinivalue_line("markers", "slow: minutes of CPU work")
