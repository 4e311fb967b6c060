"""Writes tests/golden/risk_tents.npz: input/output vectors of the reference's risk-grid tents
(packages/risk/src/optlab/risk/grid.py: `smoothing_width` :102-113, `strike_tents` :116-148, `tenor_tents` :151-153),
the shape every node vega bumps. Nothing else fixes the smoothing's constants and algebra (a wrong width, floor or
ramp passes every partition-of-unity and finite-difference check), so these vectors are taken from the reference
itself: grid.py is loaded by its file path (it imports only NumPy, and SciPy's ndtr inside `strike_tents`).

Run where the reference is mounted (it is read from /root/reference, not copied):
    python tests/golden/make_tent_fixtures.py
"""

from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np

GRID = Path("/root/reference/packages/risk/src/optlab/risk/grid.py")
OUT = Path(__file__).resolve().parent / "risk_tents.npz"


def points() -> tuple[np.ndarray, np.ndarray]:
    """(k, T): every log-strike node and points between and beyond them, at maturities below the width's floor, on
    it, on and between the tenor nodes and beyond the last."""
    log_strikes = np.log([0.6, 0.7, 0.8, 0.9, 0.95, 1.0, 1.05, 1.1, 1.2, 1.4])
    k = np.concatenate([log_strikes, np.linspace(-1.2, 0.7, 19), [-0.02, 0.013, 0.31]])
    T = np.array([0.004, 1 / 52, 0.03, 1 / 12, 0.17, 0.25, 0.4, 1.0, 1.23, 2.0, 3.7, 5.0, 6.5])
    kk, TT = np.meshgrid(k, T, indexing="ij")
    return kk.ravel(), TT.ravel()


def main() -> None:
    spec = importlib.util.spec_from_file_location("reference_risk_grid", GRID)
    grid = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(grid)
    k, T = points()
    phi, phi_k, phi_kk, phi_T = grid.strike_tents(k, T)
    ell, ell_k, ell_kk, ell_T = grid.smoothing_width(k, T)
    Tt = np.unique(np.concatenate([T, np.asarray(grid.TENORS), np.linspace(0.0, 6.0, 61)]))
    psi, dpsi = grid.tenor_tents(Tt, 1)
    np.savez_compressed(
        OUT,
        k=k, T=T,
        strike_tents=np.stack([phi, phi_k, phi_kk, phi_T]),
        smoothing_width=np.stack([ell, ell_k, ell_kk, ell_T]),
        tenor_T=Tt, tenor_tents=np.stack([psi, dpsi]),
    )
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes): {k.size} (k, T) points, {Tt.size} tenor points")


if __name__ == "__main__":
    main()
