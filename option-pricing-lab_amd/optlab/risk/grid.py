"""The risk grid: forward moneyness x tenors (packages/risk/src/optlab/risk/grid.py:25-30).

The tents themselves (Gaussian-smoothed strike hats, smootherstep tenor tents, grid.py:32-166) are evaluated on
the device by the engine's table builder; the host only names the nodes."""

from __future__ import annotations

import numpy as np

STRIKES = (0.6, 0.7, 0.8, 0.9, 0.95, 1.0, 1.05, 1.1, 1.2, 1.4)  # forward moneyness K / F
TENORS = (1 / 12, 0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 5.0)  # years
TENOR_LABELS = ("1m", "3m", "6m", "1y", "18m", "2y", "3y", "5y")
STRIKE_LABELS = tuple(f"{round(100 * m)}%" for m in STRIKES)
LOG_STRIKES = tuple(float(np.log(m)) for m in STRIKES)
SHAPE = (len(STRIKES), len(TENORS))


def node_vols(surface) -> np.ndarray:
    """Implied vols of a surface at the nodes, (n_strikes, n_tenors) (grid.py:175-183)."""
    k = np.broadcast_to(np.asarray(LOG_STRIKES)[:, None], SHAPE)
    T = np.broadcast_to(np.asarray(TENORS)[None, :], SHAPE)
    return np.asarray(surface.iv(k, T), dtype=float)
