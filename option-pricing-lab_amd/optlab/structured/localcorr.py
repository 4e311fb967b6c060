"""Local correlation for the path engine's mixer (packages/structured/src/optlab/structured/localcorr.py:37-114).

Z_i = sqrt(w) Z_0 + sqrt(1 - w) (L_low e)_i with w = clip(w0 + level + down max(-u, 0) - up max(u, 0), 0, 1) and
u = y / (sigma sqrt(max(t, 1 month))), y = ln(mean S_i / F_i) at the start of the step. L_low factors the matrix that
gives corr back at w = w0; its factors take the signs of the basket's root so that the mixer is on common random
numbers with the constant matrix where the weight has not moved."""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from optlab.structured.multiasset import correlation_root, nearest_correlation

W0 = 0.3


def low_matrix(corr, w0: float) -> np.ndarray:
    c = np.asarray(corr, dtype=float)
    low = (c - w0) / (1.0 - w0)
    np.fill_diagonal(low, 1.0)
    return nearest_correlation(np.clip(low, -0.999, 0.999))


def aligned_factors(corr, root, w0: float) -> tuple:
    """(low factors signed like `root`, sign of the market factor, angle of the common factor)."""
    low = correlation_root(low_matrix(corr, w0)).copy()
    for j in range(low.shape[1] - 1):
        if np.dot(low[:, j], root[:, j]) < 0:
            low[:, j] = -low[:, j]
    if low[:, -1].mean() < 0:
        low[:, -1] = -low[:, -1]
    sign = 1.0 if root[:, -1].mean() >= 0 else -1.0
    return low, sign, float(np.arctan2(np.sqrt(1.0 - w0) * low[:, -1].mean(), np.sqrt(w0)))


@dataclass(frozen=True)
class LocalCorrelation:
    sigma: float
    level: float = 0.0
    down: float = 0.0
    up: float = 0.0
    w0: float = W0

    def bumped(self, slope: float) -> LocalCorrelation:
        return replace(self, down=max(self.down + slope, 0.0), up=max(self.up + slope, 0.0))


@dataclass(frozen=True)
class StochasticMix:
    """The weight as a Jacobi process (localcorr.py:87-114): dw = kappa (theta - w) dt + nu sqrt(w (1 - w)) dB,
    clipped to [0, 1], with dB = rho (market factor) + sqrt(1 - rho^2) (the weight's own driver); the step's
    normals are mixed with the weight at the step's start. In the path engine as mixer 2 (the reference engine
    refuses it); the driver is a stream of its own (tag 0x40000000 of the pair's Philox counter)."""

    theta: float = W0
    nu: float = 0.5
    rho: float = -0.5
    kappa: float = 2.0
    w0: float = W0

    def bumped(self, slope: float) -> StochasticMix:
        """A steeper correlation skew for a positive slope: rho moved down by it, within [-0.99, 0.99]."""
        return replace(self, rho=float(np.clip(self.rho - slope, -0.99, 0.99)))
