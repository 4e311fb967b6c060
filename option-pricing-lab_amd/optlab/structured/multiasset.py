"""Baskets of up to eight names for the path engine (packages/structured/src/optlab/structured/multiasset.py:20-75).

Spot drivers are z = root @ raw with root root^T = corr from an eigen-decomposition (valid for singular matrices).
A name with a variance process drives it by rho_i z_i + sqrt(1 - rho_i^2) w_i, with w correlated like the spots
("same") or independent."""

from __future__ import annotations

import numpy as np


def nearest_correlation(c, iterations: int = 50) -> np.ndarray:
    """The closest PSD unit-diagonal matrix by alternating projections (packages/core/src/optlab/core/
    correlation.py:9-26)."""
    x = np.array(c, dtype=float)
    x = 0.5 * (x + x.T)
    for _ in range(iterations):
        lam, vec = np.linalg.eigh(x)
        if lam.min() >= -1e-12 and np.allclose(np.diag(x), 1.0, atol=1e-12):
            break
        x = (vec * np.maximum(lam, 0.0)) @ vec.T
        np.fill_diagonal(x, 1.0)
    lam, vec = np.linalg.eigh(x)
    if lam.min() < 0:
        x = (vec * np.maximum(lam, 1e-10)) @ vec.T
        d = np.sqrt(np.diag(x))
        x = x / np.outer(d, d)
    x = 0.5 * (x + x.T)
    np.fill_diagonal(x, 1.0)
    return x


def correlation_root(corr) -> np.ndarray:
    """L with L L^T = corr, by eigen-decomposition."""
    c = np.atleast_2d(np.asarray(corr, dtype=float))
    if c.shape[0] != c.shape[1] or not np.allclose(c, c.T, atol=1e-12):
        raise ValueError("the correlation matrix must be square and symmetric")
    if not np.allclose(np.diag(c), 1.0, atol=1e-12):
        raise ValueError("the correlation matrix must have ones on its diagonal")
    lam, vec = np.linalg.eigh(c)
    if lam.min() < -1e-10:
        raise ValueError(f"the correlation matrix is not positive semi-definite (eigenvalue {lam.min():.2e})")
    if c.shape[0] == 1:
        return np.ones((1, 1))
    return vec * np.sqrt(np.maximum(lam, 0.0))


class Basket:
    def __init__(self, names, models: dict, markets: dict, corr, vol_corr: str = "same", mixer=None):
        if vol_corr not in ("same", "independent"):
            raise ValueError(f"unknown vol_corr {vol_corr!r}, expected same or independent")
        self.names = tuple(names)
        missing = [n for n in self.names if n not in models or n not in markets]
        if missing:
            raise ValueError(f"no model or market for {', '.join(missing)}")
        self.models, self.markets, self.vol_corr, self.mixer = dict(models), dict(markets), vol_corr, mixer
        self.corr = np.atleast_2d(np.asarray(corr, dtype=float))
        if self.corr.shape != (len(self.names),) * 2:
            raise ValueError(f"the correlation matrix must be {len(self.names)}x{len(self.names)}")
        self.root = correlation_root(self.corr)
