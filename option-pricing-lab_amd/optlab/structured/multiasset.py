"""Baskets of up to eight names for the path engine (the reference's basket: packages/structured/src/optlab/
structured/multiasset.py:20-75).

The engine correlates the spot normals of a step as z = R e, e independent standard normals and R R^T the basket's
correlation. R is the symmetric eigen-root V diag(sqrt(lambda)) (it exists for singular matrices, where a Cholesky
factor does not). A name with a variance process drives it by rho_i z_i + sqrt(1 - rho_i^2) w_i, w correlated like
the spots ("same") or independent of them ("independent")."""

from __future__ import annotations

import numpy as np

_UNIT = 1e-12  # how far a diagonal entry or an asymmetry may be from exact and still count as a correlation matrix


def _is_correlation(c: np.ndarray) -> bool:
    """Unit diagonal and positive semi-definite, the smallest eigenvalue not below zero."""
    return bool(np.all(np.abs(np.diag(c) - 1.0) <= _UNIT) and np.linalg.eigvalsh(c)[0] >= 0.0)


def nearest_correlation(c, max_iter: int = 200, tol: float = 1e-12) -> np.ndarray:
    """The correlation matrix nearest to c in the Frobenius norm, by Higham's alternating projections with Dykstra's
    correction (N. J. Higham, "Computing the nearest correlation matrix -- a problem from finance", IMA J. Numer.
    Anal. 22 (2002), algorithm 3.3). A matrix that already is one comes back unchanged (symmetrised). The result is
    made exactly unit-diagonal and its eigenvalues are floored at zero, so that it always has a real root."""
    a = np.array(c, dtype=float)
    a = 0.5 * (a + a.T)
    if _is_correlation(a):
        return a
    y, dykstra = a.copy(), np.zeros_like(a)
    for _ in range(max_iter):
        r = y - dykstra
        lam, vec = np.linalg.eigh(r)
        x = (vec * np.clip(lam, 0.0, None)) @ vec.T  # onto the positive semi-definite cone
        dykstra = x - r
        y_next = x.copy()
        y_next[np.diag_indices_from(y_next)] = 1.0  # onto the unit-diagonal matrices
        moved = np.linalg.norm(y_next - y) / max(np.linalg.norm(y), 1.0)
        y = y_next
        if moved < tol:
            break
    lam, vec = np.linalg.eigh(0.5 * (y + y.T))
    x = (vec * np.clip(lam, 0.0, None)) @ vec.T
    scale = 1.0 / np.sqrt(np.diag(x))
    x = scale[:, None] * x * scale[None, :]
    x = 0.5 * (x + x.T)
    x[np.diag_indices_from(x)] = 1.0
    return x


def correlation_root(corr) -> np.ndarray:
    """R with R R^T = corr: the eigen-root V diag(sqrt(lambda)), eigenvalues in ascending order (so the last column
    is the dominant, market-like factor); eigenvalues within 1e-10 below zero count as zero."""
    c = np.atleast_2d(np.asarray(corr, dtype=float))
    n = c.shape[0]
    if c.ndim != 2 or c.shape != (n, n):
        raise ValueError(f"a correlation matrix is square, got shape {c.shape}")
    if np.max(np.abs(c - c.T)) > _UNIT:
        raise ValueError("a correlation matrix is symmetric; this one is not")
    if np.max(np.abs(np.diag(c) - 1.0)) > _UNIT:
        raise ValueError("a correlation matrix has a unit diagonal; this one does not")
    if n == 1:
        return np.ones((1, 1))
    lam, vec = np.linalg.eigh(c)
    if lam[0] < -1e-10:
        raise ValueError(f"a correlation matrix is positive semi-definite; this one has eigenvalue {lam[0]:.2e}")
    return vec * np.sqrt(np.clip(lam, 0.0, None))


class Basket:
    """Names, each with its model and market, their spot correlation and an optional mixer of the correlation
    (`optlab.structured.localcorr`). `root` is the correlation's eigen-root, formed once here."""

    def __init__(self, names, models: dict, markets: dict, corr, vol_corr: str = "same", mixer=None):
        if vol_corr not in ("same", "independent"):
            raise ValueError(f"vol_corr is 'same' or 'independent', not {vol_corr!r}")
        self.names = tuple(names)
        lacking = [n for n in self.names if n not in models or n not in markets]
        if lacking:
            raise ValueError(f"the basket lacks a model or a market for: {', '.join(lacking)}")
        self.models = {n: models[n] for n in self.names}
        self.markets = {n: markets[n] for n in self.names}
        self.vol_corr, self.mixer = vol_corr, mixer
        n = len(self.names)
        self.corr = np.atleast_2d(np.asarray(corr, dtype=float))
        if self.corr.shape != (n, n):
            raise ValueError(f"{n} names need a {n} x {n} correlation matrix, got {self.corr.shape}")
        self.root = correlation_root(self.corr)
