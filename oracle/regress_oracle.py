"""Test infrastructure (the oracle): the issuer's call regression of exercise.py, in NumPy, FP64.

Restates packages/structured/src/optlab/structured/exercise.py:108-123 (_powers), :126-140 (_basis: which terms a
fit gets), :150-171 (Rule.features, Rule.continuation) and the solve of :219 (fit), here by the normal equations
(as the device) or by np.linalg.lstsq (the reference). z^3 is formed as z z z (the device's form; the reference
writes z ** p).

Pinned in tests/test_oracle_regress.py: the terms against hand values (products of powers, hinges); _powers against
a brute-force count of the monomials; the normal equations against lstsq on well-conditioned data and an exact
polynomial recovered; _basis's choice of degree and hinges against the reference's rules (room, MAX_TERMS)."""

from __future__ import annotations

import itertools

import numpy as np

MAX_TERMS, MAX_KNOTS, MIN_PATHS_PER_TERM = 25, 9, 20  # exercise.py:28-30


def powers(d: int, degree: int) -> list[tuple[int, ...]]:
    """Exponents of the monomials of total degree <= degree in d variables, constant first, lowest degrees first."""
    out = []
    for total in range(degree + 1):
        for combo in itertools.combinations_with_replacement(range(d), total):
            p = [0] * d
            for i in combo:
                p[i] += 1
            out.append(tuple(p))
    return out


def basis(z: np.ndarray) -> tuple[list, list]:
    """(powers, knots) of a fit on normalised states z (paths x variables), as exercise.py:126-140 chooses them."""
    n, d = z.shape
    room = min(MAX_TERMS, n // MIN_PATHS_PER_TERM)
    for degree in (3 if d == 1 else 2 if d <= 4 else 1, 1, 0):
        pw = powers(d, degree)
        if len(pw) <= room:
            break
    per = max(0, min(MAX_KNOTS, (room - len(pw)) // d))
    knots = []
    for i in range(d):
        levels = np.unique(np.quantile(z[:, i], np.linspace(0.0, 1.0, per + 2)[1:-1])) if per else []
        lo, hi = z[:, i].min(), z[:, i].max()
        knots += [(i, float(k)) for k in levels if lo < k < hi]
    return pw, knots


def _cube(z, p):
    return np.where(p == 0, 1.0, np.where(p == 1, z, np.where(p == 2, z * z, z * z * z)))


def features(state, mean, scale, pw, knots) -> np.ndarray:
    """Rule.features: the terms of each state row (rows x terms)."""
    z = (np.atleast_2d(np.asarray(state, dtype=float).T).T - np.asarray(mean)) / np.asarray(scale)
    cols = []
    for p in pw:
        f = np.ones(len(z))
        for i, e in enumerate(p):
            f = f * _cube(z[:, i], e)
        cols.append(f)
    cols += [np.maximum(z[:, i] - k, 0.0) for i, k in knots]
    return np.stack(cols, axis=1)


def fit(X, y, solver="normal"):
    """The coefficients of y on X: normal equations (np.linalg.solve of X'X) or the reference's lstsq."""
    if solver == "normal":
        return np.linalg.solve(X.T @ X, X.T @ y)
    return np.linalg.lstsq(X, y, rcond=None)[0]
