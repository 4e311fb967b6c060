"""The terms of the issuer's call regression (packages/structured/src/optlab/structured/exercise.py:108-140), for
`optlab.native.regress`: which monomials of the normalised state and which hinges a fit gets, as the number of
paths reaching the call allows. The regression itself runs on the device (optlab_regress)."""

from __future__ import annotations

import itertools

import numpy as np

MAX_TERMS = 25  # terms of the regression, the constant included
MAX_KNOTS = 9  # hinges per state variable, at its deciles
MIN_PATHS_PER_TERM = 20  # fewer reaching paths than this per term: fewer hinges, a lower degree


def powers(d: int, degree: int) -> list[tuple[int, ...]]:
    """The exponent tuples of the monomials of total degree <= degree in d variables: the constant first, then by
    degree, each degree's tuples in the order of the variables' multisets."""
    out = []
    for total in range(degree + 1):
        for combo in itertools.combinations_with_replacement(range(d), total):
            exps = [0] * d
            for i in combo:
                exps[i] += 1
            out.append(tuple(exps))
    return out


def terms(z: np.ndarray) -> tuple[list, list]:
    """(powers, knots) for normalised states z (paths x variables): polynomials of degree 3 in one variable, 2 in up
    to four, 1 beyond (lower when the room of min(MAX_TERMS, paths / MIN_PATHS_PER_TERM) terms is short), then up to
    MAX_KNOTS hinges a variable at its quantiles, a level equal to a variable's end dropped."""
    n, d = z.shape
    room = min(MAX_TERMS, n // MIN_PATHS_PER_TERM)
    for degree in (3 if d == 1 else 2 if d <= 4 else 1, 1, 0):
        pw = powers(d, degree)
        if len(pw) <= room:
            break
    per = max(0, min(MAX_KNOTS, (room - len(pw)) // d))
    knots = []
    for i in range(d):
        if not per:
            continue
        levels = np.unique(np.quantile(z[:, i], np.linspace(0.0, 1.0, per + 2)[1:-1]))
        lo, hi = z[:, i].min(), z[:, i].max()
        knots += [(i, float(k)) for k in levels if lo < k < hi]
    return pw, knots
