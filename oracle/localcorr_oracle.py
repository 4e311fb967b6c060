"""CPU ORACLE -- test infrastructure only. The local-correlation calibration restated in plain NumPy loops, with the
model's prices taken from paths the caller supplies (the oracle's path engine in the tests), which the engine's
calibration is checked against.

    the spread lines (CvC, PvP on the equally weighted basket) ... packages/structured/src/optlab/structured/
                                                                   localcorr.py:280-285 (spread_payoff), :141-150
                                                                   (model_quotes: pair averages, in %)
    the fit ..................................................... localcorr.py:153-178 (_fit: Gauss-Newton on
                                                                   common random numbers)
    the local form's parameters, start and bounds ............... localcorr.py:186-189
"""

from __future__ import annotations

import numpy as np

W0 = 0.3
START = (0.0, 0.1, 0.1)
BOUNDS = ((-W0, 0.0, 0.0), (1.0 - W0, 1.5, 1.5))


def spread(kind, strike, x):
    """Per path: the option on the mean of the row less the mean of the options on its entries."""
    out = np.empty(len(x))
    for i, row in enumerate(np.asarray(x, dtype=float)):
        b = sum(row) / len(row)
        if kind == "cvc":
            out[i] = max(b - strike, 0.0) - sum(max(v - strike, 0.0) for v in row) / len(row)
        else:
            out[i] = max(strike - b, 0.0) - sum(max(strike - v, 0.0) for v in row) / len(row)
    return out


def quotes(x, lines):
    """(prices in %, standard errors) of the lines on the pairs of x (rows p and n/2 + p)."""
    h = len(x) // 2
    prices, errs = [], []
    for line in lines:
        pay = 100.0 * spread(line["kind"], line["strike"], x)
        pairs = 0.5 * (pay[:h] + pay[h:])
        prices.append(pairs.mean())
        errs.append(pairs.std(ddof=1) / np.sqrt(h))
    return np.array(prices), np.array(errs)


def fit(prices_of, goal, start=START, bounds=BOUNDS, steps=6, h=0.02):
    """(best parameters, RMS error): localcorr.py:153-178, step for step."""
    goal = np.array(goal, dtype=float)
    p = np.array(start, dtype=float)
    lower, upper = np.array(bounds[0], dtype=float), np.array(bounds[1], dtype=float)
    best, best_err = p.copy(), float("inf")
    for _ in range(steps):
        f, noise = prices_of(p)
        r = f - goal
        err = float(np.sqrt(np.mean(r * r)))
        if err < best_err:
            best, best_err = p.copy(), err
        if err < float(np.mean(noise)):
            break
        J = np.zeros((len(goal), len(p)))
        for j in range(len(p)):
            q = p.copy()
            if q[j] + h <= upper[j]:
                q[j] = q[j] + h
            else:
                q[j] = q[j] - h
            g, _ = prices_of(q)
            J[:, j] = (g - f) / (q[j] - p[j])
        step = np.linalg.lstsq(J, -r, rcond=1e-3)[0]
        p = np.minimum(np.maximum(p + 0.8 * step, lower), upper)
    return best, best_err
