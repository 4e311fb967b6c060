"""Test infrastructure (the oracle): Longstaff-Schwartz on given paths, in NumPy, FP64.

Restates packages/core/src/optlab/core/products.py:105-169 (american_lsm, _basis) and :30-35 (_result) on paths it
is handed (the spots at the exercise dates and, for the basis with v, the variance factor): the backward sweep of
regressions of the discounted continuation value of the in-the-money paths, then the rule applied to a second set
of paths. `solver` "normal" forms and solves the normal equations (as the device does, in FP64); "lstsq" is the
reference's np.linalg.lstsq. s^3 is formed as s s s (the device's form; the reference writes s**3).

Pinned in tests/test_oracle_lsm.py: the two solvers give the same rule and price on well-conditioned data; the price
of a Bermudan put under Black-Scholes matches a binomial tree that allows exercise at the same dates (a lower bound
within Monte Carlo error and the rule's small suboptimality); the immediate-exercise floor."""

from __future__ import annotations

import numpy as np


def basis(s, v=None):
    cols = [np.ones_like(s), s, s * s, s * s * s]
    if v is not None:
        cols += [v, v * v, s * v]
    return np.column_stack(cols)


def lsm_rule(S, V, disc, K, call=False, min_itm=50, solver="normal"):
    """The regression coefficients of each date (None where the date exercises nothing), from pass-1 spots S
    (rows x dates) and variances V (or None)."""
    phi = 1.0 if call else -1.0
    s = S / K
    n_ex = s.shape[1]
    value = np.maximum(phi * (s[:, -1] - 1.0), 0.0) * disc[-1]
    coefs = [None] * n_ex
    for j in range(n_ex - 2, -1, -1):
        ex = np.maximum(phi * (s[:, j] - 1.0), 0.0)
        itm = ex > 0
        if itm.sum() > min_itm:
            X = basis(s[itm, j], None if V is None else V[itm, j])
            y = value[itm] / disc[j]
            if solver == "normal":
                beta = np.linalg.solve(X.T @ X, X.T @ y)
            else:
                beta, *_ = np.linalg.lstsq(X, y, rcond=None)
            coefs[j] = beta
            cont = X @ beta
            idx = np.flatnonzero(itm)[ex[itm] > cont]
            value[idx] = ex[idx] * disc[j]
    return coefs


def lsm_pay(S, V, disc, K, coefs, call=False):
    """products.py:137-150: K cash D(t_tau) of every path of the second set."""
    phi = 1.0 if call else -1.0
    s = S / K
    n_ex = s.shape[1]
    cash = np.maximum(phi * (s[:, -1] - 1.0), 0.0)
    tau = np.full(len(s), n_ex - 1)
    for j in range(n_ex - 2, -1, -1):
        ex = np.maximum(phi * (s[:, j] - 1.0), 0.0)
        itm = ex > 0
        if coefs[j] is None or not itm.any():
            continue
        X = basis(s[itm, j], None if V is None else V[itm, j])
        cont = X @ coefs[j]
        exercise = np.zeros(len(s), dtype=bool)
        exercise[np.flatnonzero(itm)[ex[itm] > cont]] = True
        cash = np.where(exercise, ex, cash)
        tau = np.where(exercise, j, tau)
    return K * cash * np.asarray(disc)[tau]


def result(pay):
    """products.py:30-35: antithetic pairs (rows i and i + n / 2) averaged first; (mean, standard error)."""
    d = np.asarray(pay)
    if len(d) % 2 == 0:
        d = 0.5 * (d[: len(d) // 2] + d[len(d) // 2 :])
    return float(d.mean()), float(d.std(ddof=1) / np.sqrt(len(d)))


def lsm_price(S1, V1, S2, V2, disc, K, spot, call=False, min_itm=50, solver="normal"):
    """(price, se, coefs): the rule from the first set, applied to the second; the immediate-exercise floor."""
    coefs = lsm_rule(S1, V1, disc, K, call, min_itm, solver)
    price, se = result(lsm_pay(S2, V2, disc, K, coefs, call))
    immediate = K * max((1.0 if call else -1.0) * (spot / K - 1.0), 0.0)
    if immediate > price:
        return immediate, 0.0, coefs
    return price, se, coefs


def bermudan_tree(S0, K, r, q, sigma, T, n_ex, steps_per_date=40, call=False, today=True):
    """Cox-Ross-Rubinstein tree with exercise allowed only at the dates T k / n_ex, k = 1..n_ex (the dates
    american_lsm uses), and today when `today`: the value the Longstaff-Schwartz estimate approaches from below."""
    n = n_ex * steps_per_date
    dt = T / n
    u = np.exp(sigma * np.sqrt(dt))
    d = 1.0 / u
    p = (np.exp((r - q) * dt) - d) / (u - d)
    disc = np.exp(-r * dt)
    phi = 1.0 if call else -1.0
    S = S0 * u ** np.arange(n, -1, -1) * d ** np.arange(0, n + 1)
    V = np.maximum(phi * (S - K), 0.0)
    for i in range(n - 1, -1, -1):
        V = disc * (p * V[:-1] + (1 - p) * V[1:])
        if i % steps_per_date == 0 and i > 0:  # an exercise date before maturity
            S = S0 * u ** np.arange(i, -1, -1) * d ** np.arange(0, i + 1)
            V = np.maximum(V, np.maximum(phi * (S - K), 0.0))
    return max(float(V[0]), max(phi * (S0 - K), 0.0)) if today else float(V[0])  # products.py:165-168
