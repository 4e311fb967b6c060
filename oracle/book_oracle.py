"""CPU ORACLE -- test infrastructure only. A trade of a book valued in double on the oracle's paths, which the HIP book
kernel is checked against.

    paths          the oracle's path engine (liboracle.so, oracle_paths): local vol per name, one table row a step
    tables         the surface oracle's bumped Dupire table (surface_oracle.py) moved by a parallel shift and a skew
    European, arithmetic Asian and up-and-out calls on the closes; the worst-of autocall of autocall_oracle.py with
    hard comparisons (smoothing 0)
"""

from __future__ import annotations

import numpy as np

import autocall_oracle as ao
import surface_oracle as so


def trade_value(lib, paths_request_type, trade: dict, book: dict, scen: dict):
    """(price, se) of one trade under one scenario.

    trade: product, names, root (n x n), n_steps, obs_every, strike, barrier, coupon, coupon_barrier, autocall,
           memory, n_pairs, seed; book: rows (per name, surface_oracle.base_rows on the book's grid), fwd
           (names x steps + 1), disc (steps + 1), dt; scen: spot (per name), vol, skew."""
    names = list(trade["names"])
    n = trade["n_steps"]
    rows = [book["rows"][a] for a in names]
    tables = np.stack([so.bumped_local_var(r, np.zeros((10, 8)), parallel=scen["vol"], skew=scen["skew"])[:n]
                       for r in rows])
    lo = np.stack([r[0][:n] for r in rows])
    inv = np.stack([r[1][:n] for r in rows])
    fwd = np.stack([book["fwd"][a][: n + 1] * (1 + scen["spot"][a]) for a in names])
    closes = ao.closes(lib, paths_request_type, trade["n_pairs"], trade["seed"], np.asarray(trade["root"]), tables,
                       lo, inv, fwd, book["dt"])
    spot0 = np.array([book["fwd"][a][0] for a in names])
    D = book["disc"][n]
    p = trade["product"]
    if p == "autocall":
        obs = np.arange(trade["obs_every"], n + 1, trade["obs_every"])
        v = ao.phoenix(closes, spot0, obs, book["disc"][obs], trade["coupon"], trade["coupon_barrier"],
                       trade["autocall"], trade["memory"], trade["barrier"], trade["strike"], 0.0)
    else:
        S = closes[0]
        K = trade["strike"] * spot0[0]
        if p == "european":
            v = D * np.maximum(S[n] - K, 0.0)
        elif p == "asian":
            fix = S[trade["obs_every"]:: trade["obs_every"]]
            v = D * np.maximum(fix.mean(axis=0) - K, 0.0)
        else:  # up-and-out on every close after the valuation date
            alive = S[1:].max(axis=0) < trade["barrier"] * spot0[0]
            v = np.where(alive, D * np.maximum(S[n] - K, 0.0), 0.0)
    return ao.pair_mean(v)
