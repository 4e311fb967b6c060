"""Black-Scholes closed forms used by the engine's host side and its tests: Black's formula, the implied
volatility (to state errors in vol bp), the geometric Asian on equally spaced fixings (the control variate of the
arithmetic Asian, packages/native/cuda/src/closed_forms.cpp:21-33) and the daily down-and-out call with the
Broadie-Glasserman-Kou shift (closed_forms.cpp:35-41)."""

from __future__ import annotations

import numpy as np
from scipy.special import ndtr
from scipy.optimize import brentq


def black_price(F, K, T, vol, D=1.0, cp="call"):
    F, K, vol = np.asarray(F, dtype=float), np.asarray(K, dtype=float), np.asarray(vol, dtype=float)
    sd = vol * np.sqrt(T)
    d1 = (np.log(F / K) + 0.5 * sd * sd) / sd
    d2 = d1 - sd
    if cp == "call":
        return D * (F * ndtr(d1) - K * ndtr(d2))
    return D * (K * ndtr(-d2) - F * ndtr(-d1))


def bs_price(S, K, T, r, q, sigma, call=True):
    return float(black_price(S * np.exp((r - q) * T), K, T, sigma, np.exp(-r * T), "call" if call else "put"))


def implied_vol(price, F, K, T, D=1.0, cp="call"):
    """Black implied volatility by bracketing; nan when the price is outside the no-arbitrage bounds."""
    intrinsic = D * max(F - K, 0.0) if cp == "call" else D * max(K - F, 0.0)
    upper = D * F if cp == "call" else D * K
    if not (intrinsic < price < upper):
        return float("nan")
    return brentq(lambda s: float(black_price(F, K, T, s, D, cp)) - price, 1e-6, 5.0, xtol=1e-14, maxiter=200)


def geometric_asian(S, K, T, r, q, sigma, n, call=True):
    """Geometric Asian on fixings t_i = i T / n, i = 1..n, paid at T."""
    i = np.arange(1, n + 1, dtype=float)
    mean_t = float(np.mean(i * T / n))
    mean_min = float(np.mean(np.minimum.outer(i, i) * T / n))
    mu = np.log(S) + (r - q - 0.5 * sigma * sigma) * mean_t
    sd = sigma * np.sqrt(mean_min)
    d2 = (mu - np.log(K)) / sd
    d1 = d2 + sd
    eg, D = np.exp(mu + 0.5 * sd * sd), np.exp(-r * T)
    return float(D * (eg * ndtr(d1) - K * ndtr(d2)) if call else D * (K * ndtr(-d2) - eg * ndtr(-d1)))


def down_out_call(S, K, H, T, r, q, sigma):
    """Continuously monitored down-and-out call, H < K (Reiner-Rubinstein)."""
    sd = sigma * np.sqrt(T)
    lam = (r - q + 0.5 * sigma * sigma) / (sigma * sigma)
    y = np.log(H * H / (S * K)) / sd + lam * sd
    down_in = S * np.exp(-q * T) * (H / S) ** (2 * lam) * ndtr(y) - K * np.exp(-r * T) * (H / S) ** (2 * lam - 2) * ndtr(y - sd)
    return bs_price(S, K, T, r, q, sigma, True) - float(down_in)


def down_out_call_bgk(S, K, H, T, r, q, sigma, n_monitor):
    return down_out_call(S, K, H * np.exp(-0.5826 * sigma * np.sqrt(T / n_monitor)), T, r, q, sigma)


def up_out_call(S, K, H, T, r, q, sigma):
    """Continuously monitored up-and-out call, K < H (reflection principle, Reiner-Rubinstein)."""
    if H <= K:
        return 0.0
    sd = sigma * np.sqrt(T)
    lam = (r - q + 0.5 * sigma * sigma) / (sigma * sigma)
    x1 = np.log(S / H) / sd + lam * sd
    y1 = np.log(H / S) / sd + lam * sd
    y = np.log(H * H / (S * K)) / sd + lam * sd
    dq, dr = np.exp(-q * T), np.exp(-r * T)
    up_in = (
        S * dq * ndtr(x1)
        - K * dr * ndtr(x1 - sd)
        - S * dq * (H / S) ** (2 * lam) * (ndtr(-y) - ndtr(-y1))
        + K * dr * (H / S) ** (2 * lam - 2) * (ndtr(-y + sd) - ndtr(-y1 + sd))
    )
    return bs_price(S, K, T, r, q, sigma, True) - float(up_in)


def up_out_call_bgk(S, K, H, T, r, q, sigma, n_monitor):
    """Daily monitoring by the Broadie-Glasserman-Kou shift: an up barrier moves up by exp(+0.5826 sigma sqrt(dt))."""
    return up_out_call(S, K, H * np.exp(0.5826 * sigma * np.sqrt(T / n_monitor)), T, r, q, sigma)
