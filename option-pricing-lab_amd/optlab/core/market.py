"""Deterministic discount and forward curves (packages/core/src/optlab/core/market.py:14-80).

ln D(T) and ln(F(T) / S0) are piecewise linear between pillars (a pillar at T = 0 is added) and continue with
the last segment's slope beyond the last pillar. The engine reads F(t_j) at every observation date."""

from __future__ import annotations

import numpy as np


class Market:
    def __init__(self, spot: float, pillars=(1.0,), discounts=(1.0,), forwards=None):
        self.spot = float(spot)
        t = np.asarray(pillars, dtype=float)
        d = np.asarray(discounts, dtype=float)
        f = np.full_like(t, self.spot) if forwards is None else np.asarray(forwards, dtype=float)
        order = np.argsort(t)
        self._t = np.concatenate([[0.0], t[order]])
        self._log_d = np.concatenate([[0.0], np.log(d[order])])
        self._log_f = np.concatenate([[0.0], np.log(f[order] / self.spot)])

    @classmethod
    def flat(cls, spot: float, r: float = 0.0, q: float = 0.0) -> Market:
        """Constant short rate r and dividend yield q (one pillar at 1 y, extended linearly)."""
        t = np.array([1.0])
        return cls(spot, t, np.exp(-r * t), spot * np.exp((r - q) * t))

    def _curve(self, values: np.ndarray, T):
        T = np.asarray(T, dtype=float)
        slope = (values[-1] - values[-2]) / (self._t[-1] - self._t[-2])
        inside = np.interp(T, self._t, values)
        return np.where(T > self._t[-1], values[-1] + slope * (T - self._t[-1]), inside)

    def D(self, T):
        return np.exp(self._curve(self._log_d, T))

    def F(self, T):
        return self.spot * np.exp(self._curve(self._log_f, T))

    def zero_rate(self, T):
        T = np.maximum(np.asarray(T, dtype=float), 1e-8)
        return -np.log(self.D(T)) / T

    def carry_rate(self, T):
        T = np.maximum(np.asarray(T, dtype=float), 1e-8)
        return np.log(self.F(T) / self.spot) / T

    def flat_rates(self, T: float):
        """(r, q) when the short rate and the carry are constant up to T, else None (packages/native/src/optlab/
        native/gpu.py:73-81)."""
        r, carry = float(self.zero_rate(T)), float(self.carry_rate(T))
        t = np.linspace(T / 16, T, 16)
        if np.allclose(self.zero_rate(t), r, rtol=0, atol=1e-10) and np.allclose(self.carry_rate(t), carry, rtol=0, atol=1e-10):
            return r, r - carry
        return None
