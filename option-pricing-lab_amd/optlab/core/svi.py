"""Raw-SVI slices interpolated into a surface, and its Dupire local variance (packages/core/src/optlab/core/
svi.py:31-160). Only the evaluation is here (the plan layer reads it); the fit is out of the engine's scope.

Between pillars w(k, T) = (1 - alpha) w_i(k) + alpha w_{i+1}(k), alpha = (theta(T) - theta_i) / (theta_{i+1} -
theta_i), with theta(T) the ATM total variance, a PCHIP cubic in T through (0, 0) and the pillars, continued
linearly beyond the last one (where the last slice is scaled by theta(T) / theta_n)."""

from __future__ import annotations

import numpy as np
from scipy.interpolate import PchipInterpolator


def svi_w(p, k):
    a, b, rho, m, sig = p
    x = k - m
    return a + b * (rho * x + np.sqrt(x * x + sig * sig))


def svi_derivs(p, k):
    """w, dw/dk, d2w/dk2 of one raw-SVI slice."""
    a, b, rho, m, sig = p
    x = k - m
    R = np.sqrt(x * x + sig * sig)
    return a + b * (rho * x + R), b * (rho + x / R), b * sig * sig / R**3


def g_function(w, wk, wkk, k):
    """Gatheral's density factor; Dupire's local variance is w_T / g."""
    return (1 - k * wk / (2 * w)) ** 2 - wk**2 / 4 * (1 / w + 0.25) + wkk / 2


class SVISurface:
    def __init__(self, maturities, params):
        self.maturities = np.asarray(maturities, dtype=float)
        self.params = np.asarray(params, dtype=float)
        self.thetas = np.array([svi_w(p, 0.0) for p in self.params])
        t = np.concatenate([[0.0], self.maturities])
        th = np.concatenate([[0.0], self.thetas])
        self._theta = PchipInterpolator(t, th, extrapolate=False)
        self._dtheta = self._theta.derivative()
        self._slope_last = (th[-1] - th[-2]) / (t[-1] - t[-2])

    @classmethod
    def from_dict(cls, d: dict) -> SVISurface:
        return cls(d["maturities"], d["params"])

    def theta(self, T):
        T = np.asarray(T, dtype=float)
        last = self.maturities[-1]
        inside = self._theta(np.clip(T, 0.0, last))
        return np.where(T > last, self.thetas[-1] + self._slope_last * (T - last), inside)

    def dtheta(self, T):
        T = np.asarray(T, dtype=float)
        inside = self._dtheta(np.clip(T, 0.0, self.maturities[-1]))
        return np.where(T > self.maturities[-1], self._slope_last, inside)

    def _slice(self, i, k):
        if i < 0:
            z = np.zeros_like(k)
            return z, z, z
        return svi_derivs(self.params[i], k)

    def derivatives(self, k, T):
        """w, w_k, w_kk, w_T at one maturity T (a scalar: what a table row asks)."""
        k = np.asarray(k, dtype=float)
        T = float(T)
        th, dth = float(self.theta(T)), float(self.dtheta(T))
        n = len(self.maturities)
        j = int(np.clip(np.searchsorted(self.maturities, T, side="left"), 0, n))
        if j == n:
            w, wk, wkk = svi_derivs(self.params[-1], k)
            s, ds = th / self.thetas[-1], dth / self.thetas[-1]
            return s * w, s * wk, s * wkk, ds * w
        lo = self.thetas[j - 1] if j > 0 else 0.0
        hi = self.thetas[j]
        alpha, dalpha = (th - lo) / (hi - lo), dth / (hi - lo)
        w0, wk0, wkk0 = self._slice(j - 1, k)
        w1, wk1, wkk1 = self._slice(j, k)
        return (
            (1 - alpha) * w0 + alpha * w1,
            (1 - alpha) * wk0 + alpha * wk1,
            (1 - alpha) * wkk0 + alpha * wkk1,
            dalpha * (w1 - w0),
        )

    def w(self, k, T):
        return self.derivatives(k, T)[0]

    def iv(self, k, T):
        return np.sqrt(np.maximum(self.w(k, T), 0.0) / float(T))

    def local_var(self, k, T, floor=1e-4, cap=4.0):
        """Dupire local variance in total implied variance: clip(w_T / max(g, 1e-6), floor, cap), T >= 1 day."""
        T = max(float(T), 1.0 / 365)
        w, wk, wkk, wT = self.derivatives(k, T)
        g = g_function(w, wk, wkk, np.broadcast_to(k, w.shape))
        return np.clip(wT / np.maximum(g, 1e-6), floor, cap)
