"""Extended SSVI surface and its Dupire local variance (packages/core/src/optlab/core/ssvi.py:26-115).

Slice at T: w(k) = (theta + chi k + sqrt(psi^2 k^2 + 2 chi theta k + theta^2)) / 2 with chi = rho psi. theta(T) is a
PCHIP cubic through (0, 0) and the pillars, linear beyond the last; the wing slopes a = psi (1 + rho) and
b = psi (1 - rho) are PCHIP cubics in theta, flat beyond the last pillar. All derivatives are analytic."""

from __future__ import annotations

import numpy as np
from scipy.interpolate import PchipInterpolator


class SSVI:
    def __init__(self, maturities, thetas, psis, rhos):
        self.maturities = np.asarray(maturities, dtype=float)
        self.thetas = np.asarray(thetas, dtype=float)
        self.psis = np.asarray(psis, dtype=float)
        self.rhos = np.asarray(rhos, dtype=float)
        t = np.concatenate([[0.0], self.maturities])
        th = np.concatenate([[0.0], self.thetas])
        a = np.concatenate([[0.0], self.psis * (1 + self.rhos)])
        b = np.concatenate([[0.0], self.psis * (1 - self.rhos)])
        self._theta_T = PchipInterpolator(t, th, extrapolate=False)
        self._dtheta_T = self._theta_T.derivative()
        self._a = PchipInterpolator(th, a, extrapolate=False)
        self._b = PchipInterpolator(th, b, extrapolate=False)
        self._da, self._db = self._a.derivative(), self._b.derivative()
        self._slope_last = (th[-1] - th[-2]) / (t[-1] - t[-2])

    @classmethod
    def power_law(cls, rho, eta, gamma, maturities, thetas) -> SSVI:
        """phi(theta) = eta / (theta^gamma (1 + theta)^(1 - gamma)) at the pillars (ssvi.py:44-48)."""
        th = np.asarray(thetas, dtype=float)
        psi = th * eta / (th**gamma * (1 + th) ** (1 - gamma))
        return cls(maturities, th, psi, np.full_like(th, rho))

    def theta(self, T):
        T = np.asarray(T, dtype=float)
        last = self.maturities[-1]
        inside = self._theta_T(np.clip(T, 0.0, last))
        return np.where(T > last, self.thetas[-1] + self._slope_last * (T - last), inside)

    def dtheta(self, T):
        T = np.asarray(T, dtype=float)
        inside = self._dtheta_T(np.clip(T, 0.0, self.maturities[-1]))
        return np.where(T > self.maturities[-1], self._slope_last, inside)

    def _wings(self, theta):
        th = np.clip(theta, 0.0, self.thetas[-1])
        a, b = self._a(th), self._b(th)
        beyond = theta > self.thetas[-1]
        da = np.where(beyond, 0.0, self._da(th))
        db = np.where(beyond, 0.0, self._db(th))
        return a, b, da, db

    def params(self, T):
        th = self.theta(T)
        a, b, _, _ = self._wings(th)
        return th, 0.5 * (a + b), (a - b) / (a + b)

    # w and derivatives evaluate the maturity's pieces (theta, the wings and their slopes, PCHIP) on T as given and
    # broadcast against k in the arithmetic: with T of shape (rows, 1) and k of (rows, n) that is one PCHIP
    # evaluation a row instead of one an element.
    def w(self, k, T):
        k, T = np.asarray(k, dtype=float), np.asarray(T, dtype=float)
        th = self.theta(T)
        a, b, _, _ = self._wings(th)
        psi, chi = 0.5 * (a + b), 0.5 * (a - b)
        return 0.5 * (th + chi * k + np.sqrt(psi**2 * k**2 + 2 * chi * th * k + th**2))

    def iv(self, k, T):
        return np.sqrt(self.w(k, T) / np.asarray(T, dtype=float))

    def derivatives(self, k, T):
        """w, w_k, w_kk, w_T (each of the broadcast shape of k and T)."""
        k, T = np.asarray(k, dtype=float), np.asarray(T, dtype=float)
        th = self.theta(T)
        a, b, da, db = self._wings(th)
        psi, chi = 0.5 * (a + b), 0.5 * (a - b)
        dpsi, dchi = 0.5 * (da + db), 0.5 * (da - db)
        R = np.sqrt(psi**2 * k**2 + 2 * chi * th * k + th**2)
        w = 0.5 * (th + chi * k + R)
        wk = 0.5 * (chi + (psi**2 * k + chi * th) / R)
        wkk = 0.5 * th**2 * (psi**2 - chi**2) / R**3
        dw_dtheta = 0.5 * (1 + (chi * k + th) / R)
        dw_dpsi = 0.5 * psi * k**2 / R
        dw_dchi = 0.5 * k * (1 + th / R)
        wT = (dw_dtheta + dw_dpsi * dpsi + dw_dchi * dchi) * self.dtheta(T)
        return np.broadcast_arrays(w, wk, wkk, wT)

    def local_var(self, k, T, floor=1e-4, cap=4.0):
        T = np.maximum(np.asarray(T, dtype=float), 1.0 / 365)
        w, wk, wkk, wT = self.derivatives(k, T)
        g = (1 - k * wk / (2 * w)) ** 2 - wk**2 / 4 * (1 / w + 0.25) + wkk / 2
        return np.clip(wT / np.maximum(g, 1e-6), floor, cap)

    def row_grid(self, t0, n_x, sd_range=(-8.0, 6.0)):
        """(T, lo, hi, k) of the engine's local-variance rows at step starts t0: T = max(t0, 1 day), the row spans
        sd_range ATM standard deviations sqrt(w(0, T)), k[i] the n_x points of np.linspace(lo[i], hi[i], n_x) (the
        same arithmetic, element by element, in one array: the same numbers)."""
        T = np.maximum(np.asarray(t0, dtype=float), 1.0 / 365)
        sd = np.sqrt(self.w(0.0, T))
        lo, hi = sd_range[0] * sd, sd_range[1] * sd
        k = np.arange(n_x, dtype=float)[None, :] * ((hi - lo) / (n_x - 1))[:, None] + lo[:, None]
        k[:, -1] = hi
        return T, lo, hi, k

    def local_var_rows(self, t0, n_x, sd_range=(-8.0, 6.0)):
        """(lo, inv_dx, rows) of the engine's Dupire local-variance rows (native/__init__.py:129-144) in one
        vectorised pass over every row: the grid of the per-row loop exactly, the rows equal to it up to the last
        bit of a few elements (on the 3-year daily grid: 3 rows of 1095, 1e-16)."""
        T, lo, hi, k = self.row_grid(t0, n_x, sd_range)
        return lo, (n_x - 1) / (hi - lo), self.local_var(k, T[:, None])
