"""Implied total-variance surfaces written as raw SVI slices, with the Dupire local variance they imply.

The two surfaces of the library's core -- raw SVI slices interpolated in ATM total variance
(packages/core/src/optlab/core/svi.py:31-160) and extended SSVI (packages/core/src/optlab/core/ssvi.py:26-115) --
are, at any maturity T, a combination of at most two raw SVI slices:

    w(k, T) = c_0(T) S(k; p_0(T)) + c_1(T) S(k; p_1(T)),
    S(k; a, b, rho, m, s2) = a + b (rho x + sqrt(x^2 + s2)),   x = k - m   (s2 is sigma^2).

Interpolated SVI: the pillars around T with weights 1 - alpha, alpha (a zero slice before the first pillar), or the
last pillar scaled by theta(T) / theta_n beyond it; the slices themselves do not move with T. Extended SSVI: one
slice whose parameters move with T -- with theta the ATM total variance, psi and chi = rho psi from the wing slopes,
S has b = psi / 2, rho = chi / psi, m = -chi theta / psi^2, s2 = theta^2 (psi^2 - chi^2) / psi^4 and
a = theta (1 - rho^2) / 2, which expands to (theta + chi k + sqrt(psi^2 k^2 + 2 chi theta k + theta^2)) / 2.

`RowSlices` holds, per maturity, the weights, the parameters and their T-derivatives; `evaluate` gives w and its
derivatives in k (once and twice) and in T; `dupire` the local variance w_T / g with Gatheral's g. The same
per-maturity records are what the device's table builders read (one record a table row)."""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
from scipy.interpolate import PchipInterpolator

#: the zero slice (w = 0 and every derivative 0; s2 = 1 keeps sqrt(x^2 + s2) away from 0)
ZERO_SLICE = (0.0, 0.0, 0.0, 0.0, 1.0)
DAY = 1.0 / 365


class RowSlices(NamedTuple):
    """Per maturity (leading shape of T): weights c (..., 2), dc/dT, slice parameters p (..., 2, 5) as
    (a, b, rho, m, s2) and dp/dT."""

    T: np.ndarray
    c: np.ndarray
    dc: np.ndarray
    p: np.ndarray
    dp: np.ndarray

    def records(self) -> np.ndarray:
        """One row of 24 doubles per maturity: c0, c1, dc0, dc1, then p0, p1, dp0, dp1 (5 each)."""
        n = self.T.size
        return np.ascontiguousarray(np.concatenate(
            [self.c.reshape(n, 2), self.dc.reshape(n, 2), self.p.reshape(n, 10), self.dp.reshape(n, 10)], axis=1))


def _slice_terms(p, dp, k):
    """S, S_k, S_kk and S_T of one slice (p, dp: (..., 5), broadcast against k)."""
    a, b, rho, m, s2 = (p[..., i] for i in range(5))
    da, db, drho, dm, ds2 = (dp[..., i] for i in range(5))
    # The operations and their order are those of the device's evaluation (csrc/leverage.hip, slice_dupire), which
    # reproduces these numbers bit for bit.
    x = k - m
    root = np.sqrt(x * x + s2)
    lin = rho * x + root
    s_k = b * (rho + x / root)
    s_kk = b * s2 / (root * root * root)
    # S_T from the partials dS/da = 1, dS/db = rho x + root, dS/drho = b x, dS/dm = -S_k, dS/ds2 = b / (2 root)
    s_t = da + db * lin + b * drho * x - s_k * dm + 0.5 * b * ds2 / root
    return a + b * lin, s_k, s_kk, s_t


def evaluate(rows: RowSlices, k):
    """(w, w_k, w_kk, w_T) at log forward moneyness k, broadcast against the maturities' shape."""
    k = np.asarray(k, dtype=float)
    out = None
    for j in range(2):
        c, dc = rows.c[..., j], rows.dc[..., j]
        s, s_k, s_kk, s_t = _slice_terms(rows.p[..., j, :], rows.dp[..., j, :], k)
        term = (c * s, c * s_k, c * s_kk, dc * s + c * s_t)
        out = term if out is None else tuple(u + v for u, v in zip(out, term))
    return np.broadcast_arrays(*out)


def density_g(k, w, w_k, w_kk):
    """Gatheral's g: the density factor in Dupire's formula written in total variance."""
    u = 1.0 - 0.5 * k * w_k / w
    return u * u - 0.25 * w_k * w_k * (1.0 / w + 0.25) + 0.5 * w_kk


def dupire(k, w, w_k, w_kk, w_T, floor=1e-4, cap=4.0):
    """clip(w_T / max(g, 1e-6), floor, cap)."""
    return np.clip(w_T / np.maximum(density_g(k, w, w_k, w_kk), 1e-6), floor, cap)


class TermStructure:
    """theta(T): a monotone (PCHIP) cubic through (0, 0) and the pillars, continued linearly beyond the last pillar
    with the slope of the last segment; theta'(T) alongside."""

    def __init__(self, maturities, thetas):
        self.maturities = np.asarray(maturities, dtype=float)
        self.thetas = np.asarray(thetas, dtype=float)
        knots = np.concatenate([[0.0], self.maturities])
        values = np.concatenate([[0.0], self.thetas])
        self._cubic = PchipInterpolator(knots, values, extrapolate=False)
        self._slope = self._cubic.derivative()
        self.last_slope = (values[-1] - values[-2]) / (knots[-1] - knots[-2])

    def __call__(self, T):
        """(theta, theta') at T (any shape)."""
        T = np.asarray(T, dtype=float)
        end = self.maturities[-1]
        inside = np.clip(T, 0.0, end)
        past = T > end
        theta = np.where(past, self.thetas[-1] + self.last_slope * (T - end), self._cubic(inside))
        return theta, np.where(past, self.last_slope, self._slope(inside))


def row_grid(surface, t0, n_x: int, sd_range=(-8.0, 6.0)):
    """(T, lo, hi, k) of the engine's local-variance rows at step starts t0 (native/__init__.py:129-144): T = max(t0,
    1 day); a row spans sd_range ATM standard deviations sqrt(w(0, T)); k[i] is np.linspace(lo[i], hi[i], n_x)
    element for element (i dx + lo, the last point hi)."""
    T = np.maximum(np.asarray(t0, dtype=float), DAY)
    sd = np.sqrt(np.asarray(surface.w(0.0, T), dtype=float))
    lo, hi = sd_range[0] * sd, sd_range[1] * sd
    k = np.arange(n_x, dtype=float)[None, :] * ((hi - lo) / (n_x - 1))[:, None] + lo[:, None]
    k[:, -1] = hi
    return T, lo, hi, k


class SliceSurface:
    """What the two surfaces share once `slices(T)` is defined: total variance, implied vol, derivatives, Dupire,
    and the engine's local-variance rows in one vectorised pass."""

    maturities: np.ndarray
    thetas: np.ndarray
    term: TermStructure

    def slices(self, T) -> RowSlices:  # pragma: no cover - defined by each surface
        raise NotImplementedError

    def theta(self, T):
        return self.term(T)[0]

    def dtheta(self, T):
        return self.term(T)[1]

    def derivatives(self, k, T):
        """w, w_k, w_kk, w_T at (k, T), broadcast together."""
        k, T = np.asarray(k, dtype=float), np.asarray(T, dtype=float)
        return evaluate(self.slices(T), k)

    def w(self, k, T):
        return self.derivatives(k, T)[0]

    def iv(self, k, T):
        return np.sqrt(np.maximum(self.w(k, T), 0.0) / np.asarray(T, dtype=float))

    def density_g(self, k, T):
        w, w_k, w_kk, _ = self.derivatives(k, T)
        return density_g(np.broadcast_to(k, w.shape), w, w_k, w_kk)

    def local_var(self, k, T, floor=1e-4, cap=4.0):
        """Dupire local variance in total implied variance, at T >= 1 day."""
        T = np.maximum(np.asarray(T, dtype=float), DAY)
        w, w_k, w_kk, w_T = self.derivatives(k, T)
        return dupire(np.broadcast_to(k, w.shape), w, w_k, w_kk, w_T, floor, cap)

    def local_vol(self, k, T):
        return np.sqrt(self.local_var(k, T))

    def row_grid(self, t0, n_x, sd_range=(-8.0, 6.0)):
        return row_grid(self, t0, n_x, sd_range)

    def local_var_rows(self, t0, n_x, sd_range=(-8.0, 6.0)):
        """(lo, inv_dx, rows): the Dupire local variance on every row of the engine's grid, one pass."""
        T, lo, hi, k = row_grid(self, t0, n_x, sd_range)
        return lo, (n_x - 1) / (hi - lo), self.local_var(k, T[:, None])

    def row_records(self, t0, n_x, sd_range=(-8.0, 6.0)):
        """(lo, hi, inv_dx, records) of the engine's local-variance rows at step starts t0: each row's grid (the
        rows of `local_var_rows` are the Dupire variance at k = j (hi - lo) / (n_x - 1) + lo, the last point hi) and
        its slice record (`RowSlices.records`), from which the device evaluates any entry of the row."""
        T = np.maximum(np.asarray(t0, dtype=float), DAY)
        sd = np.sqrt(np.asarray(self.w(0.0, T), dtype=float))
        lo, hi = sd_range[0] * sd, sd_range[1] * sd
        return lo, hi, (n_x - 1) / (hi - lo), self.slices(T).records()
