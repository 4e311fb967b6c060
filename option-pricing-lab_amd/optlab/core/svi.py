"""Raw-SVI slices interpolated into a surface (the surface of packages/core/src/optlab/core/svi.py:31-160; the fit is
not part of the engine).

Pillar i holds a raw SVI slice p_i = (a, b, rho, m, sigma), w_i(k) = a + b (rho (k - m) + sqrt((k - m)^2 + sigma^2)),
and theta_i = w_i(0). theta(T) is a monotone cubic in T through (0, 0) and the pillars, linear beyond the last.
Between pillars i - 1 and i the surface is (1 - alpha) w_{i-1} + alpha w_i with alpha = (theta(T) - theta_{i-1}) /
(theta_i - theta_{i-1}) (a zero slice before the first pillar); beyond the last pillar it is w_n scaled by
theta(T) / theta_n. Evaluation goes through optlab.core.slices."""

from __future__ import annotations

import numpy as np

from optlab.core.slices import ZERO_SLICE, RowSlices, SliceSurface, TermStructure


def _raw_to_slice(p) -> np.ndarray:
    """(a, b, rho, m, sigma) -> the slice record's (a, b, rho, m, sigma^2)."""
    p = np.asarray(p, dtype=float)
    return np.concatenate([p[..., :4], p[..., 4:5] ** 2], axis=-1)


def svi_w(p, k):
    """Total variance of one raw SVI slice at k."""
    a, b, rho, m, sigma = p
    x = np.asarray(k, dtype=float) - m
    return a + b * (rho * x + np.sqrt(x * x + sigma * sigma))


class SVISurface(SliceSurface):
    def __init__(self, maturities, params):
        self.maturities = np.asarray(maturities, dtype=float)
        self.params = np.asarray(params, dtype=float)
        self.thetas = np.array([float(svi_w(p, 0.0)) for p in self.params])
        self.term = TermStructure(self.maturities, self.thetas)
        # the slices in record form, with the zero slice as slice -1 (index 0 here)
        self._slices = np.concatenate([np.asarray([ZERO_SLICE]), _raw_to_slice(self.params)])
        self._levels = np.concatenate([[0.0], self.thetas])

    @classmethod
    def from_dict(cls, d: dict) -> SVISurface:
        return cls(d["maturities"], d["params"])

    def to_dict(self) -> dict:
        return {"kind": "svi", "maturities": self.maturities.tolist(), "params": self.params.tolist()}

    def slices(self, T) -> RowSlices:
        T = np.asarray(T, dtype=float)
        theta, dtheta = self.term(T)
        n = len(self.maturities)
        upper = np.searchsorted(self.maturities, T, side="left")  # the pillar at or above T; n beyond the last
        past = upper >= n
        hi_idx = np.minimum(upper, n - 1) + 1  # in _slices / _levels (index 0 is the zero slice)
        lo_idx = hi_idx - 1
        span = self._levels[hi_idx] - self._levels[lo_idx]
        alpha = (theta - self._levels[lo_idx]) / span
        dalpha = dtheta / span
        scale, dscale = theta / self.thetas[-1], dtheta / self.thetas[-1]
        c = np.stack([np.where(past, scale, 1.0 - alpha), np.where(past, 0.0, alpha)], axis=-1)
        dc = np.stack([np.where(past, dscale, -dalpha), np.where(past, 0.0, dalpha)], axis=-1)
        first = np.where(past, n, lo_idx)  # beyond the last pillar: the last slice alone
        second = np.where(past, 0, hi_idx)
        p = np.stack([self._slices[first], self._slices[second]], axis=-2)
        return RowSlices(T, c, dc, p, np.zeros_like(p))
