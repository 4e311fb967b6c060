"""Extended SSVI surface (the surface of packages/core/src/optlab/core/ssvi.py:26-115; the fit is not part of the
engine).

Pillars carry theta (ATM total variance), psi and rho. theta(T) is a monotone cubic through (0, 0) and the pillars,
linear beyond the last; the wing slopes A = psi (1 + rho) and B = psi (1 - rho) are monotone cubics in theta
through (0, 0), held flat beyond the last pillar. At T the slice has psi = (A + B) / 2 and chi = rho psi = (A - B) / 2,
and it is evaluated as the raw SVI slice with the same total variance (optlab.core.slices), its parameters moving
with T through theta, A and B."""

from __future__ import annotations

import numpy as np
from scipy.interpolate import PchipInterpolator

from optlab.core.slices import RowSlices, SliceSurface, TermStructure


class SSVI(SliceSurface):
    def __init__(self, maturities, thetas, psis, rhos):
        self.maturities = np.asarray(maturities, dtype=float)
        self.thetas = np.asarray(thetas, dtype=float)
        self.psis = np.asarray(psis, dtype=float)
        self.rhos = np.asarray(rhos, dtype=float)
        self.term = TermStructure(self.maturities, self.thetas)
        level = np.concatenate([[0.0], self.thetas])
        self._wing_up = PchipInterpolator(level, np.concatenate([[0.0], self.psis * (1 + self.rhos)]),
                                          extrapolate=False)
        self._wing_down = PchipInterpolator(level, np.concatenate([[0.0], self.psis * (1 - self.rhos)]),
                                            extrapolate=False)

    @classmethod
    def power_law(cls, rho, eta, gamma, maturities, thetas) -> SSVI:
        """Plain SSVI at the pillars: psi = theta phi(theta), phi(theta) = eta theta^-gamma (1 + theta)^(gamma - 1)."""
        th = np.asarray(thetas, dtype=float)
        return cls(maturities, th, eta * th ** (1.0 - gamma) * (1.0 + th) ** (gamma - 1.0), np.full_like(th, rho))

    def _shape(self, T):
        """theta, theta', psi, chi and their T-derivatives at T."""
        theta, dtheta = self.term(T)
        top = self.thetas[-1]
        level = np.clip(theta, 0.0, top)
        moving = theta <= top  # beyond the last pillar the wings stop moving
        up, down = self._wing_up(level), self._wing_down(level)
        dup = np.where(moving, self._wing_up(level, 1), 0.0) * dtheta
        ddown = np.where(moving, self._wing_down(level, 1), 0.0) * dtheta
        return theta, dtheta, 0.5 * (up + down), 0.5 * (up - down), 0.5 * (dup + ddown), 0.5 * (dup - ddown)

    def params(self, T):
        """(theta, psi, rho) of the slice at T."""
        theta, _, psi, chi, _, _ = self._shape(T)
        return theta, psi, chi / psi

    def slices(self, T) -> RowSlices:
        T = np.asarray(T, dtype=float)
        theta, dtheta, psi, chi, dpsi, dchi = self._shape(T)
        live = psi > 0.0  # psi = 0 only at theta = 0 (T = 0), where w = 0
        ps = np.where(live, psi, 1.0)
        rho = chi / ps
        drho = (dchi * ps - chi * dpsi) / ps**2
        one = 1.0 - rho * rho
        b, db = 0.5 * ps, 0.5 * dpsi
        a = 0.5 * theta * one
        da = 0.5 * dtheta * one - theta * rho * drho
        m = -chi * theta / ps**2
        dm = -(dchi * theta + chi * dtheta) / ps**2 + 2.0 * chi * theta * dpsi / ps**3
        s2 = (theta / ps) ** 2 * one
        ds2 = 2.0 * theta * dtheta * one / ps**2 - 2.0 * theta**2 * rho * drho / ps**2 - 2.0 * theta**2 * one * dpsi / ps**3
        slice0 = np.stack([a, b, rho, m, s2], axis=-1)
        dslice0 = np.stack([da, db, drho, dm, ds2], axis=-1)
        dead = ~live[..., None]
        slice0 = np.where(dead, np.array([0.0, 0.0, 0.0, 0.0, 1.0]), slice0)
        dslice0 = np.where(dead, 0.0, dslice0)
        zero = np.broadcast_to(np.array([0.0, 0.0, 0.0, 0.0, 1.0]), slice0.shape)
        c = np.stack([np.ones_like(T), np.zeros_like(T)], axis=-1)
        return RowSlices(T, c, np.zeros_like(c), np.stack([slice0, zero], axis=-2),
                         np.stack([dslice0, np.zeros_like(dslice0)], axis=-2))

    def to_dict(self) -> dict:
        return {"maturities": self.maturities.tolist(), "thetas": self.thetas.tolist(), "psis": self.psis.tolist(),
                "rhos": self.rhos.tolist()}

    @classmethod
    def from_dict(cls, d: dict) -> SSVI:
        return cls(d["maturities"], d["thetas"], d["psis"], d["rhos"])
