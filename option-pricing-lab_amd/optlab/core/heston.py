"""Heston vanilla prices by Fourier inversion: the independent reference of the QE-M and Euler tests
(packages/core/src/optlab/core/heston.py:37-76 prices the same way).

phi(u) = E[exp(i u X_T)], X_T = ln(S_T / F_T), in the Albrecher et al. "little trap" form; the call is Lewis's
C = D (F - sqrt(F K) / pi * int_0^inf Re[exp(i u ln(F / K)) phi(u - i / 2)] / (u^2 + 1/4) du)."""

from __future__ import annotations

import numpy as np
from scipy.integrate import quad

from optlab.core.models import HestonParams


def char_function(u, T, p: HestonParams):
    u = np.asarray(u, dtype=complex)
    iu = 1j * u
    beta = p.kappa - p.rho * p.xi * iu
    d = np.sqrt(beta * beta + p.xi**2 * (iu + u * u))
    g = (beta - d) / (beta + d)
    e = np.exp(-d * T)
    C = p.kappa * p.theta / p.xi**2 * ((beta - d) * T - 2.0 * np.log((1 - g * e) / (1 - g)))
    Dv = (beta - d) / p.xi**2 * (1 - e) / (1 - g * e)
    return np.exp(C + Dv * p.v0)


def heston_call(F, K, T, p: HestonParams, D=1.0) -> float:
    k = np.log(F / K)

    def integrand(u):
        return float(np.real(np.exp(1j * u * k) * char_function(u - 0.5j, T, p)) / (u * u + 0.25))

    integral, _ = quad(integrand, 0.0, np.inf, limit=400, epsabs=1e-13, epsrel=1e-12)
    return float(D * (F - np.sqrt(F * K) / np.pi * integral))


def heston_price(F, K, T, p: HestonParams, D=1.0, cp="call") -> float:
    c = heston_call(F, K, T, p, D)
    return c if cp == "call" else c - D * (F - K)
