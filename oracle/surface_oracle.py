"""CPU ORACLE -- test infrastructure only. NumPy restatement of the risk grid's tents and of the tent-bumped
surface's Dupire local variance, which the HIP table builder (csrc/surface.hip) computes on the device.

    strike tents: Gaussian-smoothed hats in log forward moneyness ... packages/risk/src/optlab/risk/grid.py:80-140
    tenor tents: smootherstep partition of unity ..................... grid.py:36-66, 143-166
    sigma'(k, T) = sigma(k, T) + sum_ij H_ij phi_i(k, T) psi_j(T) ..... packages/risk/src/optlab/risk/surfaces.py:1-16
    + parallel + skew rotation -skew c sqrt(1/max(T, 1/12)) tanh(k/0.3)  packages/structured/src/optlab/structured/surfaces.py:10-50
    w = T sigma'^2 with exact w_k, w_kk, w_T; Dupire on the base rows . surfaces.py:44-52, 158-177, 180-205
"""

from __future__ import annotations

import numpy as np
from scipy.special import ndtr

STRIKES = (0.6, 0.7, 0.8, 0.9, 0.95, 1.0, 1.05, 1.1, 1.2, 1.4)
TENORS = (1 / 12, 0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 5.0)
LOG_STRIKES = tuple(float(np.log(m)) for m in STRIKES)
ATM, WING, SCALE, T_FLOOR = 0.04, 0.12, 0.15, 1 / 52


def smootherstep(u):
    u = np.clip(u, 0.0, 1.0)
    return u**3 * (10.0 - 15.0 * u + 6.0 * u * u), 30.0 * u * u * (1.0 - u) ** 2


def tenor_tents(T):
    """(psi, dpsi/dT), each (n_tenors, *T.shape): flat at one beyond the end tenors."""
    T = np.asarray(T, dtype=float)
    nodes = np.asarray(TENORS)
    n = len(nodes)
    psi, dpsi = np.zeros((n, *T.shape)), np.zeros((n, *T.shape))
    psi[0] = np.where(T <= nodes[0], 1.0, 0.0)
    psi[-1] = np.where(T >= nodes[-1], 1.0, 0.0)
    for i in range(n - 1):
        a, b = nodes[i], nodes[i + 1]
        inside = (T > a) & (T < b)
        s, ds = smootherstep((T - a) / (b - a))
        psi[i] += np.where(inside, 1.0 - s, 0.0)
        psi[i + 1] += np.where(inside, s, 0.0)
        dpsi[i] += np.where(inside, -ds / (b - a), 0.0)
        dpsi[i + 1] += np.where(inside, ds / (b - a), 0.0)
    for i in range(1, n - 1):
        psi[i] = np.where(T == nodes[i], 1.0, psi[i])
    return psi, dpsi


def ramps():
    """Each strike hat as const + sum_m c_m (k - a_m)_+ (grid.py:80-97)."""
    x = np.asarray(LOG_STRIKES)
    n = len(x)
    out = []
    for i in range(n):
        if i == 0:
            d = x[1] - x[0]
            out.append((1.0, [(x[0], -1 / d), (x[1], 1 / d)]))
        elif i == n - 1:
            d = x[-1] - x[-2]
            out.append((0.0, [(x[-2], 1 / d), (x[-1], -1 / d)]))
        else:
            dl, dr = x[i] - x[i - 1], x[i + 1] - x[i]
            out.append((0.0, [(x[i - 1], 1 / dl), (x[i], -1 / dl - 1 / dr), (x[i + 1], 1 / dr)]))
    return out


def width(k, T):
    """l = c(k) sqrt(max(T, floor)) and (l_k, l_kk, l_T)."""
    root = np.sqrt(np.maximum(T, T_FLOOR))
    span, a2 = WING - ATM, SCALE**2
    e = np.exp(-k * k / a2)
    c = WING - span * e
    ck = span * e * 2 * k / a2
    ckk = span * e * (2 / a2 - 4 * k * k / a2**2)
    return c * root, ck * root, ckk * root, np.where(T < T_FLOOR, 0.0, 0.5 * c / root)


def strike_tents(k, T):
    """(phi, phi_k, phi_kk, phi_T), each (n_strikes, *shape)."""
    k, T = np.broadcast_arrays(np.asarray(k, dtype=float), np.asarray(T, dtype=float))
    ell, lk, lkk, lT = width(k, T)
    out = np.zeros((4, len(STRIKES), *k.shape))
    for i, (const, rs) in enumerate(ramps()):
        out[0, i] = const
        for a, c in rs:
            u = k - a
            z = u / ell
            pdf = np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)
            cdf = ndtr(z)
            out[0, i] += c * (u * cdf + ell * pdf)
            out[1, i] += c * (cdf + pdf * lk)
            out[2, i] += c * (pdf * ((1 - z * lk) ** 2 / ell + lkk))
            out[3, i] += c * (pdf * lT)
    return out


def base_rows(surface, t0, n_x=801, sd_range=(-8.0, 6.0)):
    """(lo, inv_dx, k, T, w, w_k, w_kk, w_T) of the base surface on the engine's table grid (surfaces.py:180-205)."""
    n = len(t0)
    lo, inv, k = np.empty(n), np.empty(n), np.empty((n, n_x))
    T = np.maximum(np.asarray(t0, dtype=float), 1.0 / 365)
    out = [np.empty((n, n_x)) for _ in range(4)]
    for i, t in enumerate(T):
        sd = float(np.sqrt(surface.w(0.0, float(t))))
        a, b = sd_range[0] * sd, sd_range[1] * sd
        lo[i], inv[i] = a, (n_x - 1) / (b - a)
        k[i] = np.linspace(a, b, n_x)
        for o, v in zip(out, surface.derivatives(k[i], float(t))):
            o[i] = v
    return lo, inv, k, T, *out


SKEW_WIDTH, SKEW_T_FLOOR = 0.3, 1 / 12
SKEW_NORM = 1.0 / (np.tanh(-np.log(0.9) / SKEW_WIDTH) + np.tanh(np.log(1.1) / SKEW_WIDTH))


def skew_bump(skew, k, T):
    """The skew rotation and its derivatives in k, k twice and T."""
    if not skew:
        z = np.zeros(np.broadcast(k, T).shape)
        return z, z, z, z
    held = T < SKEW_T_FLOOR
    scale = np.sqrt(1.0 / np.maximum(T, SKEW_T_FLOOR))
    dscale = np.where(held, 0.0, -0.5 * scale / np.maximum(T, SKEW_T_FLOOR))
    a = -skew * SKEW_NORM
    th = np.tanh(k / SKEW_WIDTH)
    sech2 = 1 - th * th
    return a * scale * th, a * scale * sech2 / SKEW_WIDTH, -2 * a * scale * th * sech2 / SKEW_WIDTH**2, a * dscale * th


def bumped_local_var(rows, H, floor=1e-4, cap=4.0, parallel=0.0, skew=0.0):
    """The local variance table of the base rows bumped by H (n_strikes x n_tenors, in vol units), a parallel shift
    and a skew rotation."""
    lo, inv, k, T, w, wk, wkk, wT = rows
    Tm = np.broadcast_to(T[:, None], k.shape)
    s = np.sqrt(np.maximum(w, 1e-300) / Tm)
    sk = wk / (2 * Tm * s)
    skk = (wkk - 2 * Tm * sk**2) / (2 * Tm * s)
    sT = (wT - s**2) / (2 * Tm * s)
    phi = strike_tents(k, Tm)
    psi, dpsi = tenor_tents(Tm)
    a = np.einsum("ij,i...,j...->...", H, phi[0], psi)
    ak = np.einsum("ij,i...,j...->...", H, phi[1], psi)
    akk = np.einsum("ij,i...,j...->...", H, phi[2], psi)
    aT = np.einsum("ij,i...,j...->...", H, phi[0], dpsi) + np.einsum("ij,i...,j...->...", H, phi[3], psi)
    b, bk, bkk, bT = skew_bump(skew, k, Tm)
    a, ak, akk, aT = a + parallel + b, ak + bk, akk + bkk, aT + bT
    s2, sk2, skk2, sT2 = s + a, sk + ak, skk + akk, sT + aT
    w2, wk2 = Tm * s2**2, 2 * Tm * s2 * sk2
    wkk2, wT2 = 2 * Tm * (sk2**2 + s2 * skk2), s2**2 + 2 * Tm * s2 * sT2
    g = (1 - k * wk2 / (2 * w2)) ** 2 - wk2**2 / 4 * (1 / w2 + 0.25) + wkk2 / 2
    return np.clip(wT2 / np.maximum(g, 1e-6), floor, cap)
