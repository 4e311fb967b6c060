"""The models the path engine simulates and the time grid it steps on (packages/core/src/optlab/core/models.py).

Every name is simulated in x = ln(S_t / F(t)) by the log-Euler step x <- x - a dt / 2 + sqrt(a dt) Z with the
instantaneous variance a(t, x, v) read at the start of the step: a constant (Black-Scholes), the Dupire local
variance of a surface, max(v, 0) (Heston, full truncation) or L(t, x)^2 max(v, 0) (SLV)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HestonParams:
    v0: float
    kappa: float
    theta: float
    xi: float
    rho: float

    def as_array(self) -> np.ndarray:
        return np.array([self.v0, self.kappa, self.theta, self.xi, self.rho])


class Model:
    has_variance = False


@dataclass
class BlackScholesModel(Model):
    vol: float


@dataclass
class HestonModel(Model):
    params: HestonParams
    has_variance = True

    @property
    def heston(self) -> HestonParams:
        return self.params


@dataclass
class LocalVolModel(Model):
    surface: object  # anything with local_var(k, T) and w(k, T): SVISurface, SSVI, a bumped view


class Leverage:
    """L(t, x) on time rows: piecewise constant in t, linear in x with flat ends (models.py:75-95)."""

    def __init__(self, times, x_grids, values):
        self.times = np.asarray(times, dtype=float)
        self.x_grids = np.asarray(x_grids, dtype=float)
        self.values = np.asarray(values, dtype=float)

    def __call__(self, t, x):
        i = int(np.clip(np.searchsorted(self.times, t + 1e-12, side="right") - 1, 0, len(self.times) - 1))
        return np.interp(x, self.x_grids[i], self.values[i])


@dataclass
class SLVModel(Model):
    """Heston variance with vol-of-vol scaled by `mix`, spot variance L(t, x)^2 max(v, 0)."""

    params: HestonParams
    leverage: Leverage
    mix: float = 1.0
    has_variance = True

    @property
    def heston(self) -> HestonParams:
        p = self.params
        return HestonParams(p.v0, p.kappa, p.theta, self.mix * p.xi, p.rho)


def time_grid(times, dt_max, dt_min=None) -> np.ndarray:
    """Steps no longer than dt_max through every observation time (models.py:117-135). With dt_min the mesh is
    graded, h = min(dt_max, dt_min + t / 60), and a step that would leave less than half a step to the next
    observation lands on it; without, each interval is cut into equal steps."""
    grid = [0.0]
    for t in np.asarray(times, dtype=float):
        if dt_min is not None:
            while t - grid[-1] > 1e-12:
                h = min(dt_max, dt_min + grid[-1] / 60)
                grid.append(t if t - grid[-1] < 1.5 * h else grid[-1] + h)
            continue
        n = max(1, int(np.ceil((t - grid[-1]) / dt_max - 1e-9)))
        grid.extend(np.linspace(grid[-1], t, n + 1)[1:])
    return np.array(grid)
