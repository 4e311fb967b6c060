"""Local correlation for the path engine's mixer (packages/structured/src/optlab/structured/localcorr.py:37-114), and
its calibration to pseudo-quotes of correlation spreads with the basket simulated on the engine (localcorr.py:116-400).

Z_i = sqrt(w) Z_0 + sqrt(1 - w) (L_low e)_i with w = clip(w0 + level + down max(-u, 0) - up max(u, 0), 0, 1) and
u = y / (sigma sqrt(max(t, 1 month))), y = ln(mean S_i / F_i) at the start of the step. L_low factors the matrix that
gives corr back at w = w0; its factors take the signs of the basket's root so that the mixer is on common random
numbers with the constant matrix where the weight has not moved."""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np
from scipy.special import ndtr

from optlab.structured.multiasset import correlation_root, nearest_correlation

W0 = 0.3
MONTH = 1 / 12
ORIGIN = "synthetic"
STRIKES = {"cvc": (1.0, 1.1), "pvp": (1.0, 0.9)}  # where each spread trades, in fractions of spot
CP = {"cvc": "call", "pvp": "put"}
SEED = 2026
NOTE = ("Calibrated on pseudo-quotes of CvC and PvP, where the correlation skew is an assumption: no broker level was "
        "available")


def low_matrix(corr, w0: float) -> np.ndarray:
    c = np.asarray(corr, dtype=float)
    low = (c - w0) / (1.0 - w0)
    np.fill_diagonal(low, 1.0)
    return nearest_correlation(np.clip(low, -0.999, 0.999))


def aligned_factors(corr, root, w0: float) -> tuple:
    """(low factors signed like `root`, sign of the market factor, angle of the common factor)."""
    low = correlation_root(low_matrix(corr, w0)).copy()
    for j in range(low.shape[1] - 1):
        if np.dot(low[:, j], root[:, j]) < 0:
            low[:, j] = -low[:, j]
    if low[:, -1].mean() < 0:
        low[:, -1] = -low[:, -1]
    sign = 1.0 if root[:, -1].mean() >= 0 else -1.0
    return low, sign, float(np.arctan2(np.sqrt(1.0 - w0) * low[:, -1].mean(), np.sqrt(w0)))


@dataclass(frozen=True)
class LocalCorrelation:
    sigma: float
    level: float = 0.0
    down: float = 0.0
    up: float = 0.0
    w0: float = W0

    def bumped(self, slope: float) -> LocalCorrelation:
        return replace(self, down=max(self.down + slope, 0.0), up=max(self.up + slope, 0.0))


@dataclass(frozen=True)
class StochasticMix:
    """The weight as a Jacobi process (localcorr.py:87-114): dw = kappa (theta - w) dt + nu sqrt(w (1 - w)) dB,
    clipped to [0, 1], with dB = rho (market factor) + sqrt(1 - rho^2) (the weight's own driver); the step's
    normals are mixed with the weight at the step's start. In the path engine as mixer 2 (the reference engine
    refuses it); the driver is a stream of its own (tag 0x40000000 of the pair's Philox counter)."""

    theta: float = W0
    nu: float = 0.5
    rho: float = -0.5
    kappa: float = 2.0
    w0: float = W0

    def bumped(self, slope: float) -> StochasticMix:
        """A steeper correlation skew for a positive slope: rho moved down by it, within [-0.99, 0.99]."""
        return replace(self, rho=float(np.clip(self.rho - slope, -0.99, 0.99)))


# ---- pseudo-quotes: the members' marginals joined by a Gaussian copula (localcorr.py:216-400) ---------------------
@dataclass(frozen=True)
class CorrelationSkew:
    """The assumption the pseudo-quotes carry: every pair's correlation moved by an offset of the strike alone,
    up by slope_down per unit of strike below 1 and down by slope_up per unit above."""

    slope_down: float = 0.5
    slope_up: float = 0.25

    def offset(self, strike: float) -> float:
        return self.slope_down * max(1.0 - strike, 0.0) - self.slope_up * max(strike - 1.0, 0.0)


def marginal_quantile(surface, T: float, n_k: int = 801):
    """u -> X = S_T / F_T under the surface: the distribution of X at K is 1 + dC/dK of the undiscounted call in
    forward units C(K) = N(d1) - K N(d1 - s), s the total implied vol at k = ln K (Breeden-Litzenberger), on k in
    [-7, 6] ATM standard deviations, kept within [0, 1] and non-decreasing, inverted by interpolation in k."""
    sd = float(np.sqrt(surface.theta(T)))
    k = np.linspace(-7.0 * sd, 6.0 * sd, n_k)
    s = np.sqrt(np.maximum(np.asarray(surface.w(k, T), dtype=float), 1e-12))
    strikes = np.exp(k)
    d1 = 0.5 * s - k / s
    call = ndtr(d1) - strikes * ndtr(d1 - s)
    cdf = np.maximum.accumulate(np.clip(1.0 + np.gradient(call, strikes), 0.0, 1.0))
    rising = np.ones(n_k, dtype=bool)
    rising[1:] = np.diff(cdf) > 1e-14
    u_nodes, k_nodes = cdf[rising], k[rising]
    return lambda u: np.exp(np.interp(u, u_nodes, k_nodes))


def skewed_matrix(matrix, strike: float, skew: CorrelationSkew = CorrelationSkew()) -> np.ndarray:
    """The correlation a product around `strike` sees: every pair moved by the skew's offset, kept in [-0.95, 0.99]
    and repaired to a correlation matrix."""
    m = np.clip(np.asarray(matrix, dtype=float) + skew.offset(strike), -0.95, 0.99)
    np.fill_diagonal(m, 1.0)
    return nearest_correlation(m)


def copula_spots(quantiles, forwards, matrix, n_pairs: int, seed: int = SEED) -> np.ndarray:
    """(2 n_pairs, n) antithetic pairs of S_i(T) / S_i(0): Gaussian draws correlated by `matrix`, mapped through
    each member's quantile and scaled by its forward over spot (rows p and n_pairs + p are a pair)."""
    z = np.random.default_rng(seed).standard_normal((n_pairs, len(quantiles)))
    u = ndtr(np.vstack([z, -z]) @ correlation_root(np.asarray(matrix, dtype=float)).T)
    return np.stack([f * q(u[:, i]) for i, (f, q) in enumerate(zip(forwards, quantiles))], axis=1)


def spread_payoff(kind: str, strike: float, x) -> np.ndarray:
    """The option on the equally weighted basket less the average of the same option on every member, per path
    (x: (paths, members) of S_T / S_0). At one strike the CvC and the PvP pay the same on every path."""
    if kind not in CP:
        raise ValueError(f"unknown spread {kind!r}, expected cvc or pvp")
    x = np.asarray(x, dtype=float)
    sign = 1.0 if kind == "cvc" else -1.0
    basket = np.maximum(sign * (x.mean(axis=1) - strike), 0.0)
    return basket - np.maximum(sign * (x - strike), 0.0).mean(axis=1)


def member_vols(surfaces, forwards, strike: float, T: float) -> np.ndarray:
    """Each member's implied vol at the structure's one strike (in its own log forward moneyness)."""
    k = np.log(strike / np.asarray(forwards, dtype=float))
    return np.array([float(np.sqrt(np.asarray(sf.w(np.array([kk]), T), dtype=float)[0] / T))
                     for sf, kk in zip(surfaces, k)])


def _shares(forwards, vols):
    f = np.asarray(forwards, dtype=float)
    return f / f.sum() * np.asarray(vols, dtype=float)


def average_correlation(matrix, forwards, vols) -> float:
    """The basket variance's one correlation: the off-diagonal pairs weighted by a_i a_j, a_i = share x vol."""
    a = _shares(forwards, vols)
    wts = np.outer(a, a)
    np.fill_diagonal(wts, 0.0)
    return float(np.sum(wts * np.asarray(matrix, dtype=float)) / wts.sum())


def read_correlation(basket_price: float, kind: str, strike: float, T: float, forwards, vols) -> tuple:
    """(correlation, basket vol) read from an undiscounted basket option taken as lognormal: its implied vol s_B,
    then rho with s_B^2 = sum a_i^2 s_i^2 + rho ((sum a_i s_i)^2 - sum a_i^2 s_i^2)."""
    from optlab.core.blackscholes import implied_vol

    f = np.asarray(forwards, dtype=float)
    s_b = float(implied_vol(basket_price, f.mean(), strike, T, 1.0, CP[kind]))
    a = _shares(f, vols)
    own = float(np.sum(a * a))
    return (s_b * s_b - own) / (float(a.sum()) ** 2 - own), s_b


def spread_quote(surfaces, forwards, x, matrix, kind: str, strike: float, T: float) -> dict:
    """One pseudo-quote in % of notional, undiscounted: the spread from the copula paths (pair averages), the
    members in closed form at their own vols, the basket as their sum."""
    from optlab.core.blackscholes import black_price

    h = len(x) // 2
    pay = spread_payoff(kind, strike, x)
    pay = 0.5 * (pay[:h] + pay[h:])
    vols = member_vols(surfaces, forwards, strike, T)
    f = np.asarray(forwards, dtype=float)
    members = float(np.mean([black_price(fi, strike, T, vi, 1.0, CP[kind]) for fi, vi in zip(f, vols)]))
    spread = float(pay.mean())
    implied, s_b = read_correlation(members + spread, kind, strike, T, f, vols)
    return {"kind": kind, "strike": strike, "tenor": T, "price": 100 * spread,
            "price_se": 100 * float(pay.std(ddof=1)) / np.sqrt(h), "basket_price": 100 * (members + spread),
            "members_price": 100 * members, "basket_vol": s_b, "rho": average_correlation(matrix, f, vols),
            "implied_rho": implied}


def basket_rows(day, names, surfaces, forward, base_matrix, tenors=(0.25, 0.5, 1.0), strikes=STRIKES,
                n_pairs: int = 100_000, skew: CorrelationSkew = CorrelationSkew(), seed: int = SEED) -> list:
    """Every pseudo-quote of one basket on one day (`forward(T)`: each member's forward over spot). One copula
    simulation a (tenor, strike), shared by the CvC and the PvP of that strike."""
    label = "+".join(names)
    out = []
    for T in tenors:
        f = np.asarray(forward(T), dtype=float)
        quantiles = [marginal_quantile(sf, T) for sf in surfaces]
        for strike in sorted({k for ks in strikes.values() for k in ks}):
            matrix = skewed_matrix(base_matrix, strike, skew)
            x = copula_spots(quantiles, f, matrix, n_pairs, seed)
            out += [{"basket": label, "obs_date": day, **spread_quote(surfaces, f, x, matrix, kind, strike, T)}
                    for kind, ks in strikes.items() if strike in ks]
    return out


def targets(surfaces, markets, corr, T: float = 1.0, skew=None, names=None) -> list:
    """The three lines a tenor is calibrated on (localcorr.py:130-138): PvP 90 %, CvC 100 % and 110 %."""
    names = tuple(names or (f"n{i}" for i in range(len(surfaces))))

    def forward(t):
        return [float(m.F(t)) / float(m.spot) for m in markets]

    rows = basket_rows(None, list(names), surfaces, forward, corr, tenors=(T,), skew=skew or CorrelationSkew())
    keep = [r for r in rows if (r["kind"], r["strike"]) in (("pvp", 0.9), ("cvc", 1.0), ("cvc", 1.1))]
    return sorted(keep, key=lambda r: r["strike"])


def basket_sigma(surfaces, corr, T: float = 1.0) -> float:
    """ATM vol of the equally weighted basket from the members' ATM vols and the correlation (localcorr.py:122-127)."""
    vols = np.array([float(np.sqrt(sf.theta(T) / T)) for sf in surfaces])
    return float(np.sqrt(vols @ np.asarray(corr, dtype=float) @ vols) / len(vols))


# ---- the fit on the engine (localcorr.py:141-213) -----------------------------------------------------------------
def basket_model(surfaces, markets, corr, mixer, names=None):
    """Local vol for every member, the measured correlation, the mixer."""
    from optlab.core.models import LocalVolModel
    from optlab.structured.multiasset import Basket

    names = tuple(names or (f"n{i}" for i in range(len(surfaces))))
    return Basket(names, {n: LocalVolModel(sf) for n, sf in zip(names, surfaces)}, dict(zip(names, markets)), corr,
                  mixer=mixer)


def quotes_from_spots(x, wanted) -> tuple:
    """(prices in %, standard errors) of the lines `wanted` on spots x (paths, members) of S_T / S_0."""
    h = len(x) // 2
    prices, errs = [], []
    for r in wanted:
        pay = 100.0 * spread_payoff(r["kind"], r["strike"], x)
        pay = 0.5 * (pay[:h] + pay[h:])
        prices.append(float(pay.mean()))
        errs.append(float(pay.std(ddof=1) / np.sqrt(h)))
    return np.array(prices), np.array(errs)


def model_quotes(basket, wanted, T: float, n_paths: int, seed: int, backend: str = "auto") -> tuple:
    """The lines under the basket's mixer, the basket simulated by the path engine to T on steps of at most 1/50."""
    from optlab import native

    paths = native.simulate(basket, None, [T], n_paths, seed, dt_max=1 / 50, backend=backend)
    x = np.stack([paths[n][:, 0] / float(basket.markets[n].spot) for n in basket.names], axis=1)
    return quotes_from_spots(x, wanted)


def gauss_newton(quotes, goal, start, bounds, steps: int = 6, h: float = 0.02) -> tuple:
    """(best parameters, their RMS error) of a Gauss-Newton fit of quotes(p) -> (prices, standard errors) to goal,
    every evaluation on the same random numbers: a forward difference of h per parameter (backwards where that
    would leave the bounds), a least-squares step (singular values below 1e-3 of the largest dropped) damped by
    0.8 and clipped to the bounds; it stops once the RMS error is below the prices' mean standard error."""
    goal = np.asarray(goal, dtype=float)
    p = np.array(start, dtype=float)
    lo, hi = np.asarray(bounds[0], dtype=float), np.asarray(bounds[1], dtype=float)
    best, best_err = p.copy(), np.inf
    for _ in range(steps):
        f, noise = quotes(p)
        resid = f - goal
        err = float(np.sqrt(np.mean(resid * resid)))
        if err < best_err:
            best, best_err = p.copy(), err
        if err < float(noise.mean()):
            break
        jac = np.empty((goal.size, p.size))
        for j in range(p.size):
            q = p.copy()
            q[j] += h if p[j] + h <= hi[j] else -h
            jac[:, j] = (quotes(q)[0] - f) / (q[j] - p[j])
        p = np.clip(p + 0.8 * np.linalg.lstsq(jac, -resid, rcond=1e-3)[0], lo, hi)
    return best, best_err


FORMS = {
    # (build the mixer from the parameters, start, (lower, upper) bounds)
    "local": (lambda sigma, p: LocalCorrelation(sigma, p[0], p[1], p[2]), (0.0, 0.1, 0.1),
              ((-W0, 0.0, 0.0), (1.0 - W0, 1.5, 1.5))),
    "stochastic": (lambda sigma, p: StochasticMix(p[0], p[1], p[2]), (W0, 0.5, -0.5),
                   ((0.01, 0.05, -0.95), (0.99, 3.0, 0.95))),
}


def calibrate(surfaces, markets, corr, T: float = 1.0, n_paths: int = 40_000, seed: int = SEED, form: str = "local",
              skew=None, names=None, wanted=None, backend: str = "auto") -> dict:
    """The mixer fitted to a tenor's pseudo-quotes (localcorr.py:153-213), every model price simulated on the
    engine: {"mixer", "lines", "rmse", "origin", "note", "form"}; the lines are judged on other paths (seed + 1).
    The reference reads surfaces, markets and the correlation from a basket context; here they are arguments.
    `wanted` replaces the pseudo-quotes by lines of one's own (dicts with kind, strike, price, rho)."""
    if form not in FORMS:
        raise ValueError(f"unknown form {form!r}")
    wanted = wanted if wanted is not None else targets(surfaces, markets, corr, T, skew, names)
    sigma = basket_sigma(surfaces, corr, T)
    build, start, bounds = FORMS[form]

    def quotes(p, s=seed):
        return model_quotes(basket_model(surfaces, markets, corr, build(sigma, p), names), wanted, T, n_paths, s,
                            backend)

    p, rmse = gauss_newton(quotes, [r["price"] for r in wanted], start, bounds)
    mixer = build(sigma, p)
    fitted, se = quotes(p, seed + 1)
    lines = [{"kind": r["kind"], "strike": r["strike"], "tenor": T, "pseudo_quote": r["price"], "model": float(f),
              "se": float(e), "assumed_rho": r.get("rho")} for r, f, e in zip(wanted, fitted, se)]
    return {"mixer": mixer, "lines": lines, "rmse": rmse, "origin": ORIGIN, "note": NOTE, "form": form}
