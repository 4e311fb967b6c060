"""Local correlation for the path engine's mixer (packages/structured/src/optlab/structured/localcorr.py:37-114), and
its calibration to pseudo-quotes of correlation spreads with every basket simulated by the engine (localcorr.py:
116-400).

The mixer. A step's spot normals are Z_i = sqrt(w) Z_0 + sqrt(1 - w) (L_low e)_i, Z_0 and e independent standard
normals, so every Z_i stays a standard normal whatever w is (each name keeps its own law) and the step's correlation
is w + (1 - w) C_low off the diagonal. C_low is chosen so that w = w0 gives the measured matrix C back. With
`LocalCorrelation` the weight reads the basket's level at the step's start,

    w = clip(w0 + level + down max(-u, 0) - up max(u, 0), 0, 1),   u = y / (sigma sqrt(max(t, 1 month))),

y = ln(mean_i S_i / F_i); with `StochasticMix` it is a Jacobi process of its own.

The calibration. Correlation spreads (CvC, PvP: an option on the equally weighted basket less the average of the
same option on the members) are priced off the members' surfaces joined by a Gaussian copula whose correlation moves
with the strike (an assumed skew, since no broker level exists); the mixer's parameters are then fitted by
Gauss-Newton to three such lines with the basket simulated on the engine at every evaluation."""

from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Callable, NamedTuple

import numpy as np
from scipy.special import ndtr

from optlab.structured.multiasset import correlation_root, nearest_correlation

W0 = 0.3  # the weight at the money
MONTH = 1 / 12
ORIGIN = "synthetic"
STRIKES = {"cvc": (1.0, 1.1), "pvp": (1.0, 0.9)}  # where each spread trades, in fractions of spot
CP = {"cvc": "call", "pvp": "put"}
SEED = 2026
NOTE = ("Calibrated on pseudo-quotes of CvC and PvP, where the correlation skew is an assumption: no broker level was "
        "available")
LINES = (("pvp", 0.9), ("cvc", 1.0), ("cvc", 1.1))  # what one tenor is fitted to


# ---- the mixer's factors -----------------------------------------------------------------------------------------
def low_matrix(corr, w0: float) -> np.ndarray:
    """C_low with w0 + (1 - w0) C_low = C off the diagonal: the pairs' excess over w0 rescaled, kept within
    (-0.999, 0.999), then repaired to the nearest correlation matrix where taking w0 out left it indefinite."""
    c = np.asarray(corr, dtype=float)
    n = len(c)
    pairs = ~np.eye(n, dtype=bool)
    low = np.eye(n)
    low[pairs] = np.clip((c[pairs] - w0) / (1.0 - w0), -0.999, 0.999)
    return nearest_correlation(low)


def aligned_factors(corr, root, w0: float) -> tuple:
    """(L_low, s, a) for the engine's mixer: a factor L_low of C_low (L_low L_low^T = C_low) whose columns draw on
    the same raw normals as the basket root's, so that the mixer runs on common random numbers with the constant
    matrix. Its last column is C_low's dominant eigenvector (scaled), signed to load positively on average; the
    others are C_low's remaining eigen-columns turned by the orthogonal Q that brings them closest to the root's
    remaining columns in the Frobenius norm (orthogonal Procrustes: Q = U V^T from the SVD U S V^T of A^T R). Q
    leaves L_low L_low^T unchanged, so no law depends on it; it settles what an eigen-solver leaves arbitrary where
    eigenvalues repeat (an equicorrelated basket). s is the sign that makes the root's own last column, the market
    factor M, load positively. With N a second independent normal the mixer takes Z_0 = cos a M + sin a N and uses
    e_last = sin a M - cos a N: independent normals for any a, and a = atan2(sqrt(1 - w0) mean(L_low[:, last]),
    sqrt(w0)) makes equal loadings on the top factor give the spots at w = w0 on M alone, as the constant matrix's.
    A price under the mixer then moves away from the constant-correlation price on the same draws only where the
    weight has moved."""
    low = correlation_root(low_matrix(corr, w0))
    root = np.asarray(root, dtype=float)
    top = low[:, -1] if low[:, -1].mean() >= 0.0 else -low[:, -1]
    rest = low[:, :-1]
    if rest.shape[1]:
        u, _, vt = np.linalg.svd(rest.T @ root[:, :-1])
        rest = rest @ (u @ vt)
    low = np.column_stack([rest, top])
    market_sign = -1.0 if root[:, -1].mean() < 0.0 else 1.0
    angle = np.arctan2(np.sqrt(1.0 - w0) * low[:, -1].mean(), np.sqrt(w0))
    return low, market_sign, float(angle)


@dataclass(frozen=True)
class LocalCorrelation:
    """The weight as a function of the basket's level in standard deviations (sigma: the basket's ATM vol)."""

    sigma: float
    level: float = 0.0
    down: float = 0.0
    up: float = 0.0
    w0: float = W0

    def bumped(self, slope: float) -> LocalCorrelation:
        """Both sides' slopes raised by `slope` (each kept non-negative): a steeper skew for slope > 0."""
        return replace(self, down=max(0.0, self.down + slope), up=max(0.0, self.up + slope))


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
        return replace(self, rho=float(min(0.99, max(-0.99, self.rho - slope))))


# ---- pseudo-quotes: the members' marginals joined by a Gaussian copula (localcorr.py:216-400) ---------------------
@dataclass(frozen=True)
class CorrelationSkew:
    """The assumption the pseudo-quotes carry: every pair's correlation moved by an offset of the strike alone,
    up by slope_down per unit of strike below 1 and down by slope_up per unit above."""

    slope_down: float = 0.5
    slope_up: float = 0.25

    def offset(self, strike: float) -> float:
        below, above = max(0.0, 1.0 - strike), max(0.0, strike - 1.0)
        return self.slope_down * below - self.slope_up * above


def _w_and_slope(surface, k: np.ndarray, T: float) -> tuple:
    """Total implied variance and its k-derivative on k: the surface's own derivative where it has one, else a
    central difference of w."""
    if hasattr(surface, "derivatives"):
        w, w_k = surface.derivatives(k, T)[:2]
        return np.asarray(w, dtype=float), np.asarray(w_k, dtype=float)
    step = 1e-5
    w = np.asarray(surface.w(k, T), dtype=float)
    w_k = (np.asarray(surface.w(k + step, T), dtype=float) - np.asarray(surface.w(k - step, T), dtype=float))
    return w, w_k / (2.0 * step)


def marginal_quantile(surface, T: float, n_k: int = 801):
    """u -> X = S_T / F_T under the surface (Breeden-Litzenberger). In forward units the undiscounted call is
    C(k) = N(d1) - e^k N(d2), d1 = -k / s + s / 2, d2 = d1 - s, s = sqrt(w(k)), and n(d1) = e^k n(d2) gives
    P(X <= e^k) = 1 + dC/dK = N(-d2) + n(d2) w_k / (2 s) in closed form. It is taken on k in [-7, 6] ATM standard
    deviations, kept in [0, 1] and non-decreasing, and inverted by linear interpolation in k over the nodes where
    it strictly rises."""
    sd = float(np.sqrt(surface.theta(T)))
    k = np.linspace(-7.0 * sd, 6.0 * sd, n_k)
    w, w_k = _w_and_slope(surface, k, T)
    s = np.sqrt(np.maximum(w, 1e-12))
    d2 = -k / s - 0.5 * s
    density_term = np.exp(-0.5 * d2 * d2) / np.sqrt(2.0 * np.pi) * w_k / (2.0 * s)
    cdf = np.maximum.accumulate(np.clip(ndtr(-d2) + density_term, 0.0, 1.0))
    rises = np.concatenate([[True], cdf[1:] > cdf[:-1] + 1e-14])
    u_nodes, k_nodes = cdf[rises], k[rises]
    return lambda u: np.exp(np.interp(u, u_nodes, k_nodes))


def skewed_matrix(matrix, strike: float, skew: CorrelationSkew = CorrelationSkew()) -> np.ndarray:
    """The correlation a product around `strike` sees: every pair moved by the skew's offset, kept in [-0.95, 0.99]
    and repaired to a correlation matrix."""
    m = np.asarray(matrix, dtype=float)
    moved = np.clip(m + skew.offset(strike), -0.95, 0.99)
    moved[np.diag_indices_from(moved)] = 1.0
    return nearest_correlation(moved)


def copula_spots(quantiles, forwards, matrix, n_pairs: int, seed: int = SEED) -> np.ndarray:
    """(2 n_pairs, n) antithetic pairs of S_i(T) / S_i(0) (rows p and n_pairs + p are a pair): standard normals from
    `default_rng(seed)` correlated by `matrix`'s eigen-root, to uniforms, through each member's quantile, times its
    forward over spot."""
    g = np.random.default_rng(seed).standard_normal((n_pairs, len(quantiles)))
    g = g @ correlation_root(np.asarray(matrix, dtype=float)).T
    u = ndtr(np.concatenate([g, -g]))
    return np.column_stack([float(f) * q(u[:, j]) for j, (f, q) in enumerate(zip(forwards, quantiles))])


def spread_payoff(kind: str, strike: float, x) -> np.ndarray:
    """The option on the equally weighted basket less the average of the same option on every member, per path
    (x: (paths, members) of S_T / S_0). At one strike the CvC and the PvP pay the same on every path."""
    x = np.asarray(x, dtype=float)
    basket = x.mean(axis=1)
    if kind == "cvc":
        return np.maximum(basket - strike, 0.0) - np.maximum(x - strike, 0.0).mean(axis=1)
    if kind == "pvp":
        return np.maximum(strike - basket, 0.0) - np.maximum(strike - x, 0.0).mean(axis=1)
    raise ValueError(f"a spread is 'cvc' or 'pvp', not {kind!r}")


def _pair_estimate(pay: np.ndarray) -> tuple:
    """(mean, standard error) of a per-path value over antithetic pairs (rows p and n/2 + p)."""
    h = len(pay) // 2
    pairs = 0.5 * (pay[:h] + pay[h:])
    return float(pairs.mean()), float(pairs.std(ddof=1) / np.sqrt(h))


def member_vols(surfaces, forwards, strike: float, T: float) -> np.ndarray:
    """Each member's implied vol at the structure's one strike (in its own log forward moneyness)."""
    k = np.log(strike / np.asarray(forwards, dtype=float))
    w = [float(np.asarray(sf.w(np.array([kj]), T), dtype=float)[0]) for sf, kj in zip(surfaces, k)]
    return np.sqrt(np.array(w) / T)


def _loadings(forwards, vols) -> np.ndarray:
    """a_i = (F_i / sum F) sigma_i: a member's share of the basket times its vol."""
    f = np.asarray(forwards, dtype=float)
    return f * np.asarray(vols, dtype=float) / f.sum()


def average_correlation(matrix, forwards, vols) -> float:
    """The basket variance's one correlation: (a' C a - sum a_i^2) / ((sum a)^2 - sum a_i^2), the off-diagonal
    pairs weighted by a_i a_j."""
    a = _loadings(forwards, vols)
    m = np.asarray(matrix, dtype=float)
    own = float(np.sum(a * a * np.diag(m)))
    return (float(a @ m @ a) - own) / (float(a.sum()) ** 2 - float(np.sum(a * a)))


def read_correlation(basket_price: float, kind: str, strike: float, T: float, forwards, vols) -> tuple:
    """(correlation, basket vol) read off an undiscounted basket option taken as lognormal: its implied vol s_B
    solves s_B^2 = sum a_i^2 + rho ((sum a_i)^2 - sum a_i^2) for rho."""
    from optlab.core.blackscholes import implied_vol

    f = np.asarray(forwards, dtype=float)
    s_b = float(implied_vol(basket_price, float(f.mean()), strike, T, 1.0, CP[kind]))
    a = _loadings(f, vols)
    diag = float(np.sum(a * a))
    return (s_b * s_b - diag) / (float(a.sum()) ** 2 - diag), s_b


def _line(surfaces, forwards, x, matrix, kind: str, strike: float, T: float) -> dict:
    """One pseudo-quote, in % of notional and undiscounted: the spread from the copula paths, the members in
    closed form at their own vols at this strike, the basket as their sum."""
    from optlab.core.blackscholes import black_price

    spread, spread_se = _pair_estimate(spread_payoff(kind, strike, x))
    vols = member_vols(surfaces, forwards, strike, T)
    members = float(np.mean([black_price(float(fj), strike, T, float(vj), 1.0, CP[kind])
                             for fj, vj in zip(forwards, vols)]))
    implied, s_b = read_correlation(members + spread, kind, strike, T, forwards, vols)
    return {"kind": kind, "strike": strike, "tenor": T, "price": 100.0 * spread, "price_se": 100.0 * spread_se,
            "basket_price": 100.0 * (members + spread), "members_price": 100.0 * members, "basket_vol": s_b,
            "rho": average_correlation(matrix, forwards, vols), "implied_rho": implied}


def pseudo_quotes(surfaces, forward, base_matrix, tenors=(0.25, 0.5, 1.0), strikes=STRIKES, n_pairs: int = 100_000,
                  skew: CorrelationSkew = CorrelationSkew(), seed: int = SEED, label: str = "", day=None) -> list:
    """Every pseudo-quote of one basket (`forward(T)`: each member's forward over spot). One copula simulation a
    (tenor, strike), on the same seed for every strike; the CvC and the PvP of one strike share it."""
    by_strike: dict = {}
    for kind, ks in strikes.items():
        for k in ks:
            by_strike.setdefault(k, []).append(kind)
    rows = []
    for T in tenors:
        f = np.asarray(forward(T), dtype=float)
        quantiles = [marginal_quantile(sf, T) for sf in surfaces]
        for strike in sorted(by_strike):
            matrix = skewed_matrix(base_matrix, strike, skew)
            x = copula_spots(quantiles, f, matrix, n_pairs, seed)
            rows += [{"basket": label, "obs_date": day, **_line(surfaces, f, x, matrix, kind, strike, T)}
                     for kind in strikes if kind in by_strike[strike]]
    return rows


def targets(surfaces, markets, corr, T: float = 1.0, skew=None, names=None) -> list:
    """The three lines a tenor is calibrated on (localcorr.py:130-138): PvP 90 %, CvC 100 % and 110 %, by strike."""
    label = "+".join(names or (f"n{i}" for i in range(len(surfaces))))
    rows = pseudo_quotes(surfaces, lambda t: [float(m.F(t)) / float(m.spot) for m in markets], corr, tenors=(T,),
                         skew=skew or CorrelationSkew(), label=label)
    return sorted((r for r in rows if (r["kind"], r["strike"]) in LINES), key=lambda r: r["strike"])


def basket_sigma(surfaces, corr, T: float = 1.0) -> float:
    """ATM vol of the equally weighted basket from the members' ATM vols and the correlation (localcorr.py:122-127)."""
    v = np.sqrt(np.array([float(sf.theta(T)) for sf in surfaces]) / T)
    return float(np.sqrt(v @ np.asarray(corr, dtype=float) @ v)) / len(v)


# ---- the fit on the engine (localcorr.py:141-213) -----------------------------------------------------------------
def basket_model(surfaces, markets, corr, mixer, names=None):
    """Local vol for every member, the measured correlation, the mixer."""
    from optlab.core.models import LocalVolModel
    from optlab.structured.multiasset import Basket

    names = tuple(names or (f"n{i}" for i in range(len(surfaces))))
    return Basket(names, {n: LocalVolModel(sf) for n, sf in zip(names, surfaces)}, dict(zip(names, markets)), corr,
                  mixer=mixer)


def model_quotes(basket, wanted, T: float, n_paths: int, seed: int, backend: str = "auto") -> tuple:
    """(prices in %, standard errors) of the lines `wanted` under the basket's mixer: the basket simulated by the
    path engine to T on steps of at most 1/50."""
    from optlab import native

    paths = native.simulate(basket, None, [T], n_paths, seed, dt_max=1 / 50, backend=backend)
    x = np.column_stack([paths[n][:, 0] / float(basket.markets[n].spot) for n in basket.names])
    est = [_pair_estimate(100.0 * spread_payoff(r["kind"], r["strike"], x)) for r in wanted]
    return np.array([e[0] for e in est]), np.array([e[1] for e in est])


def gauss_newton(quotes, goal, start, bounds, steps: int = 6, h: float = 0.02) -> tuple:
    """(best parameters, their RMS error) fitting quotes(p) -> (prices, standard errors) to `goal` on common random
    numbers. A step: the Jacobian by one-sided differences of h (downwards where upwards would pass the upper bound),
    the least-squares step with singular values below 1e-3 of the largest dropped, taken at 0.8 and clipped to the
    bounds. The fit ends early once the RMS error is below the prices' mean standard error, where further steps
    would fit the noise; the best point seen is returned."""
    goal = np.asarray(goal, dtype=float)
    lo, hi = (np.asarray(b, dtype=float) for b in bounds)
    p = np.array(start, dtype=float)
    best_err, best = np.inf, p.copy()
    for _ in range(steps):
        f, noise = quotes(p)
        err = float(np.sqrt(np.mean((f - goal) ** 2)))
        if err < best_err:
            best_err, best = err, p.copy()
        if err < float(np.mean(noise)):
            break
        moves = np.where(p + h <= hi, h, -h)
        jac = np.column_stack([(quotes(p + m * e)[0] - f) / ((p + m * e)[j] - p[j])
                               for j, (m, e) in enumerate(zip(moves, np.eye(p.size)))])
        p = np.clip(p + 0.8 * np.linalg.lstsq(jac, goal - f, rcond=1e-3)[0], lo, hi)
    return best, best_err


class Form(NamedTuple):
    """A mixer family the fit can take: its constructor from (sigma, parameters), start and (lower, upper) bounds."""

    make: Callable
    start: tuple
    bounds: tuple


FORMS = {
    "local": Form(lambda sigma, p: LocalCorrelation(sigma, p[0], p[1], p[2]), (0.0, 0.1, 0.1),
                  ((-W0, 0.0, 0.0), (1.0 - W0, 1.5, 1.5))),
    "stochastic": Form(lambda sigma, p: StochasticMix(p[0], p[1], p[2]), (W0, 0.5, -0.5),
                       ((0.01, 0.05, -0.95), (0.99, 3.0, 0.95))),
}


def calibrate(surfaces, markets, corr, T: float = 1.0, n_paths: int = 40_000, seed: int = SEED, form: str = "local",
              skew=None, names=None, wanted=None, backend: str = "auto") -> dict:
    """The mixer fitted to a tenor's pseudo-quotes (localcorr.py:153-213), every model price simulated on the
    engine: {"mixer", "lines", "rmse", "origin", "note", "form"}; the lines are judged on other paths (seed + 1).
    The reference reads surfaces, markets and the correlation from a basket context; here they are arguments.
    `wanted` replaces the pseudo-quotes by lines of one's own (dicts with kind, strike, price, rho)."""
    shape = FORMS.get(form)
    if shape is None:
        raise ValueError(f"form is one of {sorted(FORMS)}, not {form!r}")
    if wanted is None:
        wanted = targets(surfaces, markets, corr, T, skew, names)
    sigma = basket_sigma(surfaces, corr, T)

    def quotes(p, s=seed):
        mixer = shape.make(sigma, p)
        return model_quotes(basket_model(surfaces, markets, corr, mixer, names), wanted, T, n_paths, s, backend)

    params, rmse = gauss_newton(quotes, [r["price"] for r in wanted], shape.start, shape.bounds)
    judged, judged_se = quotes(params, seed + 1)
    lines = [dict(kind=r["kind"], strike=r["strike"], tenor=T, pseudo_quote=r["price"], model=float(m), se=float(e),
                  assumed_rho=r.get("rho")) for r, m, e in zip(wanted, judged, judged_se)]
    return {"mixer": shape.make(sigma, params), "lines": lines, "rmse": rmse, "origin": ORIGIN, "note": NOTE,
            "form": form}
