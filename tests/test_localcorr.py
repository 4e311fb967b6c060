"""Local-correlation calibration on the host (no GPU): the pseudo-quote builders against what the mathematics fixes
(a flat surface's marginal is lognormal; a lognormal basket's correlation read back exactly; the CvC and the PvP of
one strike pay the same on every path; a zero skew leaves the copula's correlation where it was) and the
Gauss-Newton fit against the oracle's restatement (localcorr.py:153-178) on the same price function."""

from __future__ import annotations

import sys

import numpy as np
import pytest
from _support import ROOT
from scipy.stats import norm

sys.path.insert(0, str(ROOT / "oracle"))
import localcorr_oracle as lco  # noqa: E402

from optlab.core.blackscholes import black_price  # noqa: E402
from optlab.structured import localcorr as lc  # noqa: E402


class FlatSurface:
    def __init__(self, vol):
        self.vol = vol

    def theta(self, T):
        return self.vol**2 * T

    def w(self, k, T):
        return np.full(np.shape(k), self.vol**2 * T)


def test_a_flat_surface_has_the_lognormal_quantile():
    q = lc.marginal_quantile(FlatSurface(0.25), 1.0)
    u = np.linspace(0.001, 0.999, 41)
    want = np.exp(-0.5 * 0.25**2 + 0.25 * norm.ppf(u))
    np.testing.assert_allclose(q(u), want, rtol=2e-3)


@pytest.mark.parametrize("kind,strike", [("cvc", 1.0), ("cvc", 1.1), ("pvp", 0.9)])
def test_a_lognormal_basket_gives_its_correlation_back(kind, strike):
    f, vols, rho, T = np.array([1.01, 1.02, 0.99]), np.array([0.2, 0.25, 0.3]), 0.55, 1.0
    a = f / f.sum() * vols
    s_b = np.sqrt(np.sum(a * a) + rho * (a.sum() ** 2 - np.sum(a * a)))
    price = black_price(f.mean(), strike, T, s_b, 1.0, lc.CP[kind])
    got, vol = lc.read_correlation(price, kind, strike, T, f, vols)
    assert got == pytest.approx(rho, abs=1e-8) and vol == pytest.approx(s_b, rel=1e-9)


def test_one_strike_cvc_and_pvp_pay_the_same_and_spreads_are_not_positive():
    x = np.random.default_rng(1).lognormal(0.0, 0.3, (1000, 4))
    for strike in (0.9, 1.0, 1.1):
        np.testing.assert_allclose(lc.spread_payoff("cvc", strike, x), lc.spread_payoff("pvp", strike, x), atol=1e-14)
    assert np.all(lc.spread_payoff("cvc", 1.0, x) <= 1e-15)  # the option on an average is worth less (convexity)
    np.testing.assert_allclose(lc.spread_payoff("cvc", 1.0, x), lco.spread("cvc", 1.0, x), atol=1e-14)


def test_without_skew_the_pseudo_quotes_read_back_the_measured_correlation():
    from optlab.core.market import Market

    surfaces = [FlatSurface(v) for v in (0.2, 0.25, 0.3)]
    markets = [Market.flat(100.0, 0.02, 0.0)] * 3
    corr = np.full((3, 3), 0.6) + 0.4 * np.eye(3)
    lines = lc.targets(surfaces, markets, corr, 1.0, skew=lc.CorrelationSkew(0.0, 0.0))
    assert [(r["kind"], r["strike"]) for r in lines] == [("pvp", 0.9), ("cvc", 1.0), ("cvc", 1.1)]
    for r in lines:
        assert r["rho"] == pytest.approx(0.6, abs=1e-12)
        assert r["implied_rho"] == pytest.approx(0.6, abs=0.03)  # the lognormal read of a non-lognormal basket
        assert r["price"] < 0
    skewed = lc.targets(surfaces, markets, corr, 1.0)  # the default skew: more correlation below the money
    assert skewed[0]["rho"] > 0.6 > skewed[2]["rho"]


def test_the_fit_is_the_oracle_restatement_step_for_step():
    """A smooth nonlinear price function with a kink where a parameter reaches its bound: the product's fit and the
    oracle's take the same steps to the same parameters."""
    A = np.array([[1.0, -0.5, 0.3], [0.2, 1.2, -0.4], [-0.3, 0.1, 0.9]])
    target = np.array([0.25, 0.6, 0.4])

    def prices(p):
        return A @ np.tanh(p) + 0.1 * p**2, np.full(3, 1e-6)

    goal = prices(target)[0]
    got = lc.gauss_newton(prices, goal, lc.FORMS["local"][1], lc.FORMS["local"][2])
    want = lco.fit(prices, goal)
    np.testing.assert_allclose(got[0], want[0], rtol=1e-13, atol=1e-15)
    assert got[1] == pytest.approx(want[1], rel=1e-12, abs=1e-15)
    assert np.linalg.norm(got[0] - target) < 0.05  # and it gets there


def test_nearest_correlation_is_highams_published_example():
    """Higham (2002), section 4: the nearest correlation matrix to the tridiagonal [-1, 2, -1] matrix of order 4,
    printed there to four places; a matrix that already is a correlation comes back unchanged."""
    from optlab.structured.multiasset import correlation_root, nearest_correlation

    a = 2.0 * np.eye(4) - np.eye(4, k=1) - np.eye(4, k=-1)
    x = nearest_correlation(a)
    want = np.array([[1.0, -0.8084, 0.1916, 0.1068], [-0.8084, 1.0, -0.6562, 0.1916],
                     [0.1916, -0.6562, 1.0, -0.8084], [0.1068, 0.1916, -0.8084, 1.0]])
    np.testing.assert_allclose(x, want, atol=6e-5)
    assert np.all(np.diag(x) == 1.0) and np.linalg.eigvalsh(x)[0] > -1e-12
    r = correlation_root(x)
    np.testing.assert_allclose(r @ r.T, x, atol=1e-12)
    c = np.array([[1.0, 0.5, 0.2], [0.5, 1.0, 0.3], [0.2, 0.3, 1.0]])
    assert np.array_equal(nearest_correlation(c), c)


def test_low_matrix_gives_the_measured_matrix_back_at_w0_and_factors_are_aligned():
    c = np.array([[1.0, 0.6, 0.45], [0.6, 1.0, 0.5], [0.45, 0.5, 1.0]])
    low = lc.low_matrix(c, lc.W0)
    off = ~np.eye(3, dtype=bool)
    np.testing.assert_allclose(lc.W0 + (1.0 - lc.W0) * low[off], c[off], atol=1e-15)
    from optlab.structured.multiasset import correlation_root

    root = correlation_root(c)
    f, sign, angle = lc.aligned_factors(c, root, lc.W0)
    np.testing.assert_allclose(f @ f.T, low, atol=1e-12)
    assert np.all(np.einsum("ij,ij->j", f[:, :-1], root[:, :-1]) >= 0) and f[:, -1].mean() > 0
    assert sign * root[:, -1].mean() > 0
    # the angle of the common factor: tan a = sqrt(1 - w0) mean(top column) / sqrt(w0)
    assert np.tan(angle) == pytest.approx(np.sqrt(1.0 - lc.W0) * f[:, -1].mean() / np.sqrt(lc.W0), rel=1e-12)


def test_equicorrelated_factors_draw_on_the_roots_own_columns():
    """With repeated eigenvalues the eigen-solver's basis is arbitrary: the aligned factor's other columns are the
    root's, scaled by 1 / sqrt(1 - w0), and at w0 the common factor cancels the second normal N and gives every
    name the root's loading on the market normal, so the mixer's spots at w0 are the constant matrix's draws."""
    from optlab.structured.multiasset import correlation_root

    w0 = lc.W0
    for n, rho in ((3, 0.5), (4, 0.6), (5, 0.45)):
        c = np.full((n, n), rho) + (1.0 - rho) * np.eye(n)
        root = correlation_root(c)
        f, sign, a = lc.aligned_factors(c, root, w0)
        np.testing.assert_allclose(np.sqrt(1.0 - w0) * f[:, :-1], root[:, :-1], atol=1e-12)
        top = f[:, -1]
        on_n = np.sqrt(w0) * np.sin(a) - np.sqrt(1.0 - w0) * top * np.cos(a)
        on_m = np.sqrt(w0) * np.cos(a) + np.sqrt(1.0 - w0) * top * np.sin(a)
        np.testing.assert_allclose(on_n, 0.0, atol=1e-12)
        np.testing.assert_allclose(sign * on_m, root[:, -1], atol=1e-12)
