"""The Longstaff-Schwartz oracle (oracle/lsm_oracle.py) pinned on the CPU: its binomial-tree reference against the
Black-Scholes European and its own convergence; the LSM estimate on NumPy Black-Scholes paths against the tree (a
lower bound within Monte Carlo error); the normal equations against lstsq; the immediate-exercise floor; the oracle's
variance factor at the observations (the basis with v) against its deterministic recursion."""

from __future__ import annotations

import sys

import numpy as np
import pytest
from _support import ROOT, oracle_paths, oracle_paths_var
from scipy.stats import norm

sys.path.insert(0, str(ROOT / "oracle"))
import lsm_oracle as lo  # noqa: E402

S0, K, R, SIGMA, T, N_EX = 36.0, 40.0, 0.06, 0.2, 1.0, 50  # Longstaff and Schwartz (2001), table 1, first row


def bs_put(S, K, r, sigma, T):
    d1 = (np.log(S / K) + (r + 0.5 * sigma**2) * T) / (sigma * np.sqrt(T))
    return K * np.exp(-r * T) * norm.cdf(-(d1 - sigma * np.sqrt(T))) - S * norm.cdf(-d1)


def gbm_paths(rng, n_paths, S=S0, r=R, sigma=SIGMA, T=T, n_ex=N_EX):
    dt = T / n_ex
    z = rng.standard_normal((n_paths // 2, n_ex))
    z = np.vstack([z, -z])  # rows i and i + n / 2 antithetic, as the engine lays them out
    return S * np.exp(np.cumsum((r - 0.5 * sigma**2) * dt + sigma * np.sqrt(dt) * z, axis=1))


def test_tree_is_the_european_with_one_date_and_converges_with_more_steps():
    assert lo.bermudan_tree(S0, K, R, 0.0, SIGMA, T, 1, 4000, today=False) == pytest.approx(
        bs_put(S0, K, R, SIGMA, T), abs=2e-3)
    a = lo.bermudan_tree(S0, K, R, 0.0, SIGMA, T, N_EX, 40)
    b = lo.bermudan_tree(S0, K, R, 0.0, SIGMA, T, N_EX, 80)
    assert a == pytest.approx(b, abs=3e-4)
    assert a == pytest.approx(4.478, abs=2e-3)  # the paper's American value (50 dates are nearly continuous)
    assert bs_put(S0, K, R, SIGMA, T) < a


def test_lsm_on_gbm_paths_is_a_lower_bound_of_the_tree_and_the_solvers_agree():
    rng = np.random.default_rng(3)
    S1, S2 = gbm_paths(rng, 1 << 15), gbm_paths(rng, 1 << 15)
    disc = np.exp(-R * np.linspace(T / N_EX, T, N_EX))
    tree = lo.bermudan_tree(S0, K, R, 0.0, SIGMA, T, N_EX, 40)
    p_n, se_n, c_n = lo.lsm_price(S1, None, S2, None, disc, K, S0, solver="normal")
    p_l, se_l, c_l = lo.lsm_price(S1, None, S2, None, disc, K, S0, solver="lstsq")
    assert p_n == pytest.approx(p_l, rel=1e-12) and se_n == pytest.approx(se_l, rel=1e-12)
    # the monomials of s = S / K over the in-the-money paths are nearly collinear (the normal matrix's condition
    # number reaches 3e10 at the first dates): the two solvers' coefficients differ in the sixth digit, their fitted
    # continuation values (about 0.1 in units of the strike) by less than 1e-7, and the prices not at all
    s1 = S1 / K
    for j, (a, b) in enumerate(zip(c_n, c_l)):
        assert (a is None) == (b is None)
        if a is not None:
            X = lo.basis(s1[s1[:, j] < 1.0, j])
            np.testing.assert_allclose(X @ a, X @ b, rtol=0, atol=1e-7)
    assert tree - 4 * se_n - 0.03 < p_n < tree + 4 * se_n


def test_deep_in_the_money_put_is_worth_its_immediate_exercise():
    rng = np.random.default_rng(4)
    S1, S2 = gbm_paths(rng, 4096, S=20.0), gbm_paths(rng, 4096, S=20.0)
    disc = np.exp(-R * np.linspace(T / N_EX, T, N_EX))
    price, se, _ = lo.lsm_price(S1, None, S2, None, disc, K, 20.0)
    assert price == 20.0 and se == 0.0


def test_the_basis_with_a_variance_factor_has_seven_columns():
    s, v = np.array([0.9, 0.8]), np.array([0.04, 0.05])
    X = lo.basis(s, v)
    np.testing.assert_array_equal(X[:, 6], s * v)
    assert X.shape == (2, 7) and lo.basis(s).shape == (2, 4)


def test_oracle_variance_factor_at_the_observations():
    """oracle_paths_var (the paths with v kept for the basis with v): the spots are oracle_paths' own; with xi = 0
    the variance factor is deterministic and follows the Euler recursion v += kappa (theta - max(v, 0)) dt, the same
    on both legs; with xi > 0 the legs' variances differ and the antithetic pair's mean over many pairs stays near
    the recursion's value (E[v] obeys it while v stays positive)."""
    from optlab import native
    from optlab.core.market import Market
    from optlab.core.models import HestonModel, HestonParams

    market = Market.flat(100.0, 0.03, 0.0)
    times = np.linspace(0.1, 1.0, 10)
    p = native.plan(HestonModel(HestonParams(0.09, 1.5, 0.04, 0.0, -0.7)), market, times)
    S, V = oracle_paths_var(p, 64, 5)
    assert np.array_equal(S, oracle_paths(p, 64, 5)[0].T)
    v, want = 0.09, []
    for i in range(1, len(p.dt) + 1):
        v += 1.5 * (0.04 - max(v, 0.0)) * p.dt[i - 1]
        want += [v] * int(np.sum(p.obs_step == i))
    assert np.allclose(V, np.array(want)[None, :], rtol=1e-14, atol=0.0)
    q = native.plan(HestonModel(HestonParams(0.09, 1.5, 0.04, 0.3, -0.7)), market, times)
    S, V = oracle_paths_var(q, 20_000, 5)
    assert np.array_equal(S, oracle_paths(q, 20_000, 5)[0].T)
    assert not np.array_equal(V[:20_000], V[20_000:])
    assert np.max(np.abs(V.mean(axis=0) - np.array(want))) < 2e-3
