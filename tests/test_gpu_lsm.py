"""American options by Longstaff-Schwartz on the device (optlab_lsm): FP64 against the oracle's LSM on the same
paths (the engine's own, fetched through optlab_paths); FP32 against a Bermudan binomial tree; the Heston basis with
the variance factor (FP32 against FP64 on the same draws, at least the European); determinism and checked requests.
"""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest
from _support import ROOT, oracle_paths_var
from test_oracle import _golden_models

sys.path.insert(0, str(ROOT / "oracle"))
import lsm_oracle as lo  # noqa: E402

from optlab import native  # noqa: E402
from optlab.core.market import Market  # noqa: E402
from optlab.core.models import BlackScholesModel  # noqa: E402
from optlab.native import _abi  # noqa: E402

pytestmark = pytest.mark.gpu

S0, K, R, SIGMA, T, N_EX = 36.0, 40.0, 0.06, 0.2, 1.0, 50  # Longstaff and Schwartz (2001), table 1, first row


def engine_paths(model, market, times, n_paths, seed, first_pair):
    p = native.plan(model, market, times)
    spots, _, _ = native.run(p, n_paths // 2, native.seed64(seed), first_pair=first_pair, backend="fp64")
    return spots[0]


@pytest.mark.parametrize("case", ["bs_put", "bs_put_ragged", "bs_put_deep_otm", "bs_put_low_vol", "lv_put", "lv_call"])
def test_fp64_lsm_is_the_oracle_on_the_same_paths(case):
    if case == "bs_put_deep_otm":  # S0 52 against K 40: few paths and a narrow band in the money at the early dates
        model, market, strike, cp, spot = BlackScholesModel(SIGMA), Market.flat(52.0, R, 0.0), K, "put", 52.0
    elif case == "bs_put_low_vol":  # at the money, sigma 0.08: the band in the money is a few percent wide
        model, market, strike, cp, spot = BlackScholesModel(0.08), Market.flat(40.0, R, 0.0), K, "put", 40.0
    elif case.startswith("bs_put"):
        model, market, strike, cp, spot = BlackScholesModel(SIGMA), Market.flat(S0, R, 0.0), K, "put", S0
    else:
        market, models = _golden_models()
        model, spot = models["lv"], market.spot
        strike, cp = (spot * 1.05, "put") if case == "lv_put" else (spot * 0.95, "call")
    # ragged: 5001 pairs, so the last workgroup of the pricing pass and the sweep's last tile are partly empty
    n_paths, seed = (2 * 5001 if case == "bs_put_ragged" else 1 << 14), 17
    times = np.linspace(T / N_EX, T, N_EX)
    print(f"lsm fp64 {case}: engine", flush=True)
    got = native.american_lsm(model, market, strike, T, cp, N_EX, n_paths, rng=seed, backend="fp64")
    print("oracle", flush=True)
    S1 = engine_paths(model, market, times, n_paths, seed, 0)
    S2 = engine_paths(model, market, times, n_paths, seed, n_paths // 2)
    disc = market.D(times)
    want, want_se, coefs = lo.lsm_price(S1, None, S2, None, disc, strike, spot, call=cp == "call")
    assert got.price == pytest.approx(want, rel=1e-10)
    assert got.stderr == pytest.approx(want_se, rel=1e-8)
    s1 = S1 / strike
    phi = 1.0 if cp == "call" else -1.0
    for j, c in enumerate(coefs):
        if c is None:
            assert np.all(np.isnan(got.coefs[j]))
            continue
        X = lo.basis(s1[phi * (s1[:, j] - 1.0) > 0, j])
        A = X.T @ X
        # The normal equations are ill-conditioned (condition number 1e8 to 1e15 for a narrow in-the-money band):
        # another summation order moves the coefficients, and even the fitted values, far beyond rounding. What a
        # correct solve fixes is the backward error of the device's coefficients on the oracle's system: about
        # 1e-16 for a different summation order, 1e-5 for a 1e-4 slip in one coefficient.
        back = np.linalg.norm(A @ (got.coefs[j, :4] - c)) / (np.linalg.norm(A) * np.linalg.norm(c))
        assert back < 1e-12, (j, back)
        assert np.all(got.coefs[j, 4:] == 0.0)
    # the reference's lstsq on the same paths: the same exercise decisions up to ties
    ref, ref_se, _ = lo.lsm_price(S1, None, S2, None, disc, strike, spot, call=cp == "call", solver="lstsq")
    print(case, "device", got.price, "lstsq", ref, "se", ref_se, flush=True)
    assert got.price == pytest.approx(ref, rel=1e-6)


def test_fp64_lsm_with_the_variance_basis_is_the_oracle_on_its_paths():
    """The basis with v (seven terms) against the oracle's LSM on the oracle's own Heston paths and variance factor:
    the device's FP64 paths agree with those to about 1e-12, and on this seed a 1e-12 relative perturbation of S
    and V moves the oracle's price by 1e-13 at most (no exercise decision flips), so the prices must agree to 1e-10."""
    market, models = _golden_models()
    model = models["heston"]
    strike = market.spot * 1.1
    n_paths, seed = 1 << 14, 29
    times = np.linspace(T / N_EX, T, N_EX)
    print("lsm fp64 heston, basis with v: engine", flush=True)
    got = native.american_lsm(model, market, strike, T, "put", N_EX, n_paths, rng=seed, backend="fp64")
    print("oracle", flush=True)
    p = native.plan(model, market, times)
    S1, V1 = oracle_paths_var(p, n_paths // 2, native.seed64(seed), 0)
    S2, V2 = oracle_paths_var(p, n_paths // 2, native.seed64(seed), n_paths // 2)
    want, want_se, coefs = lo.lsm_price(S1, V1, S2, V2, market.D(times), strike, market.spot)
    assert got.price == pytest.approx(want, rel=1e-10)
    assert got.stderr == pytest.approx(want_se, rel=1e-8)
    s1 = S1 / strike
    for j, c in enumerate(coefs):
        if c is None:
            assert np.all(np.isnan(got.coefs[j]))
            continue
        itm = s1[:, j] < 1.0
        X = lo.basis(s1[itm, j], V1[itm, j])
        A = X.T @ X
        # backward error of the device's seven coefficients on the oracle's system; the two path sets differ by
        # about 1e-12, so 1e-9 (a 1e-4 slip in one coefficient gives about 1e-5)
        back = np.linalg.norm(A @ (got.coefs[j] - c)) / (np.linalg.norm(A) * np.linalg.norm(c))
        assert back < 1e-9, (j, back)


def test_fp32_bermudan_put_is_within_monte_carlo_error_of_the_tree():
    tree = lo.bermudan_tree(S0, K, R, 0.0, SIGMA, T, N_EX, 40)
    got = native.american_lsm(BlackScholesModel(SIGMA), Market.flat(S0, R, 0.0), K, T, "put", N_EX, 1 << 18, rng=5,
                              backend="fp32")
    assert got.timing["precision"] == "fp32"
    assert tree - 4 * got.stderr - 0.02 < got.price < tree + 4 * got.stderr, (got.price, got.stderr, tree)
    again = native.american_lsm(BlackScholesModel(SIGMA), Market.flat(S0, R, 0.0), K, T, "put", N_EX, 1 << 18,
                                rng=5, backend="fp32")
    assert again.price == got.price and np.array_equal(again.coefs, got.coefs, equal_nan=True)


def test_heston_lsm_with_the_variance_factor():
    market, models = _golden_models()
    model = models["heston"]
    strike = market.spot * 1.1
    a = native.american_lsm(model, market, strike, T, "put", 25, 1 << 16, rng=3, backend="fp32")
    b = native.american_lsm(model, market, strike, T, "put", 25, 1 << 16, rng=3, backend="fp64")
    assert np.all(np.isfinite(a.coefs[:-1, 4:7]) | np.isnan(a.coefs[:-1, 0:1]))  # the v columns are fitted
    assert abs(a.price - b.price) < 4 * np.hypot(a.stderr, b.stderr)
    # at least the European on the same draws (the American holds the European's exercise at maturity)
    S = native.simulate(model, market, [T], 1 << 16, seed=3, backend="fp64")[:, -1]
    pay = market.D(T) * np.maximum(strike - S, 0.0)
    euro = 0.5 * (pay[: len(pay) // 2] + pay[len(pay) // 2:])
    assert b.price > euro.mean() - 4 * euro.std(ddof=1) / np.sqrt(len(euro))


def test_requests_are_checked():
    market, models = _golden_models()
    with pytest.raises(ValueError, match="basis_var"):
        p = native.plan(BlackScholesModel(0.2), market, np.linspace(0.1, 1.0, 10))
        pr, keep, _ = native.paths_request(p, 1024, 1)
        rq = _abi.LsmRequest()
        rq.paths = pr
        disc = np.ones(10)
        rq.strike, rq.disc, rq.spot, rq.basis_var, rq.min_itm = 100.0, _abi.ptr(disc), 100.0, 1, 50
        native._check(_abi.load(), _abi.load().optlab_lsm(C.byref(rq), 1))
    with pytest.raises(ValueError, match="one name"):
        b, _ = __import__("test_oracle")._basket()
        native.american_lsm(b, market, 100.0, 1.0, "put", 10, 1024)
