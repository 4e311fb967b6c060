"""The stochastic mixing weight (mixer 2; localcorr.py:87-114 StochasticMix, which the reference's compiled engine
refuses) in the oracle's path restatement, pinned on the CPU:
- a weight that cannot move (nu 0, theta = w0) gives the constant-matrix local correlation's paths bit for bit,
  Heston residuals included (the weight's driver is a stream of its own);
- every name keeps its law whatever the weight does (test_structured_localcorr.py:66-79);
- with nu 0 the weight is deterministic, w_i = theta + (w0 - theta) (1 - kappa dt)^i, and sets each step's
  correlation of the names' increments to w_i + (1 - w_i) c_low;
- the reference's own NumPy basket simulator under StochasticMix (tests/golden/stochastic_mix.json, written by
  tests/golden/make_mix_fixtures.py) within Monte Carlo error on ten payoffs and two weight regimes: basket,
  worst-of and best-of options, the squared spread and each path's realized correlation (the weight's mean and
  variance); and the standard errors of the pair averages within 6 % of the reference's at the same number of
  pairs, which is where the driver's antithetic pairing shows.
Planted in a copy of this oracle, each of these mistakes fails the last test: rho's sign (22 SE), no independent
part in the weight's driver (32 SE), sqrt(w) for sqrt(w (1 - w)) (49 SE), half the mean reversion (270 SE), a
driver that is not antithetic (standard errors 12-17 % above the reference's).
"""

from __future__ import annotations

import json
import sys

import numpy as np
import pytest
from _support import ROOT, oracle_paths

from optlab import native
from optlab.core.blackscholes import black_price
from optlab.core.market import Market
from optlab.core.models import BlackScholesModel
from optlab.structured.localcorr import LocalCorrelation, StochasticMix, low_matrix
from optlab.structured.multiasset import Basket

sys.path.insert(0, str(ROOT / "tests" / "golden"))
from make_mix_fixtures import payoffs  # noqa: E402

GOLD = json.loads((ROOT / "tests" / "golden" / "stochastic_mix.json").read_text())


def gold_basket(mixer, corr=None, vols=None):
    names = tuple(GOLD["names"])
    market = Market.flat(GOLD["spot"], GOLD["r"], GOLD["q"])
    c = GOLD["rho"] if corr is None else corr
    return Basket(names, {n: BlackScholesModel(v) for n, v in zip(names, vols or GOLD["vols"])},
                  {n: market for n in names}, np.array([[1.0, c], [c, 1.0]]), "same", mixer), market


def mix_of(case):
    m = GOLD["cases"][case]["mixer"]
    return StochasticMix(m["theta"], m["nu"], m["rho"], m["kappa"], m["w0"])


def test_a_still_weight_is_the_constant_matrix_bit_for_bit():
    from test_oracle import _basket

    b, _ = _basket(mixer=True, slv=True)  # local vol, SLV (aux stream) and Black-Scholes names
    still = Basket(b.names, b.models, b.markets, b.corr, b.vol_corr, StochasticMix(0.3, 0.0, -0.5, 2.0, 0.3))
    flat = Basket(b.names, b.models, b.markets, b.corr, b.vol_corr, LocalCorrelation(0.15, 0.0, 0.0, 0.0, 0.3))
    times = [0.1, 0.25, 0.5]
    a = oracle_paths(native.plan(still, None, times, dt_max=1 / 100), 600, 5)
    c = oracle_paths(native.plan(flat, None, times, dt_max=1 / 100), 600, 5)
    assert np.array_equal(a, c)
    # and a moving weight does move the paths
    moving = Basket(b.names, b.models, b.markets, b.corr, b.vol_corr, StochasticMix(0.3, 1.0, -0.5, 2.0, 0.3))
    assert not np.array_equal(a, oracle_paths(native.plan(moving, None, times, dt_max=1 / 100), 600, 5))


@pytest.mark.parametrize("case", sorted(GOLD["cases"]))
def test_every_name_keeps_its_law_whatever_the_weight(case):
    basket, market = gold_basket(mix_of(case))
    times = GOLD["times"]
    out = oracle_paths(native.plan(basket, None, times), 1 << 15, 3)
    for a, vol in enumerate(GOLD["vols"]):
        for col in (25, 51):
            T = times[col]
            s = out[a, col]
            half = len(s) // 2
            F = float(market.F(T))
            for K in (0.85 * F, F, 1.15 * F):
                pay = np.maximum(s - K, 0.0)
                pay = 0.5 * (pay[:half] + pay[half:])
                se = pay.std(ddof=1) / np.sqrt(half)
                assert abs(pay.mean() - float(black_price(F, K, T, vol, 1.0, "call"))) < 4 * se


def test_a_deterministic_weight_sets_each_step_s_correlation():
    theta, kappa, w0, rho = 0.9, 3.0, 0.3, 0.5
    basket, _ = gold_basket(StochasticMix(theta, 0.0, 0.0, kappa, w0), corr=rho)
    times = GOLD["times"]
    dt = times[0]
    out = oracle_paths(native.plan(basket, None, times), 1 << 15, 9)
    x = np.log(out / out[:, :1, :1])  # per path the log closes; increments below
    inc = np.diff(np.concatenate([np.zeros((2, 1, x.shape[2])), x], axis=1), axis=1)
    c_low = low_matrix(np.array([[1.0, rho], [rho, 1.0]]), w0)[0, 1]
    for i in (0, 10, 51):
        w = theta + (w0 - theta) * (1.0 - kappa * dt) ** i
        want = w + (1.0 - w) * c_low
        got = np.corrcoef(inc[0, i], inc[1, i])[0, 1]
        assert abs(got - want) < 4 * (1 - want**2) / np.sqrt(inc.shape[2]), (i, got, want)
    assert want > 0.9 and abs(np.corrcoef(inc[0, 0], inc[1, 0])[0, 1] - rho) < 0.01


@pytest.mark.parametrize("case", sorted(GOLD["cases"]))
def test_the_reference_basket_simulator_within_monte_carlo_error(case):
    basket, _ = gold_basket(mix_of(case))
    out = oracle_paths(native.plan(basket, None, GOLD["times"]), 1 << 17, 7)
    half = out.shape[2] // 2
    for key, pay in payoffs(out[0].T / GOLD["spot"], out[1].T / GOLD["spot"]).items():
        d = 0.5 * (pay[:half] + pay[half:])
        mean, se = d.mean(), d.std(ddof=1) / np.sqrt(half)
        want, want_se = GOLD["cases"][case]["values"][key]
        assert abs(mean - want) < 4 * np.hypot(se, want_se), (key, mean, want, (mean - want) / np.hypot(se, want_se))
        ratio = se * np.sqrt(half) / (want_se * np.sqrt(GOLD["n_paths"] // 2))
        assert abs(ratio - 1.0) < 0.06, (key, ratio)
