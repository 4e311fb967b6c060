"""Local-correlation calibration on the engine (GPU): every model price of the Gauss-Newton fit is a basket simulated
by the HIP path engine (local vol per name, the local-correlation mixer); in FP64 the fitted level, down and up
equal the oracle restatement's fit on the oracle's paths of the same plans and the same pseudo-quotes
(localcorr.py:153-213); the FP32 fit lands within the prices' noise of it."""

from __future__ import annotations

import sys
from functools import lru_cache

import numpy as np
import pytest
from _support import ROOT, oracle_paths

sys.path.insert(0, str(ROOT / "oracle"))
import localcorr_oracle as lco  # noqa: E402

from optlab import native  # noqa: E402
from optlab.core.engines import seed64  # noqa: E402
from optlab.core.market import Market  # noqa: E402
from optlab.core.ssvi import SSVI  # noqa: E402
from optlab.structured import localcorr as lc  # noqa: E402

pytestmark = pytest.mark.gpu

PILLARS = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0])
T, N_PATHS, SEED = 1.0, 8192, 11


@lru_cache(maxsize=1)
def basket():
    surfaces = [SSVI.power_law(r, 1.2, 0.5, PILLARS, v * PILLARS) for v, r in ((0.04, -0.6), (0.05, -0.7), (0.03, -0.5))]
    markets = [Market.flat(100.0, 0.03, 0.0)] * 3
    corr = np.full((3, 3), 0.5) + 0.5 * np.eye(3)
    return surfaces, markets, corr, lc.targets(surfaces, markets, corr, T)


def oracle_prices(p, seed=SEED):
    surfaces, markets, corr, wanted = basket()
    mixer = lc.LocalCorrelation(lc.basket_sigma(surfaces, corr, T), *p)
    plan = native.plan(lc.basket_model(surfaces, markets, corr, mixer), None, [T], dt_max=1 / 50)
    out = oracle_paths(plan, N_PATHS // 2, seed64(seed))
    x = out[:, 0, :].T / np.array([m.spot for m in markets])
    return lco.quotes(x, wanted)


def test_fp64_fit_is_the_oracle_fit_on_the_same_paths():
    surfaces, markets, corr, wanted = basket()
    got = lc.calibrate(surfaces, markets, corr, T, N_PATHS, SEED, wanted=wanted, backend="fp64")
    # one model price first: the engine's lines are the oracle's on the same draws
    mixer = lc.LocalCorrelation(lc.basket_sigma(surfaces, corr, T), 0.05, 0.3, 0.2)
    f, se = lc.model_quotes(lc.basket_model(surfaces, markets, corr, mixer), wanted, T, N_PATHS, SEED, "fp64")
    f0, se0 = oracle_prices((0.05, 0.3, 0.2))
    np.testing.assert_allclose(f, f0, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(se, se0, rtol=1e-8)
    want, want_err = lco.fit(oracle_prices, [r["price"] for r in wanted])
    m = got["mixer"]
    np.testing.assert_allclose([m.level, m.down, m.up], want, rtol=1e-7, atol=1e-9)
    assert got["rmse"] == pytest.approx(want_err, rel=1e-7, abs=1e-10)
    assert len(got["lines"]) == 3 and got["form"] == "local"
    # the pseudo-quotes assume more correlation below the money than above: the fit leans the same way
    assert m.down > m.up


def test_fp32_fit_prices_within_the_noise_of_fp64():
    surfaces, markets, corr, wanted = basket()
    g64 = lc.calibrate(surfaces, markets, corr, T, N_PATHS, SEED, wanted=wanted, backend="fp64")
    g32 = lc.calibrate(surfaces, markets, corr, T, N_PATHS, SEED, wanted=wanted, backend="fp32")
    for a, b in zip(g64["lines"], g32["lines"]):
        assert abs(a["model"] - b["model"]) < 2 * a["se"], (a, b)
