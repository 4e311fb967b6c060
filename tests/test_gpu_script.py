"""Payoff scripts fused onto the device (GPU): the phoenix and the barrier reverse convertible of the preset list,
lowered to HIP, compiled by hipRTC and valued over the path engine's output in HBM, against the oracle's restatement
of the NumPy evaluator on the same engine paths brought to the host (FP64 paths: 1e-12; hard and smoothed
comparisons); FP32 paths within the FP64 price's standard error; a second pricing of a script reuses its compiled
kernel."""

from __future__ import annotations

import sys
from functools import lru_cache

import numpy as np
import pytest
from _support import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import script_oracle as so  # noqa: E402

from optlab.core.market import Market  # noqa: E402
from optlab.core.models import LocalVolModel  # noqa: E402
from optlab.core.ssvi import SSVI  # noqa: E402
from optlab.structured import script as sc  # noqa: E402
from optlab.structured.multiasset import Basket  # noqa: E402

pytestmark = pytest.mark.gpu

PILLARS = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0])
NAMES = ("A", "B", "C")


@lru_cache(maxsize=1)
def surfaces():
    return [SSVI.power_law(r, 1.2, 0.5, PILLARS, v * PILLARS) for v, r in ((0.03, -0.5), (0.04, -0.6), (0.05, -0.7))]


def pricer(which, smoothing):
    from optlab.native.script import ScriptPricer

    market = Market.flat(100.0, 0.03, 0.0)
    if which == "phoenix":
        basket = Basket(NAMES, {n: LocalVolModel(s) for n, s in zip(NAMES, surfaces())}, {n: market for n in NAMES},
                        np.full((3, 3), 0.6) + 0.4 * np.eye(3))
        return ScriptPricer(sc.phoenix_script(NAMES, 3.0, 0.25, daily_ki=756), basket, smoothing=smoothing), market
    text = sc.barrier_reverse_convertible_script("X", 1.0, daily_ki=252)
    return ScriptPricer(text, LocalVolModel(surfaces()[1]), market, smoothing=smoothing), market


@pytest.mark.parametrize("which,smoothing", [("phoenix", 0.0), ("phoenix", 0.005), ("brc", 0.0), ("brc", 0.005)])
def test_fp64_fused_payoff_is_the_evaluator_on_the_same_paths(which, smoothing):
    p, market = pricer(which, smoothing)
    n_paths, seed = 1 << 14, 2024
    got = p.price(n_paths, seed, "fp64")
    spots, wins = p.paths(n_paths, seed, "fp64")
    values = so.leg_values(p.program, None, smoothing, spots, wins, lambda t: float(market.D(t)))
    want, want_se = so.pair_mean(values)
    assert got["price"] == pytest.approx(want, rel=1e-12)
    assert got["se"] == pytest.approx(want_se, rel=1e-8)
    assert 0.5 < got["price"] < 1.2 and got["se"] > 0
    print(which, smoothing, got)


def test_fp32_paths_price_within_the_fp64_standard_error_and_the_kernel_is_cached():
    p, _ = pricer("phoenix", 0.005)
    g64 = p.price(1 << 18, 7, "fp64")
    g32 = p.price(1 << 18, 7, "fp32")
    assert abs(g32["price"] - g64["price"]) < g64["se"]
    assert g32["compile_ms"] == 0.0  # the same source: compiled once in the process
    again = p.price(1 << 18, 7, "fp64")
    assert again["price"] == g64["price"] and again["se"] == g64["se"]
