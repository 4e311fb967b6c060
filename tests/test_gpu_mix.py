"""The stochastic mixing weight (mixer 2: localcorr.py:87-114 StochasticMix, an addition the reference's compiled
engine refuses) on the device: FP32 paths of two Black-Scholes names observed weekly against the reference's own
NumPy basket simulator (tests/golden/stochastic_mix.json) within Monte Carlo error, with the pair averages'
standard errors within 6 % of the reference's; a weight that cannot move gives the constant-matrix local
correlation's paths bit for bit in both precisions. FP64 paths equal the oracle's in
test_gpu_engine.py (case basket_slv_stochmix)."""

from __future__ import annotations

import numpy as np
import pytest
from test_oracle import _basket
from test_oracle_mix import GOLD, gold_basket, mix_of, payoffs

from optlab import native
from optlab.structured.localcorr import LocalCorrelation, StochasticMix
from optlab.structured.multiasset import Basket

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", sorted(GOLD["cases"]))
def test_fp32_mix_is_the_reference_basket_simulator_within_monte_carlo_error(case):
    basket, _ = gold_basket(mix_of(case))
    print(f"stochastic mix {case}: 2^19 pairs fp32", flush=True)
    spots, _, timing = native.run(native.plan(basket, None, GOLD["times"]), 1 << 19, native.seed64(11),
                                  backend="fp32")
    assert timing["precision"] == "fp32"
    half = spots[0].shape[0] // 2
    for key, pay in payoffs(spots[0] / GOLD["spot"], spots[1] / GOLD["spot"]).items():
        d = 0.5 * (pay[:half] + pay[half:])
        mean, se = d.mean(), d.std(ddof=1) / np.sqrt(half)
        want, want_se = GOLD["cases"][case]["values"][key]
        assert abs(mean - want) < 4 * np.hypot(se, want_se), (key, mean, want, (mean - want) / np.hypot(se, want_se))
        ratio = se * np.sqrt(half) / (want_se * np.sqrt(GOLD["n_paths"] // 2))
        assert abs(ratio - 1.0) < 0.06, (key, ratio)


@pytest.mark.parametrize("backend", ["fp64", "fp32"])
def test_a_still_weight_is_the_constant_matrix_bit_for_bit(backend):
    b, _ = _basket(mixer=True, slv=True)
    still = Basket(b.names, b.models, b.markets, b.corr, b.vol_corr, StochasticMix(0.3, 0.0, -0.5, 2.0, 0.3))
    flat = Basket(b.names, b.models, b.markets, b.corr, b.vol_corr, LocalCorrelation(0.15, 0.0, 0.0, 0.0, 0.3))
    times = [0.1, 0.25, 0.5]
    print(f"still weight {backend}", flush=True)
    a, _, _ = native.run(native.plan(still, None, times, dt_max=1 / 100), 3000, 5, backend=backend)
    c, _, _ = native.run(native.plan(flat, None, times, dt_max=1 / 100), 3000, 5, backend=backend)
    for x, y in zip(a, c):
        assert np.array_equal(x, y)
