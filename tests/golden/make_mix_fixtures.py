"""Writes tests/golden/stochastic_mix.json: statistical vectors of the reference's NumPy basket simulator
(packages/structured/src/optlab/structured/multiasset.py:77-131, simulate_basket) under the stochastic mixing weight
(localcorr.py:87-114, StochasticMix). The reference's compiled engine has no such mixer, so nothing else fixes the
weight's dynamics: these are basket payoffs of two Black-Scholes names observed weekly for a year (the weekly dates
are the simulation grid), their means and standard errors over antithetic pairs, which the path engine's mixer 2
must reproduce within Monte Carlo error.

Run where the reference is mounted (it is read from /root/reference, not copied):
    python tests/golden/make_mix_fixtures.py
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/packages")
OUT = Path(__file__).resolve().parent / "stochastic_mix.json"

NAMES = ("AAA", "BBB")
VOLS = (0.2, 0.3)
RHO = 0.7
SPOT, R, Q = 100.0, 0.03, 0.01
TIMES = [k / 52 for k in range(1, 53)]
N_PATHS = 1 << 19
SEED = 20261203
CASES = {  # StochasticMix(theta, nu, rho, kappa, w0)
    "skew_down": (0.5, 1.0, -0.8, 2.0, 0.3),
    "fast_up": (0.8, 0.6, 0.6, 4.0, 0.3),
}


def payoffs(a, b):
    """Per path: the payoffs the fixture holds, on the two names' closes in units of the spot (a, b: paths x dates)."""
    bt, wt = 0.5 * (a[:, -1] + b[:, -1]), np.minimum(a[:, -1], b[:, -1])
    bh = 0.5 * (a[:, 25] + b[:, 25])  # week 26
    # each path's realized correlation of the weekly log returns: its spread over paths is the weight's variance
    ra, rb = np.diff(np.log(a), axis=1, prepend=0.0), np.diff(np.log(b), axis=1, prepend=0.0)
    rc = (ra * rb).sum(axis=1) / np.sqrt((ra * ra).sum(axis=1) * (rb * rb).sum(axis=1))
    return {
        "basket_put_0.9": np.maximum(0.9 - bt, 0.0),
        "basket_put_1.0": np.maximum(1.0 - bt, 0.0),
        "basket_call_1.0": np.maximum(bt - 1.0, 0.0),
        "basket_call_1.1": np.maximum(bt - 1.1, 0.0),
        "worst_of_put_1.0": np.maximum(1.0 - wt, 0.0),
        "best_of_call_1.0": np.maximum(np.maximum(a[:, -1], b[:, -1]) - 1.0, 0.0),
        "basket_put_1.0_week26": np.maximum(1.0 - bh, 0.0),
        "spread_sq": (a[:, -1] - b[:, -1]) ** 2,
        "realized_corr": rc,
        "realized_corr_dev_sq": (rc - 0.8) ** 2,  # centred near the mean: the weight's variance, little noise
    }


def main() -> None:
    for pkg in ("core", "structured"):
        sys.path.insert(0, str(REF / pkg / "src"))
    from optlab.core.market import Market
    from optlab.core.models import BlackScholesModel
    from optlab.structured.localcorr import StochasticMix
    from optlab.structured.multiasset import Basket, simulate_basket

    markets = {n: Market.flat(SPOT, R, Q) for n in NAMES}
    corr = np.array([[1.0, RHO], [RHO, 1.0]])
    out = {"names": NAMES, "vols": VOLS, "rho": RHO, "spot": SPOT, "r": R, "q": Q, "times": TIMES,
           "n_paths": N_PATHS, "seed": SEED, "cases": {}}
    for name, (theta, nu, rho, kappa, w0) in CASES.items():
        mixer = StochasticMix(theta, nu, rho, kappa, w0)
        basket = Basket(NAMES, {n: BlackScholesModel(v) for n, v in zip(NAMES, VOLS)}, markets, corr, "same", mixer)
        paths = simulate_basket(basket, TIMES, N_PATHS, SEED)
        a, b = paths["AAA"] / SPOT, paths["BBB"] / SPOT
        half = N_PATHS // 2
        values = {}
        for key, pay in payoffs(a, b).items():
            d = 0.5 * (pay[:half] + pay[half:])  # rows i and i + n / 2 are antithetic partners
            values[key] = [float(d.mean()), float(d.std(ddof=1) / np.sqrt(half))]
        out["cases"][name] = {"mixer": {"theta": theta, "nu": nu, "rho": rho, "kappa": kappa, "w0": w0},
                              "values": values}
        print(name, {k: round(v[0], 5) for k, v in values.items()}, flush=True)
    OUT.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
