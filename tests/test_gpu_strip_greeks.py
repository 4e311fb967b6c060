"""The strip's pathwise sensitivities on the GPU (optlab_strip_greeks): d price / d S0 and d price / d v0 through
every Heston QE-M step (d / d sigma under Black-Scholes).

Bars: FP64 equals central differences of the oracle's strip on the same draws (rel 1e-6) -- over many steps in the
quadratic branch, over one step in each branch (every path takes the exponential branch at psi ~ 7), and under
Black-Scholes; FP32 within 0.1 SE of FP64 on the same draws, and within 4 SE of FP32 bump-and-revalue on common
random numbers (paired differences over 16 seeds) where paths switch branches; against what the mathematics fixes:
Black-Scholes delta and vega in closed form, d E[S_t] / d v0 = 0 (the scheme's martingale correction), and the
v0-derivative of the Fourier Heston price."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
from _support import oracle
from scipy.stats import norm
from test_oracle import _strip

from optlab.core.heston import heston_price
from optlab.core.models import HestonParams
from optlab.native import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = _abi.load()  # raises when the library is not built: a GPU test never passes without the HIP engine
    if lib.optlab_gpu_count() < 1:
        pytest.fail("no HIP device visible")
    return lib


def _greeks(lib, rq, backend):
    shape = (rq.n_mats, rq.n_strikes + 1)
    d, dse, g, gse = (np.zeros(shape) for _ in range(4))
    code = lib.optlab_strip_greeks(C.byref(rq), backend, _abi.ptr(d), _abi.ptr(dse), _abi.ptr(g), _abi.ptr(gse))
    assert code == 0, lib.optlab_last_error()
    return d, dse, g, gse


def _oracle_prices(rq, **bump):
    saved = {k: getattr(rq, k) for k in bump}
    for k, v in bump.items():
        setattr(rq, k, v)
    assert oracle().oracle_strip(C.byref(rq), 0) == 0
    for k, v in saved.items():
        setattr(rq, k, v)
    return rq._keepalive[2].copy()


def _set(rq, **fields):
    for k, v in fields.items():
        setattr(rq, k, v)
    return rq


CASES = {
    # many steps, psi far below psi_c on every path: the quadratic branch and the tangent's passage between steps
    "heston_quadratic_64_steps": dict(n_steps=64, mats=(16, 32, 64), fields=dict(T=0.5, xi=0.3, kappa=2.0)),
    # one step from v0 = 0.01 with xi = 1: psi = 7.3 on every path, the exponential branch alone
    "heston_exponential_one_step": dict(n_steps=1, mats=(1,), fields=dict(T=0.25, xi=1.0, v0=0.01, kappa=1.5)),
    # one step in the quadratic branch (psi = 0.02)
    "heston_quadratic_one_step": dict(n_steps=1, mats=(1,), fields=dict(T=0.25, xi=0.3, v0=0.04)),
    "black_scholes": dict(model=0, n_steps=52, mats=(13, 52), fields=dict(T=1.0, sigma=0.25)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fp64_sensitivities_are_differences_of_the_oracle(lib, case):
    c = CASES[case]
    model = c.get("model", 1)
    rq, _ = _strip(model=model, n_paths=2048, n_steps=c["n_steps"], mats=c["mats"])
    _set(rq, **c["fields"])
    d, _, g, _ = _greeks(lib, rq, 0)
    h = 1e-6
    up, dn = _oracle_prices(rq, S0=rq.S0 * (1 + h)), _oracle_prices(rq, S0=rq.S0 * (1 - h))
    fd = (up - dn) / (2 * h * rq.S0)
    np.testing.assert_allclose(d, fd, rtol=1e-6, atol=1e-9)
    name = "v0" if model == 1 else "sigma"
    x = getattr(rq, name)
    h = 1e-7
    up, dn = _oracle_prices(rq, **{name: x + h}), _oracle_prices(rq, **{name: x - h})
    fd = (up - dn) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())
    assert np.abs(g[:, 2]).min() > 1.0  # at the money the parameter moves every price: nothing is compared at zero


def test_fp32_sensitivities_track_fp64_on_the_same_draws(lib):
    rq, _ = _strip(n_paths=1 << 18, n_steps=104, mats=(26, 104))
    d64, dse, g64, gse = _greeks(lib, rq, 0)
    d32, _, g32, _ = _greeks(lib, rq, 1)
    assert np.all(np.abs(d32 - d64) <= 0.1 * dse + 1e-6)
    assert np.all(np.abs(g32 - g64) <= 0.1 * gse + 1e-5)


def test_fp32_sensitivities_agree_with_bumps_on_common_random_numbers(lib):
    """xi = 0.6: about one leg-step in six takes the exponential branch, so bumped paths switch branches and the
    difference quotients carry the scheme's switches that a pathwise derivative does not; paired over 16 seeds the
    two agree within 4 standard errors of the paired mean. Bumps: 0.1 % of spot (a central difference's bias grows
    as h^2 and its spread as sqrt(h): at 1 % the bias of the at-the-money delta, 6e-4, is 4 of these standard
    errors) and 0.001 of v0."""
    diffs_d, diffs_g = [], []
    for seed in range(16):
        rq, (_, _, price, _) = _strip(n_paths=1 << 15, n_steps=104, mats=(26, 104))
        rq.seed = 1000 + seed
        d, _, g, _ = _greeks(lib, rq, 1)

        def bumped(**kw):
            saved = {k: getattr(rq, k) for k in kw}
            _set(rq, **kw)
            assert lib.optlab_strip(C.byref(rq), 1) == 0, lib.optlab_last_error()
            _set(rq, **saved)
            return price.copy()

        fd_d = (bumped(S0=100.1) - bumped(S0=99.9)) / 0.2
        fd_g = (bumped(v0=0.041) - bumped(v0=0.039)) / 0.002
        diffs_d.append(d - fd_d)
        diffs_g.append(g - fd_g)
    # the rounding of the FP32 prices themselves (7.6e-6 at 100, over a spot difference of 0.2: 4e-5) bounds how well
    # a difference quotient of them can agree with anything
    ok = []
    for diffs, rounding in ((np.array(diffs_d), 1e-4), (np.array(diffs_g), 1e-2)):
        mean, se = diffs.mean(axis=0), diffs.std(axis=0, ddof=1) / 4.0
        print("paired mean", mean.round(7).tolist(), "se", se.round(7).tolist())
        ok.append(bool(np.all(np.abs(mean) <= 4 * se + rounding)))
    assert all(ok), ok


def test_black_scholes_sensitivities_are_the_closed_forms(lib):
    rq, (strikes, _, _, _) = _strip(model=0, n_paths=1 << 20, n_steps=52, mats=(13, 52))
    _set(rq, T=1.0, sigma=0.25)
    d, dse, g, gse = _greeks(lib, rq, 1)
    S, r, sig = rq.S0, rq.r, rq.sigma
    for j, t in enumerate((0.25, 1.0)):
        for k, K in enumerate(strikes):
            d1 = (math.log(S / K) + (r + 0.5 * sig * sig) * t) / (sig * math.sqrt(t))
            assert abs(d[j, k] - norm.cdf(d1)) < 4 * dse[j, k] + 1e-5, (j, k)
            assert abs(g[j, k] - S * norm.pdf(d1) * math.sqrt(t)) < 4 * gse[j, k] + 1e-4, (j, k)
        # E[S_t] = S e^{r t}: its delta is e^{r t} and its sigma-derivative 0
        assert abs(d[j, -1] - math.exp(r * t)) < 4 * dse[j, -1]
        assert abs(g[j, -1]) < 4 * gse[j, -1]


def test_heston_v0_sensitivity_is_the_derivative_of_the_fourier_price(lib):
    """d price / d v0 against central differences of the Fourier price (within 4 SE plus the QE-M scheme's own
    bias, a fraction of a percent at 252 steps a year); d E[S_t] / d v0 within 4 SE of 0."""
    rq, (strikes, _, _, _) = _strip(n_paths=1 << 20, n_steps=252, mats=(126, 252), strikes=(80, 90, 100, 110, 120))
    _set(rq, T=1.0)
    _, _, g, gse = _greeks(lib, rq, 1)
    p = dict(v0=rq.v0, kappa=rq.kappa, theta=rq.theta, xi=rq.xi, rho=rq.rho)
    h = 1e-4
    for j, t in enumerate((0.5, 1.0)):
        F, D = rq.S0 * math.exp(rq.r * t), math.exp(-rq.r * t)
        for k, K in enumerate(strikes):
            up = heston_price(F, K, t, HestonParams(**{**p, "v0": p["v0"] + h}), D)
            dn = heston_price(F, K, t, HestonParams(**{**p, "v0": p["v0"] - h}), D)
            exact = (up - dn) / (2 * h)
            assert abs(g[j, k] - exact) < 4 * gse[j, k] + 0.01 * abs(exact), (j, k, g[j, k], exact)
        assert abs(g[j, -1]) < 4 * gse[j, -1] + 1e-6


def test_sensitivities_reject_bad_requests(lib):
    rq, _ = _strip(n_paths=1024)
    z = np.zeros((4, 6))
    rq.n_paths = 3
    assert lib.optlab_strip_greeks(C.byref(rq), 0, _abi.ptr(z), _abi.ptr(z), _abi.ptr(z), _abi.ptr(z)) == _abi.EINVAL
    rq.n_paths = 1024
    assert lib.optlab_strip_greeks(C.byref(rq), 0, _abi.ptr(z), None, _abi.ptr(z), _abi.ptr(z)) == _abi.EINVAL
    assert b"null output" in lib.optlab_last_error()
