"""The CPU oracle pinned before it is trusted: Random123's known answers for Philox4x32-10, the reference's golden
path and leverage vectors (recorded by the reference's compiled CPU engine), the compiled reference itself where it
was built here (bit for bit), closed forms and Fourier prices. A plausible slip -- a swapped Philox word, the wrong
stream of the variance residuals, a dropped drift term, a transposed correlation root -- fails one of these."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
from _support import REF, calls, digest_samples, fixture, oracle, oracle_paths, reference, simulated_context

from optlab import native
from optlab.core.blackscholes import black_price, bs_price, down_out_call_bgk, geometric_asian
from optlab.core.heston import heston_price
from optlab.core.market import Market
from optlab.core.models import BlackScholesModel, HestonModel, HestonParams, LocalVolModel
from optlab.native import _abi

# The compiled reference is looked for by its file only: importing this module (the GPU tests do) must not load it.
needs_ref = pytest.mark.skipif(not REF.exists(), reason="reference CPU engine not built (oracle/build_ref.sh)")

# Random123 kat_vectors, philox4x32 10 rounds: (counter, key) -> output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("counter,key,want", KAT)
def test_philox_known_answers(counter, key, want):
    out = np.zeros(4, np.uint32)
    oracle().oracle_philox(*counter, *key, 1, out.ctypes.data)
    assert tuple(int(v) for v in out) == want


@needs_ref
def test_philox_and_normals_equal_the_compiled_reference():
    a, b = np.zeros(4 * 4096, np.uint32), np.zeros(4 * 4096, np.uint32)
    oracle().oracle_philox(7, 1, 3, 0x80000000, 42, 9, 4096, a.ctypes.data)
    reference().ref_philox(7, 1, 3, 0x80000000, 42, 9, 4096, b.ctypes.data)
    assert np.array_equal(a, b)
    for tag in (0, 0x80000000):
        x, y = np.zeros(300 * 11), np.zeros(300 * 11)
        oracle().oracle_normals(2**33 + 5, 300, 11, tag, 2**40 + 3, x.ctypes.data)
        reference().ref_normals(2**33 + 5, 300, 11, tag, 2**40 + 3, y.ctypes.data)
        assert np.array_equal(x, y)


def test_normals_are_standard_and_the_streams_differ():
    n = 200_000
    z = np.zeros(n * 4)
    oracle().oracle_normals(0, n, 4, 0, 42, z.ctypes.data)
    aux = np.zeros(n * 4)
    oracle().oracle_normals(0, n, 4, 0x80000000, 42, aux.ctypes.data)
    for s in (z, aux):
        assert abs(s.mean()) < 4 / math.sqrt(s.size)
        assert abs(s.var() - 1) < 4 * math.sqrt(2 / s.size)
        assert np.abs(s).max() < 6.67  # |z| <= sqrt(-2 ln 2^-33)
    assert abs(np.corrcoef(z, aux)[0, 1]) < 4 / math.sqrt(z.size)
    # the first normal of a pair is r cos(2 pi v) of its first Philox block: a radius and an angle
    w = np.zeros(4, np.uint32)
    oracle().oracle_philox(0, 0, 0, 0, 42, 0, 1, w.ctypes.data)
    u, v = (float(w[0]) + 0.5) * 2**-32, (float(w[1]) + 0.5) * 2**-32
    assert z[0] == pytest.approx(math.sqrt(-2 * math.log(u)) * math.cos(2 * math.pi * v), rel=1e-14, abs=1e-15)
    assert z[1] == pytest.approx(math.sqrt(-2 * math.log(u)) * math.sin(2 * math.pi * v), rel=1e-14, abs=1e-15)


def _golden_models():
    market, surface, heston = simulated_context()
    return market, {
        "bs": BlackScholesModel(float(surface.iv(0.0, 1.0))),
        "lv": LocalVolModel(surface),
        "heston": HestonModel(heston),
    }


@pytest.mark.parametrize("name", ["bs", "lv", "heston"])
def test_oracle_reproduces_the_reference_golden_paths(name):
    """native_cpu.json paths/<name>: 20 000 paths at [0.25, 0.5, 1] on seed 7, recorded by the reference's
    compiled CPU engine (golden.py:409-420). Across compilers: rtol 1e-12 (golden.py:705)."""
    market, models = _golden_models()
    p = native.plan(models[name], market, [0.25, 0.5, 1.0])
    S = oracle_paths(p, 10_000, 7)[0].T
    samples, total = digest_samples(S)
    g = fixture()["native_cpu"]["paths"][name]
    assert list(S.shape) == g["shape"]
    np.testing.assert_allclose(samples, g["samples"], rtol=1e-12, atol=0)
    assert total == pytest.approx(g["sum"], rel=1e-13)


def test_oracle_reproduces_the_reference_golden_leverage():
    """native_cpu.json leverage: 20 000 particles, mix 0.5, rng 5, T 1 (golden.py:452-454), with the bin sums
    partitioned like the 12 threads of the recording machine."""
    market, surface, heston = simulated_context()
    rq, arrays, t0 = native.leverage_request(heston, surface, 1.0, 0.5, 20_000, 1 / 100, 1 / 2000, 81, 1.0, 5,
                                             (0.05, 8.0), "nadaraya_watson", False)
    assert oracle().oracle_leverage(C.byref(rq), 0, 12) == 0
    g = fixture()["native_cpu"]["leverage"]
    np.testing.assert_array_equal(t0, g["times"])
    for key in ("x_grids", "values"):
        samples, total = digest_samples(arrays[key])
        np.testing.assert_allclose(samples, g[key]["samples"], rtol=1e-12, atol=1e-15)
        assert total == pytest.approx(g[key]["sum"], rel=1e-12)


def _leverage_rq(local_linear=False, positive=False, n=6000, T=0.5):
    market, surface, heston = simulated_context()
    return native.leverage_request(heston, surface, T, 0.5, n, 1 / 100, 1 / 2000, 81, 1.0, 5, (0.05, 8.0),
                                   "local_linear" if local_linear else "nadaraya_watson", positive)


@needs_ref
@pytest.mark.parametrize("local_linear,positive", [(False, False), (True, True)])
def test_oracle_leverage_equals_the_compiled_reference(local_linear, positive):
    rq, a, _ = _leverage_rq(local_linear, positive)
    assert oracle().oracle_leverage(C.byref(rq), 4, 4) == 0
    rq2, b, _ = _leverage_rq(local_linear, positive)
    assert reference().ref_leverage(C.byref(rq2), 4) == 0
    assert np.array_equal(a["x_grids"], b["x_grids"]) and np.array_equal(a["values"], b["values"])


def _basket(mixer=False, slv=False):
    from optlab.core.models import Leverage, SLVModel
    from optlab.structured.localcorr import LocalCorrelation
    from optlab.structured.multiasset import Basket

    market, surface, heston = simulated_context()
    names = ("SPX", "AAA", "BBB")
    corr = np.array([[1.0, 0.8, 0.6], [0.8, 1.0, 0.7], [0.6, 0.7, 1.0]])
    models = {"SPX": LocalVolModel(surface), "AAA": HestonModel(heston), "BBB": BlackScholesModel(0.25)}
    if slv:
        t = np.array([0.0, 0.25, 0.5])
        xg = np.tile(np.linspace(-1, 1, 9), (3, 1))
        lev = Leverage(t, xg, 1.0 + 0.3 * np.abs(xg) + 0.1 * t[:, None])
        models["AAA"] = SLVModel(heston, lev, 0.5)
    mx = LocalCorrelation(sigma=0.15, level=0.0, down=0.3, up=0.1) if mixer else None
    return Basket(names, models, {n: market for n in names}, corr, mixer=mx), market


@needs_ref
@pytest.mark.parametrize("case", ["bs", "lv", "heston", "basket", "basket_slv_localcorr"])
def test_oracle_paths_equal_the_compiled_reference(case):
    """Bit for bit, on every model, baskets, SLV's aux stream, local correlation, kept columns and two windows."""
    if case.startswith("basket"):
        model, market = _basket(mixer="localcorr" in case, slv="slv" in case)
    else:
        market, models = _golden_models()
        model = models[case]
    times = [0.1, 0.25, 0.3, 0.5]
    p = native.plan(model, market, times, dt_max=1 / 100)
    kw = dict(keep=[1, 3], windows=[(0, 3), (1, 2)])
    a = oracle_paths(p, 700, 11, first_pair=5, **kw)
    b = oracle_paths(p, 700, 11, first_pair=5, lib=reference(), **kw)
    assert np.array_equal(a, b)


def test_batches_do_not_change_the_paths():
    """requests.hpp:37, test_native.py:138-161: a path depends on (seed, pair index) only."""
    market, models = _golden_models()
    p = native.plan(models["lv"], market, [0.5, 1.0], dt_max=1 / 50)
    whole = oracle_paths(p, 1000, 3)
    first = oracle_paths(p, 400, 3)
    rest = oracle_paths(p, 600, 3, first_pair=400)
    for leg in range(2):
        np.testing.assert_array_equal(whole[:, :, leg * 1000:leg * 1000 + 400], first[:, :, leg * 400:leg * 400 + 400])
        np.testing.assert_array_equal(whole[:, :, leg * 1000 + 400:(leg + 1) * 1000], rest[:, :, leg * 600:(leg + 1) * 600])


def test_windows_equal_the_min_and_max_of_the_full_columns():
    """test_paths.cpp:114-139: the window extremes reduced while simulating equal those of the stored columns."""
    market, models = _golden_models()
    times = np.linspace(0.02, 0.5, 25)
    p = native.plan(models["heston"], market, times, dt_max=1 / 100)
    out = oracle_paths(p, 500, 4, windows=[(0, 24), (5, 12)])[0]
    np.testing.assert_array_equal(out[25], out[:25].min(axis=0))
    np.testing.assert_array_equal(out[26], out[:25].max(axis=0))
    np.testing.assert_array_equal(out[27], out[5:13].min(axis=0))
    np.testing.assert_array_equal(out[28], out[5:13].max(axis=0))


def test_black_scholes_paths_match_the_formula_and_the_forward():
    m = Market.flat(100.0, 0.04, 0.01)
    p = native.plan(BlackScholesModel(0.2), m, [1.0])
    S = oracle_paths(p, 100_000, 1)[0, 0]
    K = np.array([80.0, 100.0, 120.0])
    price, se = calls(S, float(m.D(1.0)), K)
    exact = black_price(float(m.F(1.0)), K, 1.0, 0.2, float(m.D(1.0)))
    assert np.all(np.abs(price - exact) < 4 * se)
    assert abs(S.mean() / float(m.F(1.0)) - 1) < 4 * S.std() / np.sqrt(S.size) / float(m.F(1.0))


def test_heston_euler_paths_keep_the_forward_and_match_fourier():
    market, _, heston = simulated_context()
    p = native.plan(HestonModel(heston), market, [1.0])
    S = oracle_paths(p, 100_000, 3)[0, 0]
    F, D = float(market.F(1.0)), float(market.D(1.0))
    K = np.array([90.0, 100.0, 110.0])
    price, se = calls(S, D, K)
    exact = np.array([heston_price(F, k, 1.0, heston, D) for k in K])
    assert np.all(np.abs(price - exact) < 4 * se + 0.03)  # full-truncation Euler bias (test_native.py:75-82)


def _request(model, product, legs, strikes, n_steps, n_paths=200_000, H=0.0, seed=20261029, **over):
    rq = _abi.Request()
    rq.model, rq.product = model, product
    for k in range(3):
        rq.legs[k], rq.strikes[k] = legs[k], strikes[k]
    rq.H, rq.n_steps = H, n_steps
    rq.S0, rq.r, rq.q, rq.T, rq.sigma = 100.0, 0.04, 0.01, 1.0, 0.20
    h = HestonParams(0.022846642482989885, 3.0892635624407627, 0.05372458999932075, 1.307745170926575,
                     -0.6879606734541107)  # params.hpp:17-21
    rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho = h.v0, h.kappa, h.theta, h.xi, h.rho
    rq.n_paths, rq.seed = n_paths, seed
    for k, v in over.items():
        setattr(rq, k, v)
    return rq


def _oracle_price(rq):
    out = _abi.Result()
    assert oracle().oracle_price(C.byref(rq), 0, C.byref(out)) == 0, oracle().oracle_last_error()
    return out


FIXED_CASES = [
    (0, 0, (1, 2, 3), (100, 100, 100), 1),
    (0, 1, (1, 0, 0), (100, 0, 0), 12),
    (0, 2, (1, 0, 0), (100, 0, 0), 52),
    (1, 0, (1, 1, 2), (100, 119, 80), 52),
    (1, 2, (1, 0, 0), (100, 0, 0), 52),
    (1, 1, (2, 0, 0), (95, 0, 0), 20),
]


@needs_ref
@pytest.mark.parametrize("model,product,legs,strikes,n_steps", FIXED_CASES)
def test_fixed_products_equal_the_reference_schemes_on_philox_draws(model, product, legs, strikes, n_steps):
    """The reference's own make_consts / qe_step / path_unit (schemes.hpp) compiled here and driven by Philox
    stream 0 of each pair: the oracle's restatement agrees to the last bits of the sums' order."""
    rq = _request(model, product, legs, strikes, n_steps, n_paths=20_000, H=80.0)
    out = _oracle_price(rq)
    s, s2 = np.zeros(4), np.zeros(4)
    assert reference().ref_fixed_sums(C.byref(rq), s.ctypes.data, s2.ctypes.data) == 0
    n = rq.n_paths // 2
    for k in (0, 1, 3) if product == 1 else range(4):  # out[2] of the Asian is this build's control variate
        mean = s[k] / n
        assert out.price[k] == pytest.approx(mean, rel=1e-13, abs=1e-13)
        se = math.sqrt(max(s2[k] / n - mean * mean, 0.0) / (n - 1))
        assert out.se[k] == pytest.approx(se, rel=1e-9, abs=1e-15)


def test_fixed_products_against_the_validation_table():
    """test_prices.cpp:51-88: |p - ref| <= 4 sqrt(se^2 + se_ref^2) + bias, on the oracle's Philox draws."""
    mg = 100.0 * math.exp(0.03)
    rows = [
        (_request(0, 0, (1, 2, 3), (100, 100, 100), 1, 1_000_000), [(0, 9.319738, 0, 0), (1, 6.393699, 0, 0),
                                                                      (2, 2.926039, 0, 0), (3, mg, 0, 0)]),
        (_request(0, 1, (1, 0, 0), (100, 0, 0), 12, 400_000), [(0, 5.575692, 0.000114, 0), (1, 5.381250, 0, 0)]),
        (_request(0, 2, (1, 0, 0), (100, 0, 0), 252, 100_000, H=80.0), [(0, 9.245769, 0.007422, 0),
                                                                         (1, 9.319738, 0, 0)]),
        (_request(1, 0, (1, 1, 2), (100, 119, 80), 252, 100_000), [(0, 8.403656, 0, 0.02), (1, 1.010290, 0, 0.02),
                                                                    (2, 1.737996, 0, 0.02), (3, mg, 0, 0)]),
        (_request(1, 2, (1, 0, 0), (100, 0, 0), 252, 100_000, H=80.0), [(0, 8.0927, 0.0092, 0),
                                                                         (1, 8.4153, 0.0131, 0)]),
    ]
    for rq, checks in rows:
        out = _oracle_price(rq)
        for k, ref, se_ref, bias in checks:
            assert abs(out.price[k] - ref) <= 4 * math.hypot(out.se[k], se_ref) + bias, (rq.model, rq.product, k)


def test_asian_control_variate_is_unbiased_and_tighter():
    rq = _request(0, 1, (1, 0, 0), (100, 0, 0), 12, 200_000)
    out = _oracle_price(rq)
    assert out.se[2] < out.se[0] / 5
    assert abs(out.price[2] - 5.575692) < 4 * math.hypot(out.se[2], 0.000114)


def test_up_and_out_matches_the_bgk_closed_form():
    rq = _request(0, 3, (1, 0, 0), (100, 0, 0), 252, 400_000, H=130.0)
    out = _oracle_price(rq)
    from optlab.core.blackscholes import up_out_call_bgk

    exact = up_out_call_bgk(100.0, 100.0, 130.0, 1.0, 0.04, 0.01, 0.20, 252)
    assert abs(out.price[0] - exact) < 4 * out.se[0] + 0.01
    assert out.price[1] > out.price[0]


def test_closed_forms_match_the_reference_table():
    """test_prices.cpp:106-122: the closed forms against R1, R2, R3, R5 to 1e-6."""
    assert bs_price(100, 100, 1, 0.04, 0.01, 0.2, True) == pytest.approx(9.319738, abs=1e-6)
    assert bs_price(100, 100, 1, 0.04, 0.01, 0.2, False) == pytest.approx(6.393699, abs=1e-6)
    assert geometric_asian(100, 100, 1, 0.04, 0.01, 0.2, 12, True) == pytest.approx(5.381250, abs=1e-6)
    assert oracle().oracle_geometric_asian(100, 100, 1, 0.04, 0.01, 0.2, 12, 1) == pytest.approx(5.381250, abs=1e-6)
    assert down_out_call_bgk(100, 100, 80, 1, 0.04, 0.01, 0.2, 252) == pytest.approx(9.243397, abs=1e-6)
    lab = fixture()["labjs"]
    for (S, K, T, r, q, v, cp), want in zip(lab["geo_asian_cases"], lab["geo_asian"]):
        assert geometric_asian(S, K, T, r, q, v, 12, cp == "call") == pytest.approx(want, rel=1e-12)


def test_heston_fourier_matches_lab_js():
    lab = fixture()["labjs"]
    h = HestonParams(**lab["heston_params"])
    m = Market.flat(100.0, 0.04, 0.01)
    for (K, T, cp), want in zip(lab["heston_fourier_cases"], lab["heston_fourier"]):
        got = heston_price(float(m.F(T)), float(K), T, h, float(m.D(T)), cp)
        assert got == pytest.approx(want, abs=2e-7)


def _strip(model=1, n_paths=40_000, put=0, n_steps=104, mats=(13, 26, 52, 104), strikes=(80, 90, 100, 110, 120)):
    strikes = np.array(strikes, dtype=float)
    mats = np.array(mats, dtype=np.int32)
    rq = _abi.StripRequest()
    rq.model, rq.put, rq.n_steps = model, put, n_steps
    rq.S0, rq.r, rq.q, rq.T, rq.sigma = 100.0, 0.03, 0.0, 2.0, 0.2
    rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho, rq.psi_c = 0.04, 1.5, 0.04, 0.6, -0.7, 1.5
    rq.n_strikes, rq.strikes, rq.n_mats, rq.mat_step = len(strikes), _abi.ptr(strikes), len(mats), _abi.ptr(mats, _abi._PI)
    rq.n_paths, rq.seed = n_paths, 42
    price, se = np.zeros((len(mats), len(strikes) + 1)), np.zeros((len(mats), len(strikes) + 1))
    rq.price, rq.se = _abi.ptr(price), _abi.ptr(se)
    rq._keepalive = (strikes, mats, price, se)  # the request points into these arrays
    return rq, (strikes, mats, price, se)


def test_strip_at_the_last_maturity_equals_the_fixed_european():
    """The strip's last column on the same pairs is the fixed European of the same legs, bit for bit."""
    rq, (strikes, mats, price, se) = _strip(model=1, n_steps=104, mats=(52, 104), strikes=(90, 100, 110))
    assert oracle().oracle_strip(C.byref(rq), 0) == 0
    f = _request(1, 0, (1, 1, 1), (90, 100, 110), 104, 40_000, seed=42, S0=100.0, r=0.03, q=0.0, T=2.0, v0=0.04,
                 kappa=1.5, theta=0.04, xi=0.6, rho=-0.7)
    out = _oracle_price(f)
    for k in range(3):
        assert price[1, k] == out.price[k] and se[1, k] == out.se[k]
    assert price[1, 3] == out.price[3]


def test_strip_matches_fourier_and_the_forward():
    rq, (strikes, mats, price, se) = _strip(n_paths=200_000, n_steps=504, mats=(126, 504), strikes=(70, 100, 130))
    assert oracle().oracle_strip(C.byref(rq), 0) == 0
    h = HestonParams(0.04, 1.5, 0.04, 0.6, -0.7)
    for j, T in enumerate((0.5, 2.0)):
        F, D = 100 * math.exp(0.03 * T), math.exp(-0.03 * T)
        for k, K in enumerate(strikes):
            exact = heston_price(F, K, T, h, D)
            assert abs(price[j, k] - exact) < 4 * se[j, k] + 0.02
        assert abs(price[j, 3] - F) < 4 * se[j, 3]


def test_oracle_rejects_bad_requests():
    rq = _request(2, 0, (1, 0, 0), (100, 0, 0), 1)
    out = _abi.Result()
    assert oracle().oracle_price(C.byref(rq), 0, C.byref(out)) == _abi.EINVAL
    assert b"model" in oracle().oracle_last_error()
