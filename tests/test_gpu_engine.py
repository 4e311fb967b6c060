"""The HIP engine against the CPU oracle on the same Philox draws (GPU only).

Bars: the integer stream bit-exact; the FP64 mode within 1e-12 relative of the oracle (the reference CPU engine's
arithmetic) on every output; the FP32 mode within the reference's own float-against-double bounds
(test_native.py:95-111: median relative gap 1e-6, worst 1e-4 for local vol; Heston median 1e-5 with under 6 % of
paths parted by the truncation) and its prices within 0.5 standard errors of FP64 on the same normals
(test_prices.cpp:125-145)."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
from _support import digest_samples, fixture, oracle, oracle_paths, simulated_context
from test_oracle import FIXED_CASES, _basket, _golden_models, _oracle_price, _request, _strip

from optlab import native
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


def _torch():
    import torch

    return torch


def test_philox_integer_stream_is_bit_exact(lib):
    torch = _torch()
    n = 1 << 16
    for c in [(0, 0, 0, 0, 0, 0), (5, 1, 7, 0x80000000, 42, 3), (0xFFFF0000, 0xFFFFFFFF, 9, 1, 0xDEADBEEF, 77)]:
        dev = torch.empty(4 * n, dtype=torch.int32, device="cuda")
        assert lib.optlab_philox_dev(*c, n, C.c_void_p(dev.data_ptr()), None) == 0
        got = dev.cpu().numpy().view(np.uint32)
        want = np.zeros(4 * n, np.uint32)
        oracle().oracle_philox(*c, n, want.ctypes.data)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("tag", [0, 0x80000000])
def test_normals_match_the_oracle(lib, tag):
    torch = _torch()
    n_pairs, k = 20_000, 13
    want = np.zeros(n_pairs * k)
    oracle().oracle_normals(3, n_pairs, k, tag, 2**35 + 11, want.ctypes.data)
    # FP32: u and v rounded to 24 bits before the log and the angle, so |z| ~ 6 carries a few 1e-6 (absolute)
    for backend, tol, typical in ((0, 4e-15, 1e-16), (1, 2e-5, 2e-7)):
        dev = torch.empty(n_pairs * k, dtype=torch.float64, device="cuda")
        assert lib.optlab_normals_dev(3, n_pairs, k, tag, 2**35 + 11, backend, C.c_void_p(dev.data_ptr()), None) == 0
        got = dev.cpu().numpy()
        err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
        assert err.max() < tol and np.median(err) < typical


def _gpu(p, n_pairs, seed, backend, **kw):
    spots, ext, timing = native.run(p, n_pairs, seed, backend=backend, **kw)
    return spots, ext, timing


def _models():
    market, models = _golden_models()
    b, _ = _basket()
    bl, _ = _basket(mixer=True, slv=True)
    from optlab.structured.localcorr import StochasticMix
    from optlab.structured.multiasset import Basket

    bs = Basket(bl.names, bl.models, bl.markets, bl.corr, bl.vol_corr, StochasticMix(0.5, 1.0, -0.8, 2.0, 0.3))
    # every name local vol (the kernel's table form): 3 names (an exact count) and 5 (a bound of 8)
    lv = models["lv"]
    extra = {}
    for k in (3, 5):
        names = tuple(f"N{a}" for a in range(k))
        corr = np.full((k, k), 0.55) + 0.45 * np.eye(k)
        extra[f"basket_lv{k}"] = Basket(names, {n: lv for n in names}, {n: market for n in names}, corr)
    # four names of every kind but SLV (the generic form at an exact count of 4)
    names4 = b.names + ("CCC",)
    corr4 = np.eye(4)
    corr4[:3, :3] = b.corr
    corr4[3, :3] = corr4[:3, 3] = 0.3
    extra["basket4"] = Basket(names4, {**b.models, "CCC": models["lv"]}, {n: market for n in names4}, corr4)
    return market, {**models, "basket": b, "basket_slv_localcorr": bl, "basket_slv_stochmix": bs, **extra}


CASES = ["bs", "lv", "heston", "basket", "basket_slv_localcorr", "basket_slv_stochmix", "basket_lv3", "basket_lv5",
         "basket4"]


@pytest.mark.parametrize("case", CASES)
def test_fp64_paths_equal_the_oracle(lib, case):
    market, models = _models()
    p = native.plan(models[case], market, [0.1, 0.25, 0.3, 0.5], dt_max=1 / 100)
    kw = dict(keep=[0, 1, 3], windows=[(0, 3), (1, 2)])
    want = oracle_paths(p, 3000, 11, first_pair=7, **kw)
    rq, keep, _ = native.paths_request(p, 3000, 11, 7, **kw)
    got = np.empty_like(want)
    rq.out = _abi.ptr(got)
    assert lib.optlab_paths(C.byref(rq), 0, 0) == 0, lib.optlab_last_error()
    rel = np.abs(got / want - 1)
    if "heston" in case or ("basket" in case and "lv" not in case):
        # a variance within an ulp of 0 can be truncated on one side only (test_native.py:103-104)
        assert np.mean(rel > 1e-12) < 1e-3
        assert np.median(rel) < 1e-14
    else:
        assert rel.max() < 1e-12


@pytest.mark.parametrize("case", CASES)
def test_fp32_paths_track_the_oracle(lib, case):
    market, models = _models()
    p = native.plan(models[case], market, [0.5, 2.0])
    want = oracle_paths(p, 20_000, 5)
    spots, _, timing = _gpu(p, 20_000, 5, "gpu")
    assert timing["precision"] == "fp32"
    got = np.stack([s.T for s in spots])
    rel = np.abs(got / want - 1)
    if case == "heston" or "basket" in case:
        assert np.median(rel) < 1e-5 and np.mean(rel > 1e-3) < 0.06
    else:
        assert np.median(rel) < 1e-6 and rel.max() < 1e-4
    # paired price difference: a relative tolerance on top of 4 SE, as the reference's check
    S_c, S_g = want[0, -1], got[0, -1]
    h = len(S_c) // 2
    pay = lambda S: 0.5 * (np.maximum(S[:h] - 100, 0) + np.maximum(S[h:] - 100, 0))  # noqa: E731
    diff = pay(S_c) - pay(S_g)
    assert abs(diff.mean()) < 4 * diff.std(ddof=1) / np.sqrt(h) + 1e-5 * pay(S_c).mean()


@pytest.mark.parametrize("name", ["bs", "lv", "heston"])
def test_fp64_engine_reproduces_the_reference_golden_paths(lib, name):
    market, models = _golden_models()
    p = native.plan(models[name], market, [0.25, 0.5, 1.0])
    spots, _, _ = _gpu(p, 10_000, 7, "cpu")
    samples, total = digest_samples(spots[0])
    g = fixture()["native_cpu"]["paths"][name]
    np.testing.assert_allclose(samples, g["samples"], rtol=1e-12)
    assert total == pytest.approx(g["sum"], rel=1e-12)


@pytest.mark.parametrize("backend", ["cpu", "gpu"])
def test_batches_do_not_change_the_paths(lib, backend):
    market, models = _models()
    p = native.plan(models["basket_slv_localcorr"], market, [0.3, 0.6], dt_max=1 / 80)
    whole, _, _ = _gpu(p, 5000, 3, backend)
    a, _, _ = _gpu(p, 1234, 3, backend)
    b, _, _ = native.run(p, 5000 - 1234, 3, first_pair=1234, backend=backend)
    for k in range(3):
        np.testing.assert_array_equal(whole[k][:1234], a[k][:1234])
        np.testing.assert_array_equal(whole[k][1234:5000], b[k][: 5000 - 1234])
        np.testing.assert_array_equal(whole[k][5000:6234], a[k][1234:])


@pytest.mark.parametrize("case", ["bs", "lv", "heston", "basket_slv_localcorr"])
def test_a_large_grid_gives_the_paths_of_a_small_one(lib, case):
    """The first 2000 pairs of a 2^18-pair run (1024 workgroups, the constant-variance kernel for Black-Scholes)
    against a 2000-pair run, FP64 and FP32, windows included: the pair index alone fixes a path."""
    market, models = _models()
    p = native.plan(models[case], market, [0.25, 0.5], dt_max=1 / 40)
    big, m = 1 << 18, 2000
    for be in ("fp64", "fp32"):
        whole, ext_w, _ = native.run(p, big, 21, backend=be, windows=[(0, 1)])
        part, ext_p, _ = native.run(p, m, 21, backend=be, windows=[(0, 1)])
        for k in range(len(whole)):
            got = np.concatenate([whole[k][:m], whole[k][big:big + m]])
            rel = np.abs(got / part[k] - 1)
            if be == "fp64":
                assert np.median(rel) < 1e-14 and np.mean(rel > 1e-12) < 1e-3, (case, k, rel.max())
            else:
                assert np.median(rel) < 1e-6 and np.mean(rel > 1e-4) < 0.01, (case, k, rel.max())
            lo_w = np.concatenate([ext_w[k][0][0][:m], ext_w[k][0][0][big:big + m]])
            assert np.median(np.abs(lo_w / ext_p[k][0][0] - 1)) < (1e-14 if be == "fp64" else 1e-6)


def test_windows_on_the_gpu_equal_the_min_and_max_of_the_columns(lib):
    market, models = _models()
    times = np.linspace(0.02, 0.5, 25)
    p = native.plan(models["lv"], market, times, dt_max=1 / 100)
    for backend in ("cpu", "gpu"):
        spots, ext, _ = _gpu(p, 4000, 4, backend, windows=[(0, 24), (5, 12)])
        S = spots[0]
        np.testing.assert_array_equal(ext[0][0][0], S.min(axis=1))
        np.testing.assert_array_equal(ext[0][0][1], S.max(axis=1))
        np.testing.assert_array_equal(ext[0][1][0], S[:, 5:13].min(axis=1))
        np.testing.assert_array_equal(ext[0][1][1], S[:, 5:13].max(axis=1))


def test_device_resident_output_equals_the_host_copy(lib):
    torch = _torch()
    market, models = _models()
    p = native.plan(models["lv"], market, [0.25, 1.0])
    spots, _, _ = _gpu(p, 8192, 9, "gpu")
    dev, _ = native.run_device(p, 8192, 9, backend="gpu")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev[0].cpu().numpy().T, spots[0])


def _engine_price(lib, rq, backend):
    out = _abi.Result()
    assert lib.optlab_price(C.byref(rq), backend, 0, C.byref(out)) == 0, lib.optlab_last_error()
    return out


@pytest.mark.parametrize("model,product,legs,strikes,n_steps", FIXED_CASES)
def test_fp64_fixed_products_equal_the_oracle(lib, model, product, legs, strikes, n_steps):
    rq = _request(model, product, legs, strikes, n_steps, n_paths=40_000, H=80.0)
    want = _oracle_price(rq)
    got = _engine_price(lib, rq, 0)
    for k in range(4):
        assert got.price[k] == pytest.approx(want.price[k], rel=1e-12, abs=1e-14)
        assert got.se[k] == pytest.approx(want.se[k], rel=1e-9, abs=1e-14)


@pytest.mark.parametrize("model,product,legs,strikes,n_steps", FIXED_CASES)
def test_fp32_fixed_products_within_half_a_standard_error_of_fp64(lib, model, product, legs, strikes, n_steps):
    rq = _request(model, product, legs, strikes, n_steps, n_paths=200_000, H=80.0)
    d = _engine_price(lib, rq, 0)
    f = _engine_price(lib, rq, 1)
    for k in range(4):
        if d.se[k] > 0:
            assert abs(f.price[k] - d.price[k]) < 0.5 * d.se[k]


def test_fixed_products_validation_table_on_the_gpu(lib):
    """test_prices.cpp:51-88 in FP32 with 10x the CPU path counts."""
    mg = 100.0 * math.exp(0.03)
    rows = [
        (_request(0, 0, (1, 2, 3), (100, 100, 100), 1, 10_000_000), [(0, 9.319738, 0, 0), (1, 6.393699, 0, 0),
                                                                       (2, 2.926039, 0, 0), (3, mg, 0, 0)]),
        (_request(0, 1, (1, 0, 0), (100, 0, 0), 12, 4_000_000), [(0, 5.575692, 0.000114, 0), (1, 5.381250, 0, 0),
                                                                  (2, 5.575692, 0.000114, 0)]),
        (_request(0, 2, (1, 0, 0), (100, 0, 0), 252, 4_000_000, H=80.0), [(0, 9.245769, 0.007422, 0),
                                                                           (1, 9.319738, 0, 0), (3, mg, 0, 0)]),
        (_request(1, 0, (1, 1, 2), (100, 119, 80), 252, 4_000_000), [(0, 8.403656, 0, 0.02), (1, 1.010290, 0, 0.02),
                                                                      (2, 1.737996, 0, 0.02), (3, mg, 0, 0)]),
        (_request(1, 2, (1, 0, 0), (100, 0, 0), 252, 4_000_000, H=80.0), [(0, 8.0927, 0.0092, 0),
                                                                           (1, 8.4153, 0.0131, 0), (3, mg, 0, 0)]),
    ]
    for rq, checks in rows:
        out = _engine_price(lib, rq, 1)
        for k, ref, se_ref, bias in checks:
            assert abs(out.price[k] - ref) <= 4 * math.hypot(out.se[k], se_ref) + bias, (rq.model, rq.product, k)


def test_fp64_strip_equals_the_oracle(lib):
    rq, (strikes, mats, price, se) = _strip(n_paths=60_000)
    assert oracle().oracle_strip(C.byref(rq), 0) == 0
    want_p, want_se = price.copy(), se.copy()
    price[:] = 0
    assert lib.optlab_strip(C.byref(rq), 0) == 0, lib.optlab_last_error()
    np.testing.assert_allclose(price, want_p, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(se, want_se, rtol=1e-9, atol=1e-14)


def test_fp32_strip_against_fp64_and_fourier(lib):
    """Heston QE strip in FP32: within 0.5 SE of the FP64 kernel on the same normals, and within 4 SE (+0.02 QE
    allowance, test_prices.cpp:80-88) of the Fourier price."""
    rq, (strikes, mats, price, se) = _strip(n_paths=1_000_000, n_steps=504, mats=(63, 252, 504),
                                            strikes=(70, 85, 100, 115, 130))
    assert lib.optlab_strip(C.byref(rq), 0) == 0
    p64, se64 = price.copy(), se.copy()
    assert lib.optlab_strip(C.byref(rq), 1) == 0
    assert np.all(np.abs(price - p64) < 0.5 * se64 + 1e-12)
    h = HestonParams(0.04, 1.5, 0.04, 0.6, -0.7)
    for j, T in enumerate((0.25, 1.0, 2.0)):
        F, D = 100 * math.exp(0.03 * T), math.exp(-0.03 * T)
        for k, K in enumerate(strikes):
            assert abs(price[j, k] - heston_price(F, K, T, h, D)) < 4 * se[j, k] + 0.02
        assert abs(price[j, 5] - F) < 4 * se[j, 5]


def _lev(n, T=0.5, local_linear=False):
    market, surface, heston = simulated_context()
    return native.leverage_request(heston, surface, T, 0.5, n, 1 / 100, 1 / 2000, 81, 1.0, 5, (0.05, 8.0),
                                   "local_linear" if local_linear else "nadaraya_watson", False)


@pytest.mark.parametrize("local_linear", [False, True])
def test_fp64_leverage_matches_the_oracle(lib, local_linear):
    rq, want, _ = _lev(20_000, local_linear=local_linear)
    assert oracle().oracle_leverage(C.byref(rq), 0, 0) == 0
    rq, got, _ = _lev(20_000, local_linear=local_linear)
    assert lib.optlab_leverage(C.byref(rq), 0, 0) == 0, lib.optlab_last_error()
    np.testing.assert_allclose(got["x_grids"], want["x_grids"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(got["values"], want["values"], rtol=1e-9, atol=1e-12)


def test_fp32_leverage_is_within_a_fraction_of_its_own_sampling_noise(lib):
    """FP32 particles against FP64 on the same draws (the golden case: 20 000 particles, 1 y): the largest gap of
    the leverage, over the whole grid, is measured against the estimator's own sampling noise, the largest gap
    between the FP64 leverages of two seeds; the x grids (quantiles) against the golden CPU digest within the
    reference GPU's own gap to it (7e-4, native_gpu.json against native_cpu.json)."""
    market, surface, heston = simulated_context()

    def lev(backend, seed):
        rq, got, _ = native.leverage_request(heston, surface, 1.0, 0.5, 20_000, 1 / 100, 1 / 2000, 81, 1.0, seed,
                                             (0.05, 8.0), "nadaraya_watson", False)
        assert lib.optlab_leverage(C.byref(rq), backend, 0) == 0, lib.optlab_last_error()
        return np.asarray(got["x_grids"]).copy(), np.asarray(got["values"]).copy()

    x32, v32 = lev(1, 5)
    x64, v64 = lev(0, 5)
    _, v64b = lev(0, 6)
    rounding = np.max(np.abs(v32 - v64))
    noise = np.max(np.abs(v64b - v64))
    print("FP32 - FP64 leverage", rounding, "FP64 seed 5 - seed 6", noise, flush=True)
    assert rounding < 0.25 * noise
    g = fixture()["native_cpu"]["leverage"]
    samples, _ = digest_samples(x32)
    assert np.max(np.abs(samples - np.array(g["x_grids"]["samples"]))) < 7e-4


def test_bad_requests_raise_without_touching_the_device(lib):
    rq = _request(1, 7, (1, 0, 0), (100, 0, 0), 10)
    out = _abi.Result()
    assert lib.optlab_price(C.byref(rq), 1, 0, C.byref(out)) == _abi.EINVAL
    assert b"product" in lib.optlab_last_error()
    rq = _request(1, 0, (1, 0, 0), (100, 0, 0), 10)
    assert lib.optlab_price(C.byref(rq), 2, 0, C.byref(out)) == _abi.EINVAL


def test_strip_table_cache_never_serves_a_stale_table(lib):
    """Calls alternate strikes, precisions and maturities; each must equal a fresh call of the same request."""
    a, (_, _, pa, _) = _strip(n_paths=8192, strikes=(80, 100, 120))
    b, (_, _, pb, _) = _strip(n_paths=8192, strikes=(90, 95, 105, 110), mats=(26, 104))
    results = []
    for rq, price, backend in ((a, pa, 1), (b, pb, 1), (a, pa, 0), (a, pa, 1), (b, pb, 0), (b, pb, 1), (b, pb, 0),
                               (b, pb, 1)):
        assert lib.optlab_strip(C.byref(rq), backend) == 0
        results.append(price.copy())
    np.testing.assert_array_equal(results[0], results[3])
    np.testing.assert_array_equal(results[1], results[5])
    # the same strip in FP64 then FP32 twice over: the two precisions share the device table's buffer
    np.testing.assert_array_equal(results[4], results[6])
    np.testing.assert_array_equal(results[5], results[7])
    assert not np.array_equal(results[0], results[2])  # FP32 against FP64 of the same strip: not bit-equal


def test_black_scholes_fp64_workload_equals_the_oracle_and_the_closed_forms(lib):
    """The plain FP64 Black-Scholes workload: S0 100, r .03, q .01, sigma .2, T 1, 252 daily steps, 2^18 paths, seed
    42; European call K 100 (one exact step), arithmetic Asian call K 100 with the geometric control variate,
    up-and-out call K 100 B 130 on every close. FP64 GPU = the oracle (1e-12); the European within 4 SE of Black-Scholes, the geometric
    Asian within 4 SE of its closed form, the up-and-out within 4 SE of the continuous price shifted by the
    Broadie-Glasserman-Kou correction."""
    from optlab.core.blackscholes import black_price, geometric_asian, up_out_call_bgk

    common = dict(S0=100.0, r=0.03, q=0.01, T=1.0, sigma=0.20, n_paths=1 << 18, seed=42)
    F, D = 100.0 * math.exp(0.02), math.exp(-0.03)
    # the fixed European under Black-Scholes is one exact step of T / n_steps (schemes.hpp:190-198 with
    # make_consts, :116-117): the whole year in one step; the Asian and the barrier read the 252 daily closes
    euro = _request(0, 0, (1, 0, 0), (100, 0, 0), 1, **common)
    asian = _request(0, 1, (1, 0, 0), (100, 0, 0), 252, **common)
    barrier = _request(0, 3, (1, 0, 0), (100, 0, 0), 252, H=130.0, **common)
    for rq in (euro, asian, barrier):
        want, got = _oracle_price(rq), _engine_price(lib, rq, 0)
        for k in range(4):
            assert got.price[k] == pytest.approx(want.price[k], rel=1e-12, abs=1e-14), (rq.product, k)
    e = _engine_price(lib, euro, 0)
    assert abs(e.price[0] - black_price(F, 100.0, 1.0, 0.2, D)) < 4 * e.se[0]
    a = _engine_price(lib, asian, 0)
    geo = geometric_asian(100.0, 100.0, 1.0, 0.03, 0.01, 0.2, 252)
    assert abs(a.price[1] - geo) < 4 * a.se[1]
    assert abs(a.price[2] - a.price[0]) < 4 * a.se[0] and a.se[2] < 0.2 * a.se[0]  # the variate cuts the SE
    b = _engine_price(lib, barrier, 0)
    assert abs(b.price[0] - up_out_call_bgk(100.0, 100.0, 130.0, 1.0, 0.03, 0.01, 0.2, 252)) < 4 * b.se[0] + 0.01
