"""The allocation-free strip (optlab_strip_noalloc) against optlab_strip through the C ABI: the same Heston QE-M and
Black-Scholes strips in FP64 and FP32, two calls in a row on the shared device array, the second smaller than the
first (fewer paths, fewer maturities and strikes), so what the first call left behind must not leak into the
second. The two entry points run the same kernels; with the same grid they give the same bits, and a grid capped
by the array's size changes only the order of the FP64 sums (1e-12)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from optlab.core.market import Market
from optlab.core.models import BlackScholesModel, HestonModel, HestonParams
from optlab.native import _abi, gpu

pytestmark = pytest.mark.gpu

HESTON = HestonModel(HestonParams(0.04, 1.5, 0.04, 0.6, -0.7))
MARKET = Market.flat(100.0, 0.03, 0.0)


def _pair(model, strikes, mat_steps, n_steps, T, n_paths, backend, stream):
    """(optlab_strip's price, se) and (optlab_strip_noalloc's price, se) for one request."""
    import torch

    lib = _abi.load()
    rq, out = gpu.strip_request(model, MARKET, strikes, mat_steps, n_steps, T, n_paths, 42)
    assert lib.optlab_strip(C.byref(rq), backend) == 0, lib.optlab_last_error()
    want = out["price"].copy(), out["se"].copy()
    shape = out["price"].shape
    price = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
    se = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    status = lib.optlab_strip_noalloc(C.byref(rq), backend, C.c_void_p(price.data_ptr()), C.c_void_p(se.data_ptr()),
                                      C.c_void_p(stream.cuda_stream))
    assert status == 0, lib.optlab_last_error()
    stream.synchronize()
    return want, (price.cpu().numpy(), se.cpu().numpy())


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("model", ["heston", "bs"])
def test_noalloc_strip_equals_the_strip_on_two_calls_in_a_row(backend, model):
    import torch

    m = HESTON if model == "heston" else BlackScholesModel(0.2)
    stream = torch.cuda.Stream()
    big = _pair(m, np.arange(70.0, 131.0, 3.0), [63 * j for j in range(1, 9)], 504, 2.0, 1 << 18, backend, stream)
    small = _pair(m, [90.0, 100.0, 110.0], [16, 48, 64], 64, 0.5, 5002, backend, stream)
    for (want_p, want_se), (got_p, got_se) in (big, small):
        assert np.all(np.isfinite(got_p)) and np.all(np.isfinite(got_se))
        np.testing.assert_allclose(got_p, want_p, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(got_se, want_se, rtol=1e-9, atol=1e-14)
    assert small[1][0].shape == (3, 4) and big[1][0].shape == (8, 22)


def test_noalloc_strip_checks_its_request():
    import torch

    lib = _abi.load()
    rq, _ = gpu.strip_request(HESTON, MARKET, [100.0], [10], 10, 0.1, 1000, 1)
    buf = torch.empty(2, dtype=torch.float64, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    assert lib.optlab_strip_noalloc(C.byref(rq), 1, None, p, None) == _abi.EINVAL
    assert b"null device output" in lib.optlab_last_error()
    assert lib.optlab_strip_noalloc(C.byref(rq), 2, p, p, None) == _abi.EINVAL  # backend
    rq.n_mats = 17
    assert lib.optlab_strip_noalloc(C.byref(rq), 1, p, p, None) == _abi.EINVAL
