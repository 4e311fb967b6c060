"""The SLV particle workload on the GPU: the bandwidth rule 0.25 std(ln S) N^-1/5 on 256 bins and a uniform daily
grid, in FP64 against the oracle, and the calibrated leverage repricing the SSVI surface on SLV paths."""

from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest
from _support import oracle

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402

from optlab import native  # noqa: E402
from optlab.native import _abi  # noqa: E402

pytestmark = pytest.mark.gpu


def _rq(n, T):
    market, surface, heston = bench.slv_inputs()
    return native.leverage_request(heston, surface, T, 1.0, n, 1 / 365, None, 81, 1.0, 17, (0.05, 8.0),
                                   "nadaraya_watson", False, n_bins=256, bandwidth=("std", 0.25))


def _device(n, T, backend, source="rows", regression="nadaraya_watson"):
    """(x_grids, values) of the engine: rows (the host's Dupire table) or slices (one record a step)."""
    lib = _abi.load()
    market, surface, heston = bench.slv_inputs()
    rq, out, _ = native.leverage_request(heston, surface, T, 1.0, n, 1 / 365, None, 81, 1.0, 17, (0.05, 8.0),
                                         regression, False, n_bins=256, bandwidth=("std", 0.25), lv_source=source)
    if source == "rows":
        status = lib.optlab_leverage_ex(C.byref(rq), C.byref(rq._options), backend)
    else:
        status = lib.optlab_leverage_slices(C.byref(rq), C.byref(rq._options), C.byref(rq._slices), backend)
    assert status == 0, lib.optlab_last_error()
    return out["x_grids"].copy(), out["values"].copy()


@pytest.mark.parametrize("form", ["persistent", "five kernels"])
@pytest.mark.parametrize("regression", ["nadaraya_watson", "local_linear"])
def test_fp64_std_bandwidth_leverage_matches_the_oracle(form, regression, monkeypatch):
    if form == "five kernels":
        monkeypatch.setenv("OPTLAB_SLV_WORKGROUPS", "0")
    market, surface, heston = bench.slv_inputs()
    rq, want, _ = native.leverage_request(heston, surface, 0.5, 1.0, 16_384, 1 / 365, None, 81, 1.0, 17,
                                          (0.05, 8.0), regression, False, n_bins=256, bandwidth=("std", 0.25))
    assert oracle().oracle_leverage_ex(C.byref(rq), C.byref(rq._options), 0, 0) == 0
    x, v = _device(16_384, 0.5, 0, regression=regression)
    np.testing.assert_allclose(x, want["x_grids"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v, want["values"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("backend", [0, 1])
def test_persistent_launch_equals_the_five_kernels_bit_for_bit(backend, monkeypatch):
    """Integer bin totals, min / max ranges and a per-workgroup regression in the five-kernel form's arithmetic:
    the leverage does not depend on the form, on the number of workgroups or on the pairs a lane."""
    got = {}
    for groups in ("0", "3", "16", "128"):
        monkeypatch.setenv("OPTLAB_SLV_WORKGROUPS", groups)
        got[groups] = _device(20_002, 0.4, backend)  # 10 001 pairs: partly empty last lanes
    for groups in ("3", "16", "128"):
        np.testing.assert_array_equal(got[groups][0], got["0"][0])
        np.testing.assert_array_equal(got[groups][1], got["0"][1])


@pytest.mark.parametrize("backend", [0, 1])
def test_slice_records_give_the_rows_leverage_bit_for_bit(backend):
    """The local variance evaluated on the device from one slice record a step equals the host's Dupire rows."""
    a = _device(16_384, 0.5, backend, "rows")
    b = _device(16_384, 0.5, backend, "slices")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


SHORT_TO_LONG = [0.125, 0.25, 0.375, 0.75, 1.5]


def _bias_bp(dt, seed=17):
    """Mean and largest |implied-vol gap| (bp) at K 79..121 of calls on SLV paths against the surface, by maturity:
    the leverage calibrated on a uniform grid of step dt (2^18 particles, 256 bins, bandwidth 0.25 std(ln S)
    N^-1/5, to 1.5 y), repriced on 2^20 FP32 paths of the same step."""
    from optlab.core.blackscholes import implied_vol

    market, surface, heston = bench.slv_inputs()
    model = native.calibrate_leverage(heston, surface, market, 1.5, 1.0, 1 << 18, dt_max=dt, dt_min=None, rng=seed,
                                      backend="gpu", n_bins=256, bandwidth=("std", 0.25))
    S = native.simulate(model, market, SHORT_TO_LONG, 1 << 20, seed=23, dt_max=dt, dt_min=None, backend="gpu")
    K = np.asarray(bench.SLV_STRIKES)[3:18]
    h = S.shape[0] // 2
    mean, worst = [], []
    for j, T in enumerate(SHORT_TO_LONG):
        F, D = float(market.F(T)), float(market.D(T))
        pay = D * np.maximum(S[:, j][:, None] - K, 0.0)
        p = (0.5 * (pay[:h] + pay[h:])).mean(axis=0)
        bp = (np.array([implied_vol(a, F, k, T, D) for a, k in zip(p, K)]) - surface.iv(np.log(K / F), T)) * 1e4
        mean.append(float(np.mean(bp)))
        worst.append(float(np.max(np.abs(bp))))
    return np.array(mean), np.array(worst)


def test_the_leverage_reprices_the_surface_on_the_daily_grid():
    """The SLV workload's calibration (365 steps a year, 256 bins, bandwidth 0.25 std(ln S) N^-1/5) reprices the
    surface at K 79..121 within the daily Euler step's bias: measured 17-21 bp on average at 0.125 y falling to 6 bp
    at 1.5 y (largest single strike 12 bp from 0.75 y)."""
    mean, worst = _bias_bp(1 / 365)
    print("daily grid: mean bp", np.round(mean, 1), "largest", np.round(worst, 1), flush=True)
    assert np.all(np.abs(mean) < 30)
    assert np.all(worst[SHORT_TO_LONG.index(0.75):] < 20)


def test_the_repricing_bias_is_the_daily_step():
    """The same calibration and repricing on a step of 1/1000 y: the bias falls to a few bp at every maturity
    (measured 1-7 bp), most at the shortest, where the daily grid's is largest: its cause is the time step, not the
    particles, bins or bandwidth (which the CPU oracle varied without effect)."""
    daily, _ = _bias_bp(1 / 365)
    fine, _ = _bias_bp(1 / 1000)
    print("mean bp daily", np.round(daily, 1), "1/1000", np.round(fine, 1), flush=True)
    assert np.all(np.abs(fine) < 10)
    assert daily[0] - fine[0] > 8


def test_calibration_graph_launches_repeat_bit_for_bit():
    """The calibration's steps run as one captured HIP graph: two calls, with other engine work between them, give
    the same leverage bit for bit (integer bin sums, fixed-order regression)."""
    a = bench.slv_calibrate(1 << 15, 0.5)
    bench.slv_calibrate(1 << 12, 0.25)
    b = bench.slv_calibrate(1 << 15, 0.5)
    np.testing.assert_array_equal(np.asarray(a.leverage.values), np.asarray(b.leverage.values))
