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


def test_the_leverage_reprices_the_surface():
    """test_native.py:63-72's bar for SLV: |price - target| < 4 SE + 0.03 on every call of the grid."""
    model = bench.slv_calibrate(1 << 17, 1.0)
    mats = [0.25, 0.5, 0.75, 1.0]
    r = bench.slv_reprice(model, 1 << 18, mats, bench.SLV_STRIKES)
    assert np.all(np.abs(r["price_gap"]) < 4 * r["se"] + 0.03), np.abs(r["price_gap"]).max()
    assert np.nanmax(np.abs(r["vol_bp"][:, 3:18])) < 100  # K 79..121: within a vol point


def test_calibration_graph_launches_repeat_bit_for_bit():
    """The calibration's steps run as one captured HIP graph: two calls, with other engine work between them, give
    the same leverage bit for bit (integer bin sums, fixed-order regression)."""
    a = bench.slv_calibrate(1 << 15, 0.5)
    bench.slv_calibrate(1 << 12, 0.25)
    b = bench.slv_calibrate(1 << 15, 0.5)
    np.testing.assert_array_equal(np.asarray(a.leverage.values), np.asarray(b.leverage.values))
