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


def test_fp64_std_bandwidth_leverage_matches_the_oracle():
    lib = _abi.load()
    rq, want, _ = _rq(16_384, 0.5)
    assert oracle().oracle_leverage_ex(C.byref(rq), C.byref(rq._options), 0, 0) == 0
    rq, got, _ = _rq(16_384, 0.5)
    assert lib.optlab_leverage_ex(C.byref(rq), C.byref(rq._options), 0) == 0, lib.optlab_last_error()
    np.testing.assert_allclose(got["x_grids"], want["x_grids"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(got["values"], want["values"], rtol=1e-9, atol=1e-12)


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
