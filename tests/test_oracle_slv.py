"""The oracle's SLV bandwidth rule pinned without a GPU: rule 1 (coef std(ln S) N^-1/5 over the particles at each
step) against NumPy's population standard deviation of the oracle's own particles, and against rule 0 (the
reference's bandwidth, leverage.cpp:63) at the bandwidth scale that makes the two equal. A plausible slip -- the
standard deviation of the wrong sample (the variance factor, or the particles after the step), pairs counted for
particles, the sample instead of the population variance -- fails one of these."""

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

COEF = 0.25


def _request(n_particles, T, dt, bandwidth, atm_vol=None):
    market, surface, heston = bench.slv_inputs()
    return native.leverage_request(heston, surface, T, 1.0, n_particles, dt, None, 81, 1.0, 17, (0.05, 8.0),
                                   "nadaraya_watson", False, n_bins=256, bandwidth=bandwidth, atm_vol=atm_vol)


def _trace(rq, x_step=-1):
    n = rq.n_steps
    bw, sd = np.zeros(n), np.zeros(n)
    x = np.zeros(2 * rq.n_pairs)
    assert oracle().oracle_leverage_trace(C.byref(rq), C.byref(rq._options), 2, 2, _abi.ptr(bw), _abi.ptr(sd),
                                          int(x_step), _abi.ptr(x)) == 0
    return bw, sd, x


@pytest.mark.parametrize("step", [1, 7, 36])
def test_rule_one_bandwidth_is_coef_times_the_particles_population_std(step):
    n_particles = 6_000
    rq, _, _ = _request(n_particles, 0.1, 1 / 365, ("std", COEF))
    assert rq.n_steps == 37
    bw, sd, x = _trace(rq, step)
    assert x.std() > 0 and np.all(x[: n_particles // 2] != x[n_particles // 2:])  # both legs of every pair
    want = COEF * np.std(x) * float(n_particles) ** -0.2  # population std over all 2 n_pairs particles
    np.testing.assert_allclose(sd[step], np.std(x), rtol=1e-11)
    np.testing.assert_allclose(bw[step], want, rtol=1e-11)
    # the slips this rules out, each far outside the tolerance
    assert abs(np.std(x, ddof=1) / np.std(x) - 1) > 1e-5  # sample variance
    assert abs((n_particles / 2) ** -0.2 / n_particles ** -0.2 - 1) > 0.1  # pairs for particles
    assert bw[0] == 0.0 and np.all(bw[1:] > 0)


def test_rule_one_equals_rule_zero_at_the_equivalent_scale_bit_for_bit():
    """Two steps of a quarter year with atm_vol 2: rule 0's bandwidth at step 1 is scale * 2 * sqrt(0.25) * N^-0.2
    = scale * N^-0.2 exactly, so scale = coef * std (the oracle's own std, a product by a power of two) must give
    rule 1's leverage bit for bit."""
    rq1, got1, _ = _request(4_000, 0.5, 0.25, ("std", COEF), atm_vol=2.0)
    assert rq1.n_steps == 2 and list(np.asarray(got1["values"]).shape) == [2, 81]
    bw1, sd1, _ = _trace(rq1)
    rq0, got0, _ = _request(4_000, 0.5, 0.25, None, atm_vol=2.0)
    rq0.bandwidth_scale = COEF * sd1[1]
    bw0, _, _ = _trace(rq0)
    assert bw0[1] == bw1[1]
    np.testing.assert_array_equal(got0["x_grids"], got1["x_grids"])
    np.testing.assert_array_equal(got0["values"], got1["values"])
    rq0.bandwidth_scale = COEF * sd1[1] * (1 + 1e-6)  # and a different scale moves the leverage
    _trace(rq0)
    assert not np.array_equal(got0["values"][1], got1["values"][1])


def _record_entry(rec, k, floor=1e-4, cap=4.0):
    """One table entry from a slice record, in the device's operation order (scalar IEEE doubles, no fused
    multiply-add): what csrc/leverage.hip's slice_dupire computes."""
    import math

    acc = None
    for j in range(2):
        c, dc = rec[j], rec[2 + j]
        a, b, rho, m, s2 = rec[4 + 5 * j: 9 + 5 * j]
        da, db, drho, dm, ds2 = rec[14 + 5 * j: 19 + 5 * j]
        x = k - m
        root = math.sqrt(x * x + s2)
        lin = rho * x + root
        s_k = b * (rho + x / root)
        s_kk = b * s2 / (root * root * root)
        s_t = da + db * lin + b * drho * x - s_k * dm + 0.5 * b * ds2 / root
        t = (c * (a + b * lin), c * s_k, c * s_kk, dc * (a + b * lin) + c * s_t)
        acc = t if acc is None else tuple(u + v for u, v in zip(acc, t))
    w, wk, wkk, wT = acc
    u = 1.0 - 0.5 * k * wk / w
    g = u * u - 0.25 * wk * wk * (1.0 / w + 0.25) + 0.5 * wkk
    return min(max(wT / max(g, 1e-6), floor), cap)


@pytest.mark.parametrize("which", ["ssvi", "svi"])
def test_slice_records_give_the_host_table_bit_for_bit(which):
    """The records the particle method hands the device (one a step) against the Dupire rows the host builds:
    every sampled entry, evaluated in the device's order, is the table's number exactly -- the layout (weights,
    slices, T-derivatives), the grid (j dx + lo, the last point hi) and the formula all pinned."""
    from _support import simulated_context

    if which == "ssvi":
        _, surface, _ = bench.slv_inputs()
        t0 = np.arange(0, 3 * 365) / 365.0
    else:
        _, surface, _ = simulated_context()
        t0 = np.concatenate([np.arange(0, 400) / 200.0, [2.5, 3.1, 4.0]])  # before, between and past the pillars
    lo, inv, rows = surface.local_var_rows(t0, native.N_X)
    lo2, hi, inv2, rec = surface.row_records(t0, native.N_X)
    np.testing.assert_array_equal(lo, lo2)
    np.testing.assert_array_equal(inv, inv2)
    rng = np.random.default_rng(3)
    n = native.N_X
    for i in rng.choice(len(t0), 40, replace=False).tolist() + [0, len(t0) - 1]:
        for j in sorted(set(rng.choice(n, 25).tolist() + [0, 1, n // 2, n - 2, n - 1])):
            k = hi[i] if j == n - 1 else j * ((hi[i] - lo[i]) / (n - 1)) + lo[i]
            assert _record_entry(rec[i].tolist(), float(k)) == rows[i, j], (which, i, j)
