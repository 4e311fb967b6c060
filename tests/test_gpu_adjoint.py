"""The worst-of autocall and its adjoint Greeks on the GPU: device-built bumped local-variance tables and their node
derivatives against the surface oracle; the FP64 price against the oracle's paths and script; FP64 adjoint deltas
and node vegas against central differences of the oracle; FP32 adjoint against FP32 bump-and-revalue on common
random numbers (within 4 standard errors); bit-reproducibility; the torch.autograd.Function."""

from __future__ import annotations

import sys
from functools import lru_cache

import numpy as np
import pytest
from _support import ROOT, oracle

sys.path.insert(0, str(ROOT / "oracle"))
import autocall_oracle as ao  # noqa: E402
import surface_oracle as so  # noqa: E402

from optlab.core.market import Market  # noqa: E402
from optlab.core.ssvi import SSVI  # noqa: E402
from optlab.native import _abi  # noqa: E402

pytestmark = pytest.mark.gpu

PILLARS = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0])
SCALES, RHOS = (0.03, 0.04, 0.05), (-0.5, -0.6, -0.7)
SEED = 20240


def surfaces():
    return [SSVI.power_law(r, 1.2, 0.5, PILLARS, s * PILLARS) for s, r in zip(SCALES, RHOS)]


def corr(n=3, rho=0.6):
    return np.full((n, n), rho) + (1 - rho) * np.eye(n)


@lru_cache(maxsize=1)
def note():
    from optlab.native.adjoint import WorstOfAutocall

    return WorstOfAutocall(surfaces(), [Market.flat(100.0, 0.03, 0.0)] * 3, corr())


@lru_cache(maxsize=1)
def oracle_rows():
    n = note()
    return [so.base_rows(s, n.times[:-1]) for s in surfaces()]


@lru_cache(maxsize=3)
def oracle_table(a):
    return so.bumped_local_var(oracle_rows()[a], np.zeros((10, 8)))


def oracle_price(n_pairs, seed, nodes=None, spot=None):
    """The oracle's price of the note on its own paths (double), tables from the surface oracle."""
    n = note()
    rows = oracle_rows()
    H = np.zeros((3, 10, 8)) if nodes is None else np.asarray(nodes).reshape(3, 10, 8)
    tables = np.stack([so.bumped_local_var(r, H[a]) if np.any(H[a]) else oracle_table(a) for a, r in enumerate(rows)])
    lo = np.stack([r[0] for r in rows])
    inv = np.stack([r[1] for r in rows])
    f = np.ones(3) if spot is None else np.asarray(spot, dtype=float)
    c = ao.closes(oracle(), _abi.PathsRequest, n_pairs, seed, n.root, tables, lo, inv, n.fwd * f[:, None], n.dt)
    p = n.product
    v = ao.phoenix(c, n.fixing, n.obs_step, n.disc, n.coupon, p.coupon_barrier, p.autocall, p.memory, p.ki_barrier,
                   p.put_strike, p.smoothing)
    return ao.pair_mean(v)


def test_device_tables_match_the_surface_oracle():
    from optlab.risk.surfaces import DeviceSurface

    s = surfaces()[2]
    t0 = np.arange(0, 300, 7) / 252
    dev = DeviceSurface(s, t0)
    rows = so.base_rows(s, t0)
    rng = np.random.default_rng(5)
    H = rng.normal(0, 0.01, (10, 8))
    for nodes, par, skew in ((None, 0.0, 0.0), (H, 0.0, 0.0), (None, 0.01, 0.0), (None, 0.0, 0.02), (H, -0.005, 0.01)):
        got = dev.local_var(nodes, par, skew)
        want = so.bumped_local_var(rows, np.zeros((10, 8)) if nodes is None else nodes, parallel=par, skew=skew)
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-15)
    # a bump of every node by h is the parallel bump h
    np.testing.assert_allclose(dev.local_var(np.full((10, 8), 0.01)), dev.local_var(None, 0.01), rtol=1e-12)
    dev.close()


def test_node_gradient_is_the_derivative_of_the_table():
    from optlab.risk.surfaces import DeviceSurface

    s = surfaces()[1]
    t0 = np.arange(0, 756, 9) / 252
    dev = DeviceSurface(s, t0)
    rows = so.base_rows(s, t0)
    rng = np.random.default_rng(8)
    tbar = rng.normal(0, 1, (len(t0), 801))
    got = dev.node_gradient(tbar)
    h = 1e-6
    for i, j in ((5, 3), (0, 7), (3, 5), (9, 1), (6, 6), (2, 0)):
        H = np.zeros((10, 8))
        H[i, j] = h
        fd = np.sum(tbar * (so.bumped_local_var(rows, H) - so.bumped_local_var(rows, -H))) / (2 * h)
        assert got[i, j] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(got).max()), (i, j)
    dev.close()


def test_fp64_price_is_the_oracle_price_on_the_same_paths():
    n_pairs = 1024
    got, se, _, _ = note().price(2 * n_pairs, SEED, backend="fp64")
    want, want_se = oracle_price(n_pairs, SEED)
    assert got == pytest.approx(want, rel=1e-12)
    assert se == pytest.approx(want_se, rel=1e-8)
    # a spot bump and a node bump are priced on the same paths too
    H = np.zeros((3, 10, 8))
    H[1, 5, 6] = 0.01
    f = np.array([1.0, 1.01, 1.0])
    got, _, _, _ = note().price(2 * n_pairs, SEED, backend="fp64", nodes=H, spot=f)
    assert got == pytest.approx(oracle_price(n_pairs, SEED, H, f)[0], rel=1e-12)


def test_fp64_adjoint_is_the_derivative_of_the_oracle_price():
    n_pairs = 512
    g = note().adjoint(2 * n_pairs, SEED, backend="fp64")
    assert g.price == pytest.approx(oracle_price(n_pairs, SEED)[0], rel=1e-12)
    # the local-variance table is piecewise linear in x: a difference step must be small enough that few paths cross
    # a grid node between the two bumps (at 1e-5 the secant is still 5e-4 away from the derivative, at 1e-7 1e-8)
    h = 1e-7
    for a in range(3):
        f = np.ones(3)
        f[a] = 1 + h
        up = oracle_price(n_pairs, SEED, spot=f)[0]
        f[a] = 1 - h
        dn = oracle_price(n_pairs, SEED, spot=f)[0]
        fd = (up - dn) / (2 * h * 100.0)
        assert g.delta[a] == pytest.approx(fd, rel=1e-5, abs=1e-9), a
    order = np.argsort(-np.abs(g.vega).ravel())[:6]  # the six largest node vegas
    for flat in order:
        a, i, j = np.unravel_index(flat, g.vega.shape)
        H = np.zeros((3, 10, 8))
        H[a, i, j] = h
        fd = (oracle_price(n_pairs, SEED, H)[0] - oracle_price(n_pairs, SEED, -H)[0]) / (2 * h) * 0.01
        assert g.vega[a, i, j] == pytest.approx(fd, rel=1e-5, abs=1e-10), (a, i, j)


def test_fp32_adjoint_agrees_with_fp32_bumps_on_common_random_numbers():
    n_paths = 1 << 16
    nt = note()
    g = nt.adjoint(n_paths, SEED, backend="fp32")
    order = np.argsort(-np.abs(g.vega).ravel())[:8]
    nodes = [np.unravel_index(f, g.vega.shape) for f in order]
    b = nt.bump(n_paths, SEED, backend="fp32", nodes=nodes)
    assert g.price == b.price
    for a in range(3):
        s = np.hypot(g.delta_se[a], b.delta_se[a])
        assert abs(g.delta[a] - b.delta[a]) < 4 * s + 1e-4 * abs(b.delta[a]), (a, g.delta[a], b.delta[a], s)
    for a, i, j in nodes:
        s = np.hypot(g.vega_se[a, i, j], b.vega_se[a, i, j])
        assert abs(g.vega[a, i, j] - b.vega[a, i, j]) < 4 * s + 0.02 * abs(b.vega[a, i, j]), (a, i, j)
    # FP32 against FP64 on the same draws: within 4 standard errors
    p64 = nt.price(n_paths, SEED, backend="fp64")
    assert abs(g.price - p64[0]) < 4 * g.se


def test_adjoint_greeks_are_bit_reproducible():
    nt = note()
    g1 = nt.adjoint(1 << 15, 77, backend="fp32")
    g2 = nt.adjoint(1 << 15, 77, backend="fp32")
    assert g1.price == g2.price
    np.testing.assert_array_equal(g1.delta, g2.delta)
    np.testing.assert_array_equal(g1.vega, g2.vega)
    np.testing.assert_array_equal(g1.vega_se, g2.vega_se)


def test_checkpoint_intervals_give_the_same_greeks_and_the_autotuner_picks_one():
    """Every 4 or every 8 steps (3 names): the same price and deltas (the forward pass is the same) and the same
    node vegas up to the 2^-32 fixed point of the table adjoints; with checkpoint_every = 0 the autotuner runs one
    of the candidates, and its Greeks are that candidate's."""
    nt = note()
    g4 = nt.adjoint(1 << 14, 31, backend="fp64", checkpoint_every=4)
    g8 = nt.adjoint(1 << 14, 31, backend="fp64", checkpoint_every=8)
    assert g4.seconds["checkpoint_every"] == 4 and g8.seconds["checkpoint_every"] == 8
    assert g8.price == pytest.approx(g4.price, rel=1e-12)
    np.testing.assert_allclose(g8.delta, g4.delta, rtol=1e-12)
    np.testing.assert_allclose(g8.vega, g4.vega, rtol=1e-9, atol=1e-12)
    by = {k: nt.adjoint(1 << 14, 31, backend="fp32", checkpoint_every=k) for k in (4, 8)}
    g = nt.adjoint(1 << 14, 31, backend="fp32")
    k = g.seconds["checkpoint_every"]
    assert k in by
    assert g.price == by[k].price
    np.testing.assert_array_equal(g.vega, by[k].vega)
    # FP32: the same paths, so the two intervals differ by far less than the vegas' standard errors
    assert np.all(np.abs(by[8].vega - by[4].vega) <= 0.05 * by[4].vega_se + 1e-9)
    rq = nt.request(1024, 1)
    rq._prec = 1
    rq.checkpoint_every = 16  # one name only
    from optlab.core.engines import EngineError

    with pytest.raises(EngineError, match="checkpoint_every"):
        nt._call(True, rq)


def test_one_name_note_at_every_checkpoint_interval():
    """One name: intervals 4, 8 and 16 give the same price and delta and the same vegas up to the fixed point."""
    from optlab.native.adjoint import WorstOfAutocall

    nt = WorstOfAutocall(surfaces()[:1], [Market.flat(100.0, 0.03, 0.0)], corr(1))
    g = {k: nt.adjoint(1 << 13, 5, backend="fp64", checkpoint_every=k) for k in (4, 8, 16)}
    for k in (8, 16):
        assert g[k].seconds["checkpoint_every"] == k
        assert g[k].price == pytest.approx(g[4].price, rel=1e-12)
        np.testing.assert_allclose(g[k].delta, g[4].delta, rtol=1e-12)
        np.testing.assert_allclose(g[k].vega, g[4].vega, rtol=1e-9, atol=1e-12)
    auto = nt.adjoint(1 << 13, 5, backend="fp32")
    assert auto.seconds["checkpoint_every"] in (4, 8, 16)
    nt.close()


def test_autograd_function_carries_the_adjoint():
    import torch

    from optlab.native.adjoint import autocall_function

    nt = note()
    fn = autocall_function()
    spots = torch.tensor(nt.spots, dtype=torch.float64, requires_grad=True)
    nodes = torch.zeros((3, 10, 8), dtype=torch.float64, requires_grad=True)
    price = fn.apply(spots, nodes, nt, 1 << 14, 5, "fp32")
    price.backward()
    g = nt.adjoint(1 << 14, 5, backend="fp32")
    assert price.item() == g.price
    np.testing.assert_allclose(spots.grad.numpy(), g.delta, rtol=1e-12)
    np.testing.assert_allclose(nodes.grad.numpy(), g.vega / 0.01, rtol=1e-12)


def test_requests_are_checked():
    from optlab.core.engines import EngineError

    nt = note()
    rq = nt.request(1024, 1)
    rq.n_obs = 0
    rq._prec = 1
    with pytest.raises(EngineError, match="n_obs"):
        nt._call(False, rq)
    rq = nt.request(1024, 1, nodes=np.zeros((3, 10, 8)))
    rq._prec = 1
    with pytest.raises(EngineError, match="nodes must be NULL"):
        nt._call(True, rq)
