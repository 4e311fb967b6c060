"""The worst-of autocall's adjoint through the SLV, Heston and Black-Scholes steps on the GPU (optlab_autocall_ex):
the FP64 price against the oracle's paths of the same kinds (1e-12); the FP64 deltas and node vegas against
central differences of the oracle's price (rel 1e-5), an SLV name's vegas with E[v | x] = LV / L^2 held (its
squared leverage moves to L^2 LV_bumped / LV under a bump, the oracle's tables moved the same way); FP32 against
FP32 bump-and-revalue on common random numbers (4 SE); both checkpoint intervals; zero vegas for the names whose
variance does not come from a surface."""

from __future__ import annotations

from functools import lru_cache

import numpy as np
import pytest
from _support import oracle
from test_gpu_adjoint import SEED, ao, corr, so, surfaces

from optlab.core.market import Market
from optlab.core.models import BlackScholesModel, HestonModel, HestonParams, Leverage, SLVModel
from optlab.native import _abi

pytestmark = pytest.mark.gpu

SLV_PARAMS = HestonParams(0.04, 1.5, 0.04, 0.6, -0.6)


def leverage(level):
    """A leverage with a skew and a term structure (rows every quarter, 41 points in x)."""
    times = np.arange(0.0, 3.0, 0.25)
    x = np.linspace(-1.5, 1.0, 41)
    grids = np.repeat(x[None, :], len(times), axis=0)
    values = np.stack([level * (1.0 + 0.25 * np.tanh(-1.5 * x)) * (1.0 + 0.05 * t) for t in times])
    return Leverage(times, grids, values)


@lru_cache(maxsize=1)
def slv_note():
    """Three names: SLV (mix 0.8), local vol, Heston; variance normals correlated as the spots (`same`)."""
    from optlab.native.adjoint import WorstOfAutocall

    models = [SLVModel(SLV_PARAMS, leverage(0.95), mix=0.8), None,
              HestonModel(HestonParams(0.05, 2.0, 0.05, 0.5, -0.5))]
    return WorstOfAutocall(surfaces(), [Market.flat(100.0, 0.03, 0.0)] * 3, corr(), models=models, same=True)


@lru_cache(maxsize=1)
def heston_note():
    """Two names without SLV (Heston, Black-Scholes): the variance normals come from stream 0 after the spots'."""
    from optlab.native.adjoint import WorstOfAutocall

    models = [HestonModel(HestonParams(0.04, 1.5, 0.04, 0.6, -0.7)), BlackScholesModel(0.22)]
    return WorstOfAutocall(surfaces()[:2], [Market.flat(100.0, 0.03, 0.0)] * 2, corr(2), models=models)


@lru_cache(maxsize=4)
def base_rows(n_names):
    note = slv_note()
    return [so.base_rows(s, note.times[:-1]) for s in surfaces()[:n_names]]


@lru_cache(maxsize=4)
def base_table(a):
    return so.bumped_local_var(base_rows(3)[a], np.zeros((10, 8)))


def oracle_price(note, n_pairs, seed, nodes=None, spot=None):
    """The oracle's price of the note on its own paths of the note's kinds: local vol from the surface oracle's
    tables, SLV from the note's squared leverage (moved to L^2 LV_bumped / LV under a node bump)."""
    arr = note.model_arrays
    n = note.n_names
    rows = base_rows(3)[:n]
    H = np.zeros((n, 10, 8)) if nodes is None else np.asarray(nodes).reshape(n, 10, 8)
    tables = []
    for a in range(n):
        kind = arr["kind"][a]
        moved = bool(np.any(H[a]))
        if kind == 1:
            tables.append(so.bumped_local_var(rows[a], H[a]) if moved else base_table(a))
        elif kind == 3:
            lev2 = arr["lev2"][a]
            tables.append(lev2 / base_table(a) * so.bumped_local_var(rows[a], H[a]) if moved else lev2)
        else:
            tables.append(np.ones_like(base_table(a)))  # not read by these kinds
    lo = np.stack([r[0] for r in rows])
    inv = np.stack([r[1] for r in rows])
    f = np.ones(n) if spot is None else np.asarray(spot, dtype=float)
    c = ao.closes(oracle(), _abi.PathsRequest, n_pairs, seed, note.root, np.stack(tables), lo, inv,
                  note.fwd * f[:, None], note.dt, kinds=arr["kind"], cvar=arr["cvar"], heston=arr["heston"],
                  same=int(arr["same"]))
    p = note.product
    v = ao.phoenix(c, note.fixing, note.obs_step, note.disc, note.coupon, p.coupon_barrier, p.autocall, p.memory,
                   p.ki_barrier, p.put_strike, p.smoothing)
    return ao.pair_mean(v)


@pytest.mark.parametrize("which", ["slv", "heston"])
def test_fp64_price_is_the_oracle_price_on_the_same_paths(which):
    note = slv_note() if which == "slv" else heston_note()
    n_pairs = 1024
    got, se, _, _ = note.price(2 * n_pairs, SEED, backend="fp64")
    want, want_se = oracle_price(note, n_pairs, SEED)
    assert got == pytest.approx(want, rel=1e-12)
    assert se == pytest.approx(want_se, rel=1e-8)
    if which == "slv":  # a node bump of the SLV name moves its leverage; the price follows the oracle's
        H = np.zeros((3, 10, 8))
        H[0, 4, 5] = 0.01
        got, _, _, _ = note.price(2 * n_pairs, SEED, backend="fp64", nodes=H)
        assert got == pytest.approx(oracle_price(note, n_pairs, SEED, H)[0], rel=1e-12)
        assert abs(got - want) > 1e-5


@pytest.mark.parametrize("which", ["slv", "heston"])
def test_fp64_adjoint_is_the_derivative_of_the_oracle_price(which):
    note = slv_note() if which == "slv" else heston_note()
    n, n_pairs = note.n_names, 512
    g = note.adjoint(2 * n_pairs, SEED, backend="fp64")
    assert g.price == pytest.approx(oracle_price(note, n_pairs, SEED)[0], rel=1e-12)
    h = 1e-7
    for a in range(n):
        f = np.ones(n)
        f[a] = 1 + h
        up = oracle_price(note, n_pairs, SEED, spot=f)[0]
        f[a] = 1 - h
        dn = oracle_price(note, n_pairs, SEED, spot=f)[0]
        assert g.delta[a] == pytest.approx((up - dn) / (2 * h * 100.0), rel=1e-5, abs=1e-9), a
    kinds = note.model_arrays["kind"]
    for a in range(n):
        if kinds[a] not in (1, 3):
            assert np.all(g.vega[a] == 0.0), a  # no surface in this name's variance
    if which != "slv":
        return
    assert np.abs(g.vega[0]).max() > 1e-4 and np.abs(g.vega[1]).max() > 1e-4
    for a in (0, 1):  # the four largest node vegas of the SLV name and of the local-vol name
        for flat in np.argsort(-np.abs(g.vega[a]).ravel())[:4]:
            i, j = np.unravel_index(flat, g.vega[a].shape)
            H = np.zeros((n, 10, 8))
            H[a, i, j] = h
            fd = (oracle_price(note, n_pairs, SEED, H)[0] - oracle_price(note, n_pairs, SEED, -H)[0]) / (2 * h) * 0.01
            assert g.vega[a, i, j] == pytest.approx(fd, rel=1e-5, abs=1e-10), (a, i, j)


def test_checkpoint_intervals_give_the_same_greeks():
    note = slv_note()
    g4 = note.adjoint(1 << 13, 31, backend="fp64", checkpoint_every=4)
    g8 = note.adjoint(1 << 13, 31, backend="fp64", checkpoint_every=8)
    assert g8.price == pytest.approx(g4.price, rel=1e-12)
    np.testing.assert_allclose(g8.delta, g4.delta, rtol=1e-12)
    np.testing.assert_allclose(g8.vega, g4.vega, rtol=1e-9, atol=1e-12)
    f4 = note.adjoint(1 << 13, 31, backend="fp32", checkpoint_every=4)
    f8 = note.adjoint(1 << 13, 31, backend="fp32", checkpoint_every=8)
    assert f4.price == f8.price
    assert np.all(np.abs(f8.vega - f4.vega) <= 0.05 * f4.vega_se + 1e-9)


def test_fp32_adjoint_agrees_with_fp32_bumps_on_common_random_numbers():
    note = slv_note()
    n_paths = 1 << 16
    g = note.adjoint(n_paths, SEED, backend="fp32")
    nodes = []
    for a in (0, 1):
        for flat in np.argsort(-np.abs(g.vega[a]).ravel())[:4]:
            nodes.append((a,) + tuple(int(v) for v in np.unravel_index(flat, g.vega[a].shape)))
    b = note.bump(n_paths, SEED, backend="fp32", nodes=nodes)
    assert g.price == b.price
    for a in range(note.n_names):
        s = np.hypot(g.delta_se[a], b.delta_se[a])
        assert abs(g.delta[a] - b.delta[a]) < 4 * s + 1e-4 * abs(b.delta[a]), (a, g.delta[a], b.delta[a], s)
    for a, i, j in nodes:
        s = np.hypot(g.vega_se[a, i, j], b.vega_se[a, i, j])
        assert abs(g.vega[a, i, j] - b.vega[a, i, j]) < 4 * s + 0.02 * abs(b.vega[a, i, j]), (a, i, j)


def test_models_are_checked():
    from optlab.core.engines import EngineError

    note = slv_note()
    rq = note.request(1024, 1)
    rq._prec = 0
    kinds = note.model_arrays["kind"]
    saved = kinds[2]
    kinds[2] = 7
    try:
        with pytest.raises(EngineError, match="kind must be"):
            note._call(False, rq)
    finally:
        kinds[2] = saved
    note._call(False, rq)
