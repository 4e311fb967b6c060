"""The book of trades under scenarios on the GPU: every product and scenario move in FP64 against the oracle on the
same paths (1e-12); the persistent batched launch (captured in a HIP graph, and direct) against the per-trade loop
(bit-identical); the book's autocall
against the autocall engine; results independent of the work split; the benchmark book's VaR run end to end at a
reduced path count."""

from __future__ import annotations

import sys

import numpy as np
import pytest
from _support import ROOT, oracle

sys.path.insert(0, str(ROOT / "oracle"))
import book_oracle as bo  # noqa: E402
import surface_oracle as so  # noqa: E402

from optlab.core.market import Market  # noqa: E402
from optlab.core.ssvi import SSVI  # noqa: E402
from optlab.native import _abi  # noqa: E402
from optlab.risk.book import Book, Scenarios, Trade, pnl, revalue, synthetic_book, synthetic_scenarios, var_es  # noqa: E402
from optlab.structured.multiasset import correlation_root  # noqa: E402

pytestmark = pytest.mark.gpu

PILLARS = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0])


def small_book(n_steps=189, n_paths=512):
    surfaces = [SSVI.power_law(r, 1.2, 0.5, PILLARS, s * PILLARS) for s, r in ((0.03, -0.5), (0.04, -0.6), (0.05, -0.7))]
    markets = [Market.flat(100.0, 0.03, q) for q in (0.0, 0.01, 0.02)]
    corr = np.full((3, 3), 0.6) + 0.4 * np.eye(3)
    trades = [
        Trade("european", (1,), 150, 1, 0.95, n_paths=n_paths, seed=11),
        Trade("asian", (2,), 168, 21, 1.02, n_paths=n_paths, seed=12),
        Trade("barrier", (0,), 189, 1, 1.0, 1.25, n_paths=n_paths, seed=13),
        Trade("autocall", (0, 1, 2), 189, 63, 1.0, 0.6, 0.02, 0.7, 1.0, True, n_paths=n_paths, seed=14),
        Trade("autocall", (2,), 126, 63, 1.0, 0.8, 0.03, 0.9, 1.05, False, n_paths=n_paths, seed=15),
    ]
    return Book(surfaces, markets, corr, trades, n_steps=n_steps)


SCEN = Scenarios(np.array([[0.0, 0.0, 0.0], [0.03, -0.02, 0.01], [-0.05, 0.0, 0.04]]),
                 np.array([0.0, 0.015, -0.02]), np.array([0.0, -0.005, 0.01]))


def wide_book(n_steps=126, n_paths=512):
    """Five names, and autocalls on 2, 4 and 5 of them: the kernels of those name counts (the small book has trades
    on 1 and 3 names only)."""
    params = ((0.03, -0.5), (0.04, -0.6), (0.05, -0.7), (0.035, -0.55), (0.045, -0.65))
    surfaces = [SSVI.power_law(r, 1.2, 0.5, PILLARS, s * PILLARS) for s, r in params]
    markets = [Market.flat(100.0, 0.03, q) for q in (0.0, 0.01, 0.02, 0.005, 0.015)]
    corr = np.full((5, 5), 0.5) + 0.5 * np.eye(5)
    trades = [
        Trade("autocall", (3, 1), 126, 42, 1.0, 0.6, 0.02, 0.7, 1.0, True, n_paths=n_paths, seed=21),
        Trade("autocall", (0, 2, 4, 1), 126, 63, 1.0, 0.65, 0.03, 0.75, 1.0, False, n_paths=n_paths, seed=22),
        Trade("autocall", (4, 3, 2, 1, 0), 84, 21, 1.0, 0.6, 0.025, 0.7, 1.02, True, n_paths=n_paths, seed=23),
        Trade("european", (4,), 105, 1, 1.05, n_paths=n_paths, seed=24),
    ]
    return Book(surfaces, markets, corr, trades, n_steps=n_steps)


WIDE_SCEN = Scenarios(np.array([[0.0] * 5, [0.02, -0.01, 0.03, -0.02, 0.01]]), np.array([0.0, -0.01]),
                      np.array([0.0, 0.005]))


def test_fp64_book_is_the_oracle_on_the_same_paths():
    _fp64_book_against_the_oracle("batched", 128, 2)


def test_fp64_book_in_the_per_trade_loop_is_the_oracle_on_the_same_paths():
    _fp64_book_against_the_oracle("loop", 256, 1)


def test_fp64_book_on_2_4_and_5_names_is_the_oracle_on_the_same_paths():
    _fp64_book_against_the_oracle("batched", 256, 1, wide_book(), WIDE_SCEN)


def _fp64_book_against_the_oracle(mode, block, ppl, book=None, scen=SCEN):
    book = book or small_book()
    print(f"book fp64 {mode} block {block} ppl {ppl}: engine call", flush=True)
    rev = revalue(book, scen, backend="fp64", mode=mode, block=block, pairs_per_lane=ppl)
    print("engine call returned; oracle", flush=True)
    t = np.arange(book.n_steps + 1) * book.dt
    rows = [so.base_rows(s, t[:-1]) for s in book.surfaces]
    ob = {"rows": rows, "fwd": np.stack([m.F(t) for m in book.markets]), "disc": book.markets[0].D(t), "dt": book.dt}
    for i, tr in enumerate(book.trades):
        sub = book.corr[np.ix_(tr.names, tr.names)]
        root = correlation_root(sub) if len(tr.names) > 1 else np.ones((1, 1))
        td = dict(product=tr.product, names=tr.names, root=root, n_steps=tr.n_steps, obs_every=tr.obs_every,
                  strike=tr.strike, barrier=tr.barrier, coupon=tr.coupon, coupon_barrier=tr.coupon_barrier,
                  autocall=tr.autocall, memory=tr.memory, n_pairs=tr.n_paths // 2, seed=tr.seed)
        for s in range(len(scen.vol_shift)):
            sc = {"spot": scen.spot_shock[s], "vol": scen.vol_shift[s], "skew": scen.skew[s]}
            want, want_se = bo.trade_value(oracle(), _abi.PathsRequest, td, ob, sc)
            assert rev.value[i, s] == pytest.approx(want, rel=1e-12, abs=1e-13), (tr.product, s)
            assert rev.se[i, s] == pytest.approx(want_se, rel=1e-8, abs=1e-13), (tr.product, s)


def test_batched_launch_equals_the_per_trade_loop():
    book = small_book(n_paths=4096)
    print("book fp32 batched launch", flush=True)
    a = revalue(book, SCEN, mode="batched", block=256, pairs_per_lane=1)
    print("book fp32 per-trade loop", flush=True)
    b = revalue(book, SCEN, mode="loop", block=256, pairs_per_lane=1)
    np.testing.assert_array_equal(a.value, b.value)
    np.testing.assert_array_equal(a.se, b.se)
    for rep in range(3):  # graph launches again, with direct launches and synchronisations between them
        print(f"book fp32 direct launch, then graph launch {rep}", flush=True)
        c = revalue(book, SCEN, mode="direct", block=256, pairs_per_lane=1)
        d = revalue(book, SCEN, mode="batched", block=256, pairs_per_lane=1)
        np.testing.assert_array_equal(c.value, b.value)
        np.testing.assert_array_equal(d.value, b.value)
    assert np.all(a.finish_ms > 0) and np.all(b.finish_ms > 0)


def test_work_split_does_not_move_the_prices():
    book = small_book(n_paths=4096)
    ref = revalue(book, SCEN, block=256, pairs_per_lane=1)
    for block, ppl in ((128, 4), (512, 2), (256, 8), (0, 0)):
        got = revalue(book, SCEN, block=block, pairs_per_lane=ppl)
        np.testing.assert_allclose(got.value, ref.value, rtol=1e-12, atol=1e-15)
        assert got.block in (128, 256, 512) and got.pairs_per_lane in (1, 2, 4, 8)


def test_book_autocall_is_the_autocall_engine():
    from optlab.native.adjoint import Phoenix, WorstOfAutocall

    book = small_book(n_paths=8192)
    base = Scenarios(np.zeros((1, 3)), np.zeros(1), np.zeros(1))
    rev = revalue(book, base)
    tr = book.trades[3]
    note = WorstOfAutocall(book.surfaces, book.markets, book.corr,
                           Phoenix(autocall=1.0, coupon_barrier=0.7, coupon_rate=0.08, frequency_months=3,
                                   maturity_months=9, memory=True, ki_barrier=0.6, put_strike=1.0, smoothing=0.0))
    assert note.n_steps == tr.n_steps
    price, se, _, _ = note.price(tr.n_paths, tr.seed)
    assert rev.value[3, 0] == pytest.approx(price, rel=2e-6)
    assert rev.se[3, 0] == pytest.approx(se, rel=1e-3)
    note.close()


def test_benchmark_book_var_runs_end_to_end():
    book = synthetic_book(path_scale=64)
    scen = synthetic_scenarios(len(book.surfaces), n=40)
    rev = revalue(book, scen)
    assert rev.value.shape == (512, 41) and np.all(np.isfinite(rev.value))
    assert np.all(rev.value >= 0)
    p = pnl(book, rev)
    r = var_es(p)
    # ES at 97.5 % is the mean loss beyond the 97.5 % quantile: at least that quantile's loss (it may sit below the
    # 99 % VaR, the reference's two levels differ)
    q975 = -np.quantile(p, 0.025, method="lower")
    assert r["var"] > 0 and r["es"] >= q975 - 1e-12 and r["var"] >= q975 - 1e-12
    assert rev.name_steps > 0 and rev.kernel_ms > 0


def test_benchmark_book_at_its_configuration():
    """The benchmark's book at its configuration: 512 trades of seed 7 at 2^14-2^18 paths by product x 250 scenarios
    and the base, FP32, one launch. Properties: every value finite and non-negative; the European calls at the base
    scenario reprice their surface (local vol from the SSVI surface against Black-Scholes at its implied vol, within
    4 SE and the daily step's bias); a trade's numbers do not depend on the rest of the batch (a sample revalued alone
    with the same work split gives the same bits); the VaR is an actual scenario's loss."""
    from optlab.core.blackscholes import black_price

    book = synthetic_book()
    paths = np.array([t.n_paths for t in book.trades])
    assert paths.min() == 1 << 14 and paths.max() == 1 << 18
    scen = synthetic_scenarios(len(book.surfaces), n=250)
    print("benchmark book: 512 trades x 251 scenarios, FP32 batched launch", flush=True)
    rev = revalue(book, scen)
    assert rev.value.shape == (512, 251) and np.all(np.isfinite(rev.value)) and np.all(rev.value >= 0)
    assert np.all(rev.se >= 0) and np.all(np.isfinite(rev.se))
    gaps = []
    for i, tr in enumerate(book.trades):
        if tr.product != "european":
            continue
        a = tr.names[0]
        T = tr.n_steps * book.dt
        F, D = float(book.markets[a].F(T)), float(book.markets[a].D(T))
        K = tr.strike * float(book.markets[a].F(0.0))
        want = black_price(F, K, T, float(book.surfaces[a].iv(np.log(K / F), T)), D)
        gaps.append((rev.value[i, 0] - want, rev.se[i, 0]))
    gap, se = np.array(gaps).T
    print(f"European calls: {len(gap)}, max |gap| / SE {np.max(np.abs(gap) / se):.2f}, max |gap| {np.max(np.abs(gap)):.4f}",
          flush=True)
    assert np.all(np.abs(gap) <= 4 * se + 0.03)
    pick = [0, 1, 2, 3, 100, 257, 400, 511]
    sub = Book(book.surfaces, book.markets, book.corr, [book.trades[i] for i in pick], book.n_steps, book.dt)
    sub._device = book._device
    alone = revalue(sub, scen, block=rev.block, pairs_per_lane=rev.pairs_per_lane)
    np.testing.assert_array_equal(alone.value, rev.value[pick])
    np.testing.assert_array_equal(alone.se, rev.se[pick])
    r = var_es(pnl(book, rev))
    assert r["var"] > 0 and r["var"] == -pnl(book, rev)[r["var_index"]]


def test_benchmark_book_sampled_cells_are_the_oracle_in_fp64():
    """FP64 parity with the oracle on sampled (trade, scenario) cells of the benchmark book at 256 pairs: a trade of
    every product and an autocall on every name count present, under the base and two drawn scenarios."""
    book = synthetic_book()
    scen_all = synthetic_scenarios(len(book.surfaces), n=250)
    rows = [0, 17, 203]
    scen = Scenarios(scen_all.spot_shock[rows], scen_all.vol_shift[rows], scen_all.skew[rows])
    pick, seen = [], set()
    for i, tr in enumerate(book.trades):
        key = (tr.product, len(tr.names))
        if key not in seen:
            seen.add(key)
            pick.append(i)
    assert {k[0] for k in seen} == {"european", "asian", "barrier", "autocall"}
    from dataclasses import replace

    trades = [replace(book.trades[i], n_paths=512) for i in pick]
    sub = Book(book.surfaces, book.markets, book.corr, trades, book.n_steps, book.dt)
    print(f"benchmark book: {len(trades)} sampled trades x 3 scenarios in FP64 against the oracle", flush=True)
    _fp64_book_against_the_oracle("batched", 256, 1, sub, scen)
