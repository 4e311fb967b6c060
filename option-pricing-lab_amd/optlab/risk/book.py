"""A book of trades revalued under scenarios on the MI355X engine in one persistent launch (optlab_book), and the
99 % VaR of its P&L.

    revaluation   packages/risk/src/optlab/risk/engine.py:600-645 (every scenario simulated on the trade's seed)
    products      packages/structured/src/optlab/structured/engine.py:199-330 on the path engine's spots: European,
                  arithmetic Asian, up-and-out on every close, worst-of autocall (the phoenix of compiler.py)
    VaR and ES    packages/risk/src/optlab/risk/var.py:240-255 (numpy's "lower" quantile: an actual scenario)
    moves         spot x (1 + shock) per name, implied vols + parallel + skew tilt (structured/surfaces.py:18-50),
                  local-variance tables rebuilt on the device (risk/surfaces.py:141-165)

`synthetic_book` draws the book of the benchmark: 512 trades from seed 7 (40 % European, 20 % Asian, 20 % daily
barrier, 20 % worst-of autocall on 1 to 5 names; maturities uniform in 0.25..5 y; 2^14 to 2^18 paths by product:
a power of two drawn per trade in 2^14..2^16 for Europeans, 2^15..2^16 for Asians, 2^16..2^17 for barriers and
2^17..2^18 for autocalls)
on 5 names with SSVI surfaces, and 250 scenarios (spot N(0, 2 %) per name, parallel vol N(0, 1.5 pt), skew tilt
N(0, 0.5 pt)) after the base scenario."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from optlab.core.engines import EngineError
from optlab.native import _abi, precision_code
from optlab.risk.surfaces import DeviceSurface
from optlab.structured.multiasset import correlation_root

LEVEL, ES_LEVEL = 0.99, 0.975
PRODUCTS = ("european", "asian", "barrier", "autocall")
STEPS_PER_YEAR = 252


@dataclass
class Trade:
    product: str
    names: tuple
    n_steps: int
    obs_every: int
    strike: float  # vanillas: in units of the spot; autocall: the put strike
    barrier: float = 0.0  # up-and-out level in units of the spot, or the autocall's knock-in level
    coupon: float = 0.02
    coupon_barrier: float = 0.7
    autocall: float = 1.0
    memory: bool = True
    quantity: float = 1.0
    n_paths: int = 1 << 14
    seed: int = 1


@dataclass
class Book:
    surfaces: list  # implied-vol surfaces of the names
    markets: list
    corr: np.ndarray  # correlation of the names
    trades: list
    n_steps: int = 5 * STEPS_PER_YEAR
    dt: float = 1.0 / STEPS_PER_YEAR
    _device: list = field(default_factory=list, repr=False)

    def device_surfaces(self) -> list:
        if not self._device:
            t0 = np.arange(self.n_steps) * self.dt
            self._device = [DeviceSurface(s, t0) for s in self.surfaces]
        return self._device


@dataclass
class Scenarios:
    spot_shock: np.ndarray  # (n, n_names) relative
    vol_shift: np.ndarray  # (n,) vol units
    skew: np.ndarray  # (n,)


def synthetic_book(seed: int = 7, n_trades: int = 512, n_names: int = 5, path_scale: int = 1) -> Book:
    """The benchmark's book (module docstring); `path_scale` divides every trade's paths (a smaller run of the same
    queue)."""
    from optlab.core.market import Market
    from optlab.core.ssvi import SSVI

    rng = np.random.default_rng(seed)
    pillars = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0, 5.0])
    scales = np.linspace(0.03, 0.05, n_names)
    rhos = np.linspace(-0.5, -0.7, n_names)
    surfaces = [SSVI.power_law(float(r), 1.2, 0.5, pillars, float(s) * pillars) for s, r in zip(scales, rhos)]
    markets = [Market.flat(100.0, 0.03, float(q)) for q in np.linspace(0.0, 0.02, n_names)]
    corr = np.full((n_names, n_names), 0.6) + 0.4 * np.eye(n_names)
    kinds = rng.choice(4, size=n_trades, p=[0.4, 0.2, 0.2, 0.2])
    log2_paths = {0: (14, 16), 1: (15, 16), 2: (16, 17), 3: (17, 18)}  # by product, both ends included
    trades = []
    for i, k in enumerate(kinds):
        T = float(rng.uniform(0.25, 5.0))
        qty = float(rng.choice([-1.0, 1.0]) * rng.integers(1, 11))
        lo2, hi2 = log2_paths[int(k)]
        n_paths = max(2, (1 << int(rng.integers(lo2, hi2 + 1))) // path_scale)
        if k == 3:
            every = STEPS_PER_YEAR // 4
            n_steps = max(1, round(T * 4)) * every
            m = int(rng.integers(1, min(5, n_names) + 1))
            names = tuple(sorted(rng.choice(n_names, size=m, replace=False).tolist()))
            trades.append(Trade("autocall", names, n_steps, every, 1.0, 0.6, 0.02, 0.7, 1.0, True, qty, n_paths, 1000 + i))
        elif k == 1:
            every = STEPS_PER_YEAR // 12
            n_steps = max(1, round(T * 12)) * every
            trades.append(Trade("asian", (int(rng.integers(n_names)),), n_steps, every, float(rng.uniform(0.9, 1.1)),
                                quantity=qty, n_paths=n_paths, seed=1000 + i))
        elif k == 2:
            n_steps = max(1, round(T * STEPS_PER_YEAR))
            trades.append(Trade("barrier", (int(rng.integers(n_names)),), n_steps, 1, float(rng.uniform(0.9, 1.1)),
                                float(rng.uniform(1.2, 1.5)), quantity=qty, n_paths=n_paths, seed=1000 + i))
        else:
            n_steps = max(1, round(T * STEPS_PER_YEAR))
            trades.append(Trade("european", (int(rng.integers(n_names)),), n_steps, 1,
                                float(rng.uniform(0.8, 1.2)), quantity=qty, n_paths=n_paths, seed=1000 + i))
    return Book(surfaces, markets, corr, trades)


def synthetic_scenarios(n_names: int, n: int = 250, seed: int = 7) -> Scenarios:
    """The base scenario then n draws: spot N(0, 2 %) per name, parallel vol N(0, 1.5 pt), skew tilt N(0, 0.5 pt)."""
    rng = np.random.default_rng(seed + 1)
    spot = np.vstack([np.zeros(n_names), rng.normal(0.0, 0.02, (n, n_names))])
    vol = np.concatenate([[0.0], rng.normal(0.0, 0.015, n)])
    skew = np.concatenate([[0.0], rng.normal(0.0, 0.005, n)])
    return Scenarios(spot, vol, skew)


@dataclass
class Revaluation:
    value: np.ndarray  # (n_trades, n_scen)
    se: np.ndarray
    finish_ms: np.ndarray  # per trade
    tables_ms: float
    kernel_ms: float
    total_ms: float
    name_steps: int
    block: int
    pairs_per_lane: int


def revalue(book: Book, scen: Scenarios, backend: str = "gpu", mode: str = "batched", block: int = 0,
            pairs_per_lane: int = 0) -> Revaluation:
    """Every trade under every scenario: mode "batched" (one persistent launch draining the work queue, captured in a
    HIP graph with its reset and reduction), "direct" (the same launch without the graph) or "loop" (one launch per
    trade)."""
    lib = _abi.load()
    dev = book.device_surfaces()
    n_names, ns = len(dev), book.n_steps
    t = np.arange(ns + 1) * book.dt
    fwd = np.ascontiguousarray(np.stack([np.asarray(m.F(t), dtype=float) for m in book.markets]))
    disc = np.ascontiguousarray(np.asarray(book.markets[0].D(t), dtype=float))
    handles = np.array([d.handle for d in dev], dtype=np.int32)
    trades = (_abi.Trade * len(book.trades))()
    for c, tr in zip(trades, book.trades):
        c.product = PRODUCTS.index(tr.product)
        c.n_names = len(tr.names)
        for k, a in enumerate(tr.names):
            c.names[k] = a
        sub = book.corr[np.ix_(tr.names, tr.names)]
        root = correlation_root(sub) if len(tr.names) > 1 else np.ones((1, 1))
        for k, v in enumerate(np.asarray(root, dtype=float).ravel()):
            c.root[k] = v
        c.n_steps, c.obs_every, c.strike, c.barrier = tr.n_steps, tr.obs_every, tr.strike, tr.barrier
        c.coupon, c.coupon_barrier, c.autocall, c.memory = tr.coupon, tr.coupon_barrier, tr.autocall, int(tr.memory)
        c.quantity, c.n_pairs, c.seed = tr.quantity, tr.n_paths // 2, tr.seed
    spot = np.ascontiguousarray(scen.spot_shock, dtype=float)
    vol = np.ascontiguousarray(scen.vol_shift, dtype=float)
    skew = np.ascontiguousarray(scen.skew, dtype=float)
    n_scen = len(vol)
    value = np.empty((len(book.trades), n_scen))
    se = np.empty_like(value)
    finish = np.empty(len(book.trades))
    rq = _abi.BookRequest()
    rq.n_names, rq.n_steps, rq.dt = n_names, ns, book.dt
    rq.surface, rq.fwd, rq.disc = _abi.ptr(handles, _abi._PI), _abi.ptr(fwd), _abi.ptr(disc)
    rq.n_trades, rq.trades = len(book.trades), trades
    rq.n_scen, rq.spot_shock, rq.vol_shift, rq.skew = n_scen, _abi.ptr(spot), _abi.ptr(vol), _abi.ptr(skew)
    rq.mode = {"batched": 0, "loop": 1, "direct": 2}[mode]
    rq.block, rq.pairs_per_lane = block, pairs_per_lane
    rq.value, rq.se, rq.finish_ms = _abi.ptr(value), _abi.ptr(se), _abi.ptr(finish)
    code = lib.optlab_book(C.byref(rq), precision_code(backend))
    if code:
        raise EngineError(lib.optlab_last_error().decode())
    return Revaluation(value, se, finish, rq.tables_ms, rq.kernel_ms, rq.total_ms, rq.name_steps, rq.block_used,
                       rq.pairs_per_lane_used)


def pnl(book: Book, rev: Revaluation) -> np.ndarray:
    """P&L of each scenario against scenario 0 (the base), weighted by the trades' quantities."""
    q = np.array([t.quantity for t in book.trades])
    return q @ (rev.value[:, 1:] - rev.value[:, :1])


def var_es(pnl_values, level: float = LEVEL, es_level: float = ES_LEVEL) -> dict:
    """VaR (the loss at `level`, positive; numpy's "lower" quantile, an actual scenario) and ES (the mean loss beyond
    the `es_level` quantile): var.py:240-255."""
    x = np.asarray(pnl_values, dtype=float)
    q = float(np.quantile(x, 1 - level, method="lower"))
    qe = float(np.quantile(x, 1 - es_level, method="lower"))
    return {"var": -q, "es": -float(x[x <= qe].mean()), "n": int(len(x)), "var_index": int(np.flatnonzero(x == q)[0])}
