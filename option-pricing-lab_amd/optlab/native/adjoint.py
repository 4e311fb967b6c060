"""Worst-of autocall (Phoenix) under Dupire local vol on the MI355X engine: price, pathwise adjoint Greeks (delta per
name, vega per node of the risk grid per surface) carried by a torch.autograd.Function, and the bump-and-revalue
comparator on common random numbers.

    product      packages/structured/src/optlab/structured/compiler.py:114-161, 270-295 and presets.py:24-31
                 (phoenix: quarterly autocall at 100 %, coupon barrier 70 % with memory, put knocked in at 60 %),
                 the knock-in read on every close (downside.observation = "daily"); every comparison smoothed by a
                 ramp of relative width `smoothing` (structured/script/evaluator.py:93-108, greeks.py:175)
    bump Greeks  packages/structured/src/optlab/structured/greeks.py:164-313 (central differences; a spot bump
                 scales the paths at sticky moneyness, so deltas need no new surface) and
                 packages/risk/src/optlab/risk/greeks.py:199-244 with risk/surfaces.py:95-170 (node vegas: the
                 surface bumped by one tent of the grid, its local-variance table rebuilt, the paths simulated
                 again on the same seed)
    paths        packages/native/cuda/include/optlab_cuda/paths.hpp:274-361 (local vol, n names, stream 0)

Everything runs on the current HIP device through liboptlab_mi.so: the surfaces' rows are uploaded once
(`optlab.risk.surfaces.DeviceSurface`), every bumped table is built on the device, and no step of the simulation or of
the adjoint runs on the host."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from optlab.core.engines import EngineError
from optlab.native import _abi, precision_code
from optlab.risk import grid
from optlab.risk.surfaces import DeviceSurface
from optlab.structured.multiasset import correlation_root

N_BATCHES = 16
FWD_BLOCK = 256


@dataclass(frozen=True)
class Phoenix:
    """The phoenix preset (presets.py:24-31) with the coupon rate's default (schema.py:219-229) and a daily knock-in."""

    autocall: float = 1.0
    coupon_barrier: float = 0.7
    coupon_rate: float = 0.08  # per annum
    frequency_months: int = 3
    maturity_months: int = 36
    memory: bool = True
    ki_barrier: float = 0.6
    put_strike: float = 1.0
    smoothing: float = 0.005


@dataclass
class Greeks:
    price: float
    se: float
    delta: np.ndarray  # dV / dS_a per unit of spot
    delta_se: np.ndarray
    vega: np.ndarray  # (n_names, n_strikes, n_tenors): dV per vol point of each node
    vega_se: np.ndarray
    n_paths: int
    seconds: dict = field(default_factory=dict)


def _batch_legs(n_legs: int) -> np.ndarray:
    """The legs in each of the 16 price batches (consecutive forward workgroups of 256 legs)."""
    blocks = -(-n_legs // FWD_BLOCK)
    edges = [min(n_legs, (b * blocks // N_BATCHES) * FWD_BLOCK) for b in range(N_BATCHES + 1)]
    return np.diff(edges).astype(float)


def _batch_se(values: np.ndarray, legs: np.ndarray) -> tuple[float, float]:
    """Mean and standard error by batch means, batches weighted by their legs."""
    w = legs / legs.sum()
    used = legs > 0
    mean = float(np.sum(w * values))
    m = int(used.sum())
    var = float(np.sum((w[used] * (values[used] - mean)) ** 2)) * m / max(m - 1, 1)
    return mean, float(np.sqrt(var))


class WorstOfAutocall:
    """A worst-of Phoenix on len(surfaces) names (1 to 5): local vol from each name's implied surface, constant
    correlation, uniform steps (`steps_per_year`), performances against `fixings` (default: the spots)."""

    def __init__(self, surfaces, markets, corr, product: Phoenix = Phoenix(), steps_per_year: int = 252,
                 n_x: int = 801, fixings=None):
        n = len(surfaces)
        if not 1 <= n <= 5 or len(markets) != n:
            raise ValueError("1 to 5 names, one market each")
        self.product = product
        T = product.maturity_months / 12
        self.n_steps = int(round(T * steps_per_year))
        self.dt = T / self.n_steps
        t = np.arange(self.n_steps + 1) * self.dt
        self.times = t
        self.surfaces = [DeviceSurface(s, t[:-1], n_x) for s in surfaces]
        self.root = np.ascontiguousarray(correlation_root(np.asarray(corr, dtype=float)), dtype=float)
        self.fwd = np.ascontiguousarray(np.stack([np.asarray(m.F(t), dtype=float) for m in markets]))
        self.spots = self.fwd[:, 0].copy()
        self.fixing = np.ascontiguousarray(self.spots if fixings is None else np.asarray(fixings, dtype=float))
        n_periods = product.maturity_months // product.frequency_months
        self.obs_step = np.array([round(j * self.n_steps / n_periods) for j in range(1, n_periods + 1)], dtype=np.int32)
        self.disc = np.ascontiguousarray(np.asarray(markets[0].D(t[self.obs_step]), dtype=float))
        self.coupon = product.coupon_rate * product.frequency_months / 12
        self.n_names = n
        self._lib = _abi.load()

    # ---- requests ------------------------------------------------------------------------------------------------
    def request(self, n_paths: int, seed: int, first_pair: int = 0, nodes=None, spot=None) -> _abi.AutocallRequest:
        if n_paths < 2 or n_paths % 2:
            raise ValueError("n_paths must be even and >= 2")
        rq = _abi.AutocallRequest()
        handles = np.array([s.handle for s in self.surfaces], dtype=np.int32)
        keep = [handles]
        rq.n_names, rq.n_steps, rq.dt = self.n_names, self.n_steps, self.dt
        rq.surface = _abi.ptr(handles, _abi._PI)
        rq.root, rq.fwd, rq.fixing = _abi.ptr(self.root), _abi.ptr(self.fwd), _abi.ptr(self.fixing)
        rq.n_obs, rq.obs_step, rq.disc = len(self.obs_step), _abi.ptr(self.obs_step, _abi._PI), _abi.ptr(self.disc)
        p = self.product
        rq.autocall, rq.coupon_barrier, rq.coupon, rq.memory = p.autocall, p.coupon_barrier, self.coupon, int(p.memory)
        rq.ki_barrier, rq.put_strike, rq.smoothing = p.ki_barrier, p.put_strike, p.smoothing
        rq.n_pairs, rq.first_pair, rq.seed = n_paths // 2, first_pair, seed
        if nodes is not None:
            h = np.ascontiguousarray(nodes, dtype=float).reshape(self.n_names, _abi.GRID_NODES)
            keep.append(h)
            rq.nodes = _abi.ptr(h)
        if spot is not None:
            f = np.ascontiguousarray(spot, dtype=float).reshape(self.n_names)
            keep.append(f)
            rq.spot = _abi.ptr(f)
        rq._keepalive = keep
        return rq

    def _call(self, fn, rq) -> None:
        code = fn(C.byref(rq), rq._prec)
        if code:
            raise EngineError(self._lib.optlab_last_error().decode())

    def price(self, n_paths: int, seed: int, backend: str = "gpu", nodes=None, spot=None, first_pair: int = 0):
        """(price, se, the 16 batch prices, timings): one forward simulation."""
        rq = self.request(n_paths, seed, first_pair, nodes, spot)
        rq._prec = precision_code(backend)
        self._call(self._lib.optlab_autocall, rq)
        return rq.price, rq.se, np.array(rq.batch[:]), {"forward_ms": rq.forward_ms, "total_ms": rq.total_ms}

    def adjoint(self, n_paths: int, seed: int, backend: str = "gpu", first_pair: int = 0,
                checkpoint_every: int = 0) -> Greeks:
        """Price, deltas and every node vega from one forward and one adjoint pass. `checkpoint_every`: steps
        between checkpoints of the forward state (4, 8 up to 3 names, 16 for one name); 0 lets the engine's
        autotuner choose per shape (timings["checkpoint_every"] is the interval used)."""
        rq = self.request(n_paths, seed, first_pair)
        rq._prec = precision_code(backend)
        rq.checkpoint_every = int(checkpoint_every)
        shape = (self.n_names, _abi.GRID_STRIKES, _abi.GRID_TENORS)
        vega, vega_se = np.empty(shape), np.empty(shape)
        rq.vega, rq.vega_se = _abi.ptr(vega), _abi.ptr(vega_se)
        self._call(self._lib.optlab_adjoint, rq)
        n = self.n_names
        return Greeks(rq.price, rq.se, np.array(rq.delta[:n]), np.array(rq.delta_se[:n]), vega, vega_se, n_paths,
                      {"forward_ms": rq.forward_ms, "backward_ms": rq.backward_ms, "total_ms": rq.total_ms,
                       "checkpoint_every": rq.checkpoint_used})

    def bump(self, n_paths: int, seed: int, backend: str = "gpu", h_spot: float = 0.01, h_vol: float = 0.01,
             nodes=None) -> Greeks:
        """Central differences on common random numbers (greeks.py:164-313): per name a spot bump of +-h_spot
        (relative), and per node (all of them, or the (name, i, j) listed in `nodes`) a tent bump of +-h_vol vol.
        Vegas per vol point; standard errors by batch means of the CRN differences."""
        import time

        prec_name = backend
        legs = _batch_legs(n_paths)
        t0 = time.perf_counter()
        base, base_se, _, _ = self.price(n_paths, seed, prec_name)
        delta, delta_se = np.zeros(self.n_names), np.zeros(self.n_names)
        for a in range(self.n_names):
            f = np.ones(self.n_names)
            f[a] = 1 + h_spot
            _, _, up, _ = self.price(n_paths, seed, prec_name, spot=f)
            f[a] = 1 - h_spot
            _, _, dn, _ = self.price(n_paths, seed, prec_name, spot=f)
            delta[a], delta_se[a] = _batch_se((up - dn) / (2 * h_spot * self.spots[a]), legs)
        shape = (self.n_names, _abi.GRID_STRIKES, _abi.GRID_TENORS)
        vega, vega_se = np.full(shape, np.nan), np.full(shape, np.nan)
        todo = [(a, i, j) for a in range(self.n_names) for i in range(shape[1]) for j in range(shape[2])]
        if nodes is not None:
            todo = [tuple(x) for x in nodes]
        for a, i, j in todo:
            H = np.zeros(shape)
            H[a, i, j] = h_vol
            _, _, up, _ = self.price(n_paths, seed, prec_name, nodes=H)
            _, _, dn, _ = self.price(n_paths, seed, prec_name, nodes=-H)
            vega[a, i, j], vega_se[a, i, j] = _batch_se((up - dn) * 0.01 / (2 * h_vol), legs)
        seconds = {"total_s": time.perf_counter() - t0, "revaluations": 1 + 2 * self.n_names + 2 * len(todo)}
        return Greeks(base, base_se, delta, delta_se, vega, vega_se, n_paths, seconds)

    def close(self) -> None:
        for s in self.surfaces:
            s.close()


def autocall_function():
    """The torch.autograd.Function of the price in the spots and the node bumps: forward runs the engine's
    forward and adjoint passes (the Greeks come at the price of one backward pass), backward hands the stored
    Greeks on. PyTorch only carries the gradients; every number is computed by the HIP engine."""
    import torch

    class AutocallPrice(torch.autograd.Function):
        @staticmethod
        def forward(ctx, spots, nodes, note: WorstOfAutocall, n_paths: int, seed: int, backend: str):
            if torch.any(nodes != 0):
                raise EngineError("adjoint Greeks are taken at the unbumped surfaces: nodes must be zero")
            factors = (spots.detach().double().cpu().numpy() / note.spots).astype(float)
            rq = note.request(n_paths, seed, spot=factors)
            rq._prec = precision_code(backend)
            shape = (note.n_names, _abi.GRID_STRIKES, _abi.GRID_TENORS)
            vega = np.empty(shape)
            rq.vega = _abi.ptr(vega)
            note._call(note._lib.optlab_adjoint, rq)
            delta = torch.tensor(np.array(rq.delta[: note.n_names]), dtype=spots.dtype, device=spots.device)
            dvdh = torch.tensor(vega / 0.01, dtype=nodes.dtype, device=nodes.device)  # per unit of vol
            ctx.save_for_backward(delta, dvdh)
            return torch.tensor(rq.price, dtype=spots.dtype, device=spots.device)

        @staticmethod
        def backward(ctx, grad):
            delta, dvdh = ctx.saved_tensors
            return grad * delta, grad * dvdh, None, None, None, None

    return AutocallPrice
