"""Worst-of autocall (Phoenix) on the MI355X engine, each name under Dupire local vol (the default), SLV, Heston or
Black-Scholes: price, pathwise adjoint Greeks (delta per name, vega per node of the risk grid per surface) carried by
a torch.autograd.Function, and the bump-and-revalue comparator on common random numbers.

    product      packages/structured/src/optlab/structured/compiler.py:114-161, 270-295 and presets.py:24-31
                 (phoenix: quarterly autocall at 100 %, coupon barrier 70 % with memory, put knocked in at 60 %),
                 the knock-in read on every close (downside.observation = "daily"); every comparison smoothed by a
                 ramp of relative width `smoothing` (structured/script/evaluator.py:93-108, greeks.py:175)
    bump Greeks  packages/structured/src/optlab/structured/greeks.py:164-313 (central differences; a spot bump
                 scales the paths at sticky moneyness, so deltas need no new surface) and
                 packages/risk/src/optlab/risk/greeks.py:199-244 with risk/surfaces.py:95-170 (node vegas: the
                 surface bumped by one tent of the grid, its local-variance table rebuilt, the paths simulated
                 again on the same seed)
    paths        packages/native/cuda/include/optlab_cuda/paths.hpp:196-361 (local vol, n names, stream 0; Heston
                 and SLV variances by Euler with full truncation, their normals on the aux stream when a name is
                 SLV)
    SLV vegas    an SLV name's node vegas hold the conditional expectation E[v | x] = LV / L^2 of its calibration
                 (packages/native/cuda/src/leverage.cpp:42-97 divides the local variance by it): a surface bump moves
                 the squared leverage to L^2 LV_bumped / LV at every entry of the table, with no new calibration
                 (the reference recalibrates: structured/greeks.py:164-313, `slv_vega`)

Everything runs on the current HIP device through liboptlab_mi.so: the surfaces' rows are uploaded once
(`optlab.risk.surfaces.DeviceSurface`), every bumped table is built on the device, and no step of the simulation or of
the adjoint runs on the host."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from optlab.core.engines import EngineError
from optlab.core.models import BlackScholesModel, HestonModel, LocalVolModel, SLVModel
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
    correlation, uniform steps (`steps_per_year`), performances against `fixings` (default: the spots).

    `models` (optional, one a name; None keeps local vol): an SLVModel (its leverage resampled on the surface's
    table grid, row of step i the leverage's last time row at or before t_i), a HestonModel or a
    BlackScholesModel. `same`: the variance normals correlated by `corr` as the spot normals."""

    def __init__(self, surfaces, markets, corr, product: Phoenix = Phoenix(), steps_per_year: int = 252,
                 n_x: int = 801, fixings=None, models=None, same: bool = False):
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
        self.models = self._models(models, same) if models is not None else None

    # ---- the names' models ---------------------------------------------------------------------------------------
    def _lev2_rows(self, leverage, a: int) -> np.ndarray:
        """The squared leverage on name a's table grid, (n_steps, n_x)."""
        from optlab.native import _leverage_curves

        curves = list(_leverage_curves(leverage))
        times = np.asarray(leverage.times, dtype=float)
        which = np.clip(np.searchsorted(times, self.times[:-1] + 1e-12, side="right") - 1, 0, len(times) - 1)
        k = self.surfaces[a].k()
        out = np.empty_like(k)
        for i, w in enumerate(which):
            _, _, x, y = curves[w]
            out[i] = np.interp(k[i], x, y) ** 2
        return out

    def _models(self, models, same: bool) -> _abi.AutocallModels:
        n = self.n_names
        if len(models) != n:
            raise ValueError("one model a name")
        kind, cvar, heston = np.ones(n, np.int32), np.zeros(n), np.zeros((n, 5))
        lev2 = None
        for a, m in enumerate(models):
            if m is None or isinstance(m, LocalVolModel):
                continue
            if isinstance(m, BlackScholesModel):
                kind[a], cvar[a] = 0, m.vol**2
            elif isinstance(m, (HestonModel, SLVModel)):
                h = m.heston  # SLV: vol-of-vol scaled by the mix
                heston[a] = (h.v0, h.kappa, h.theta, h.xi, h.rho)
                kind[a] = 2
                if isinstance(m, SLVModel):
                    kind[a] = 3
                    if lev2 is None:
                        lev2 = np.zeros((n, self.n_steps, self.surfaces[a].n_x))
                    lev2[a] = self._lev2_rows(m.leverage, a)
            else:
                raise EngineError(f"{type(m).__name__} is not a model of the autocall engine")
        out = _abi.AutocallModels()
        keep = [kind, cvar, heston] + ([lev2] if lev2 is not None else [])
        out.kind, out.cvar, out.heston = _abi.ptr(kind, _abi._PI), _abi.ptr(cvar), _abi.ptr(heston)
        out.lev2 = _abi.ptr(lev2) if lev2 is not None else None
        out.same = int(same)
        out._keepalive = keep
        self.model_arrays = {"kind": kind, "cvar": cvar, "heston": heston, "lev2": lev2, "same": bool(same)}
        return out

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

    def _call(self, adjoint: bool, rq) -> None:
        models = C.byref(self.models) if self.models is not None else None
        code = self._lib.optlab_autocall_ex(C.byref(rq), models, int(adjoint), rq._prec)
        if code:
            raise EngineError(self._lib.optlab_last_error().decode())

    def price(self, n_paths: int, seed: int, backend: str = "gpu", nodes=None, spot=None, first_pair: int = 0):
        """(price, se, the 16 batch prices, timings): one forward simulation."""
        rq = self.request(n_paths, seed, first_pair, nodes, spot)
        rq._prec = precision_code(backend)
        self._call(False, rq)
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
        self._call(True, rq)
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
            note._call(True, rq)
            delta = torch.tensor(np.array(rq.delta[: note.n_names]), dtype=spots.dtype, device=spots.device)
            dvdh = torch.tensor(vega / 0.01, dtype=nodes.dtype, device=nodes.device)  # per unit of vol
            ctx.save_for_backward(delta, dvdh)
            return torch.tensor(rq.price, dtype=spots.dtype, device=spots.device)

        @staticmethod
        def backward(ctx, grad):
            delta, dvdh = ctx.saved_tensors
            return grad * delta, grad * dvdh, None, None, None, None

    return AutocallPrice
