"""The MI355X path engine behind optlab's engine-by-name interface (packages/native/src/optlab/native/__init__.py).

Same functions, arguments and outputs as the reference module: `plan` turns a model (or a Basket) into arrays,
`run` simulates pairs of it, `simulate` is `models.simulate` on the engine, `calibrate_leverage` is the SLV particle
method. Draws are Philox4x32-10 keyed by the seed and counted by (pair, stream, rank), so a path depends on
(seed, pair index) only -- not on batches or launches -- and pair p gives rows p (Z) and n_pairs + p (-Z).

Backends: every simulation runs on the GPU through hand-written HIP kernels (liboptlab_mi.so, C ABI in
include/optlab_mi.h). "gpu" is the reference's float engine (FP32 path state); "cpu", the reference's double
engine, is served by the double-precision mode of the same kernels, which reproduces the reference CPU engine's
numbers; "fp64"/"fp32" name the two modes directly; "auto" is "gpu". There is no CPU fallback: without the built
library or a device, every call raises EngineError.
"""

from __future__ import annotations

import ctypes as C
import os
from collections import OrderedDict
from dataclasses import asdict, dataclass
from typing import Any

import numpy as np

from optlab.core.engines import EngineError, default_engine, seed64
from optlab.core.models import (
    BlackScholesModel,
    HestonModel,
    HestonParams,
    Leverage,
    LocalVolModel,
    SLVModel,
    time_grid,
)
from optlab.native import _abi

__all__ = [
    "MAX_NAMES",
    "MAX_WINDOWS",
    "N_X",
    "Device",
    "Plan",
    "available",
    "calibrate_leverage",
    "cpu_threads",
    "devices",
    "gpu_available",
    "plan",
    "run",
    "simulate",
    "summary",
    "unsupported",
]

_KIND = {"const": 0, "table": 1, "heston": 2, "slv": 3}
N_X = 801
MAX_NAMES = 8
MAX_WINDOWS = 2
_PRECISION = {"cpu": 0, "fp64": 0, "gpu64": 0, "gpu": 1, "fp32": 1, "auto": 1}


def _lib():
    """The HIP engine library, or None when it is not built."""
    try:
        return _abi.load()
    except OSError:
        return None


def _need_lib():
    lib = _lib()
    if lib is None:
        raise EngineError(f"the HIP engine ({_abi.library_path()}) is not built: run build.py")
    return lib


def _check(lib, status: int) -> None:
    if status != _abi.OK:
        msg = (lib.optlab_last_error() or b"").decode(errors="replace")
        if status == _abi.EINVAL:
            raise ValueError(msg)
        raise EngineError(msg or f"engine error {status}")


def precision_code(backend: str) -> int:
    """0 (double) or 1 (float) for a backend name."""
    if backend not in _PRECISION:
        raise ValueError(f"backend must be cpu, gpu, auto, fp64 or fp32, not {backend!r}")
    return _PRECISION[backend]


def available() -> bool:
    """The engine library is built (it still needs a device to run)."""
    return _lib() is not None


def gpu_available() -> bool:
    lib = _lib()
    return lib is not None and lib.optlab_gpu_count() > 0


@dataclass(frozen=True)
class Device:
    name: str
    compute_capability: str
    sm_count: int
    clock_mhz: int
    memory_gb: float
    free_gb: float
    integrated: bool
    watchdog: bool


def devices() -> list[Device]:
    lib = _lib()
    if lib is None:
        return []
    out = []
    for i in range(lib.optlab_gpu_count()):
        d = _abi.Device()
        if lib.optlab_device_info(i, C.byref(d)) == _abi.OK:
            out.append(Device(d.name.decode(), f"{d.major}.{d.minor}", d.sm_count, d.clock_mhz, d.total_mem / 2**30,
                              d.free_mem / 2**30, bool(d.integrated), bool(d.kernel_timeout)))
    return out


def cpu_threads() -> int:
    lib = _lib()
    return int(lib.optlab_cpu_threads()) if lib is not None else (os.cpu_count() or 1)


# --- models to tables (native/__init__.py:112-187) ---------------------------------------------------------------

_TABLES: OrderedDict = OrderedDict()


def _cached(key, obj, build):
    hit = _TABLES.get(key)
    if hit is not None and hit[0] is obj:
        _TABLES.move_to_end(key)
        return hit[1]
    rows = build()
    _TABLES[key] = (obj, rows)
    while len(_TABLES) > 16:
        _TABLES.popitem(last=False)
    return rows


def local_var_rows(surface, t0: np.ndarray, n_x: int, sd_range=(-8.0, 6.0)):
    """(lo, inv_dx, rows): the Dupire local variance at each step start on n_x uniform points of x spanning
    sd_range ATM standard deviations sqrt(w(0, T)), T = max(t, 1 day) (native/__init__.py:129-144)."""

    def build():
        if hasattr(surface, "local_var_rows"):
            return surface.local_var_rows(t0, n_x, sd_range)
        lo, inv, rows = np.empty(len(t0)), np.empty(len(t0)), np.empty((len(t0), n_x))
        for i, t in enumerate(t0):
            T = max(float(t), 1.0 / 365)
            sd = float(np.sqrt(surface.w(0.0, T)))
            a, b = sd_range[0] * sd, sd_range[1] * sd
            lo[i], inv[i] = a, (n_x - 1) / (b - a)
            rows[i] = surface.local_var(np.linspace(a, b, n_x), T)
        return lo, inv, rows

    return _cached((id(surface), "lv", n_x, t0.tobytes(), tuple(sd_range)), surface, build)


def leverage_rows(leverage, t0: np.ndarray, n_x: int):
    """(lo, inv_dx, squared leverage rows, row of each step) (native/__init__.py:147-167)."""

    def build():
        times = np.asarray(leverage.times, dtype=float)
        if hasattr(leverage, "rows"):
            src = [(x0, x0 + (len(y) - 1) / inv, np.linspace(x0, x0 + (len(y) - 1) / inv, len(y)), y)
                   for x0, inv, y in leverage.rows]
        else:
            src = [(xs[0], xs[-1], xs, vs) for xs, vs in zip(leverage.x_grids, leverage.values)]
        lo, inv, rows = np.empty(len(src)), np.empty(len(src)), np.empty((len(src), n_x))
        for r, (a, b, xs, vs) in enumerate(src):
            grid = np.linspace(a, b, n_x)
            lo[r], inv[r] = a, (n_x - 1) / (b - a)
            rows[r] = np.interp(grid, xs, vs) ** 2
        step = np.clip(np.searchsorted(times, t0 + 1e-12, side="right") - 1, 0, len(times) - 1).astype(np.int32)
        return lo, inv, rows, step

    return _cached((id(leverage), "lev", n_x, t0.tobytes()), leverage, build)


def _describe(model, t0: np.ndarray, n_x: int):
    if isinstance(model, BlackScholesModel):
        return _KIND["const"], model.vol**2, (0.0, 1.0, 1.0, 0.0, 0.0), None, None
    if isinstance(model, LocalVolModel):
        sd = getattr(model, "sd_range", (-8.0, 6.0))
        lo, inv, rows = local_var_rows(model.surface, t0, n_x, sd)
        return _KIND["table"], 0.0, (0.0, 1.0, 1.0, 0.0, 0.0), (lo, inv, rows), np.arange(len(t0), dtype=np.int32)
    if isinstance(model, SLVModel):
        h = model.heston
        lo, inv, rows, step = leverage_rows(model.leverage, t0, n_x)
        return _KIND["slv"], 0.0, (h.v0, h.kappa, h.theta, h.xi, h.rho), (lo, inv, rows), step
    if isinstance(model, HestonModel):
        h = model.heston
        return _KIND["heston"], 0.0, (h.v0, h.kappa, h.theta, h.xi, h.rho), None, None
    raise NotImplementedError(f"{type(model).__name__} is not in the path engine")


_KNOWN = (BlackScholesModel, LocalVolModel, SLVModel, HestonModel)


def unsupported(model) -> str | None:
    """Why the engine cannot simulate this model or basket, or None (native/__init__.py:190-212). An addition: the
    stochastic mixing weight (`StochasticMix`, localcorr.py:87-108), which the reference engine refuses, runs here."""
    from optlab.structured.localcorr import LocalCorrelation, StochasticMix
    from optlab.structured.multiasset import Basket

    if isinstance(model, Basket):
        if len(model.names) > MAX_NAMES:
            return f"more than {MAX_NAMES} names"
        bad = [n for n, m in model.models.items() if not isinstance(m, _KNOWN)]
        if bad:
            return f"{', '.join(type(model.models[n]).__name__ for n in bad)} not in the path engine"
        if model.mixer is not None and not isinstance(model.mixer, (LocalCorrelation, StochasticMix)):
            return f"{type(model.mixer).__name__} is not in the path engine (local correlation and StochasticMix only)"
        return None
    if not isinstance(model, _KNOWN):
        return f"{type(model).__name__} is not in the path engine"
    if isinstance(model, SLVModel) and not isinstance(model.leverage, Leverage) and not hasattr(model.leverage, "rows"):
        return "this leverage is not in the path engine"
    return None


@dataclass
class Plan:
    """Everything of a simulation but the pairs: built once, run on any range of pairs."""

    names: tuple
    kind: np.ndarray
    cvar: np.ndarray
    heston: np.ndarray
    root: np.ndarray
    same: int
    dt: np.ndarray
    t0: np.ndarray
    obs_step: np.ndarray
    fwd: np.ndarray
    step_row: np.ndarray | None
    lo: np.ndarray | None
    inv_dx: np.ndarray | None
    table: np.ndarray | None
    mixer: int
    mix: tuple
    low_root: np.ndarray | None
    n_x: int

    @property
    def n_obs(self) -> int:
        return len(self.obs_step)

    @property
    def n_steps(self) -> int:
        return len(self.dt)


def plan(model, market, times, dt_max=1 / 252, dt_min=1 / 2000, n_x: int = N_X) -> Plan:
    """The simulation of one model or a Basket as arrays for the engine (native/__init__.py:243-302)."""
    from optlab.structured.multiasset import Basket

    why = unsupported(model)
    if why:
        raise NotImplementedError(why)
    times = np.asarray(times, dtype=float)
    if isinstance(model, Basket):
        names, models = tuple(model.names), [model.models[n] for n in model.names]
        markets = [model.markets[n] for n in model.names]
        root, same = np.asarray(model.root, dtype=float), int(model.vol_corr == "same")
    else:
        names, models, markets, root, same = ("_",), [model], [market], np.eye(1), 1
    if all(isinstance(m, BlackScholesModel) for m in models):
        dt_max, dt_min = np.inf, None
    grid = time_grid(times, dt_max, dt_min)
    obs = np.searchsorted(grid, times).astype(np.int32)
    dt, t0 = np.diff(grid), grid[:-1].copy()
    k = len(models)
    kind, cvar, hes = np.empty(k, np.int32), np.empty(k), np.empty((k, 5))
    lo, inv, tab, rows_of = [], [], [], np.zeros((k, len(dt)), np.int32)
    offset = 0
    for a, m in enumerate(models):
        kind[a], cvar[a], hes[a], rows, step = _describe(m, t0, n_x)
        if rows is not None:
            lo.append(rows[0]), inv.append(rows[1]), tab.append(rows[2])
            rows_of[a] = step + offset
            offset += len(rows[0])
    fwd = np.stack([np.asarray(mk.F(times), dtype=float) for mk in markets]) if len(times) else np.zeros((k, 0))
    mixer, mix, low = 0, (0.0,) * 8, None
    if isinstance(model, Basket) and model.mixer is not None:
        from optlab.structured.localcorr import aligned_factors

        from optlab.structured.localcorr import StochasticMix

        mx = model.mixer
        low, sign, turn = aligned_factors(model.corr, model.root, mx.w0)
        if isinstance(mx, StochasticMix):  # mixer 2: (w0, theta, nu, rho, kappa, market sign, cos, sin of the turn)
            mixer, mix = 2, (mx.w0, mx.theta, mx.nu, mx.rho, mx.kappa, sign, np.cos(turn), np.sin(turn))
        else:
            mixer, mix = 1, (mx.w0, mx.level, mx.down, mx.up, mx.sigma, sign, np.cos(turn), np.sin(turn))
        low = np.ascontiguousarray(low, dtype=float)
    return Plan(
        names,
        kind,
        cvar,
        np.ascontiguousarray(hes),
        np.ascontiguousarray(root),
        same,
        dt,
        t0,
        obs,
        np.ascontiguousarray(fwd),
        rows_of if tab else None,
        np.concatenate(lo) if lo else None,
        np.concatenate(inv) if inv else None,
        np.ascontiguousarray(np.concatenate(tab)) if tab else None,
        mixer,
        tuple(float(x) for x in mix),
        low,
        n_x,
    )


def paths_request(p: Plan, n_pairs: int, seed: int, first_pair: int = 0, keep=None, windows=()):
    """(request, keep-alive arrays, n_keep) of a run: the optlab_paths_request of include/optlab_mi.h over the
    plan's arrays. `out` is left null; the caller points it at host memory or passes a device buffer."""
    keep = np.arange(p.n_obs) if keep is None else np.asarray(keep, dtype=int)
    if len(windows) > MAX_WINDOWS:
        raise ValueError(f"at most {MAX_WINDOWS} windows")
    col_of = np.full(p.n_obs, -1, np.int32)
    col_of[keep] = np.arange(len(keep), dtype=np.int32)
    win_lo = np.array([w[0] for w in windows], np.int32)
    win_hi = np.array([w[1] for w in windows], np.int32)
    n_cols = len(keep) + 2 * len(windows)
    arrays = {
        "dt": np.ascontiguousarray(p.dt, dtype=float),
        "t0": np.ascontiguousarray(p.t0, dtype=float),
        "obs_step": np.ascontiguousarray(p.obs_step, dtype=np.int32),
        "root": np.ascontiguousarray(p.root, dtype=float),
        "kind": np.ascontiguousarray(p.kind, dtype=np.int32),
        "cvar": np.ascontiguousarray(p.cvar, dtype=float),
        "heston": np.ascontiguousarray(p.heston, dtype=float),
        "step_row": None if p.step_row is None else np.ascontiguousarray(p.step_row, dtype=np.int32),
        "row_lo": None if p.lo is None else np.ascontiguousarray(p.lo, dtype=float),
        "row_inv_dx": None if p.inv_dx is None else np.ascontiguousarray(p.inv_dx, dtype=float),
        "table": None if p.table is None else np.ascontiguousarray(p.table, dtype=float),
        "fwd": np.ascontiguousarray(p.fwd, dtype=float),
        "low_root": None if p.low_root is None else np.ascontiguousarray(p.low_root, dtype=float),
        "col_of": col_of,
        "win_lo": win_lo,
        "win_hi": win_hi,
    }
    rq = _abi.PathsRequest()
    rq.n_names, rq.n_steps, rq.n_obs, rq.n_x = len(p.names), len(p.dt), p.n_obs, int(p.n_x)
    rq.n_rows = 0 if p.lo is None else len(p.lo)
    rq.n_pairs, rq.first_pair, rq.seed = int(n_pairs), int(first_pair), int(seed) & (2**64 - 1)
    ints = ("obs_step", "kind", "step_row", "col_of", "win_lo", "win_hi")
    for name, a in arrays.items():
        setattr(rq, name, _abi.ptr(a, _abi._PI if name in ints else _abi._PD))
    rq.same, rq.mixer = int(p.same), int(p.mixer)
    for i, v in enumerate(p.mix):
        rq.mix[i] = float(v)
    rq.n_cols, rq.n_keep, rq.n_windows = n_cols, len(keep), len(windows)
    rq._keepalive = arrays  # the request points into these arrays: they live as long as it does
    return rq, arrays, len(keep)


def run(p: Plan, n_pairs: int, seed: int, first_pair: int = 0, backend: str = "gpu", threads: int = 0, keep=None,
        windows=()):
    """(spots, extremes, timing) as native/__init__.py:317-374: spots[a] is a (2 n_pairs, len(keep)) Fortran-order
    view of name a at the kept observations, extremes[a][w] the (lowest, highest) spot over window w."""
    lib = _need_lib()
    prec = precision_code(backend)
    rq, keepalive, m = paths_request(p, n_pairs, seed, first_pair, keep, windows)
    k = len(p.names)
    out = np.empty((k, rq.n_cols, 2 * n_pairs))
    rq.out = _abi.ptr(out)
    _check(lib, lib.optlab_paths(C.byref(rq), prec, int(threads)))
    del keepalive
    spots = [out[a, :m].T for a in range(k)]
    extremes = [[(out[a, m + 2 * w], out[a, m + 2 * w + 1]) for w in range(len(windows))] for a in range(k)]
    name = "gpu" if backend == "auto" else backend
    return spots, extremes, {"kernel_ms": rq.kernel_ms, "total_ms": rq.total_ms, "backend": name,
                             "precision": "fp64" if prec == 0 else "fp32"}


def run_device(p: Plan, n_pairs: int, seed: int, first_pair: int = 0, backend: str = "gpu", keep=None, windows=(),
               out=None, stream=None):
    """`run` with the output left in HBM: a torch float64 CUDA tensor (names, n_cols, 2 n_pairs). `out` may be a
    preallocated tensor of that shape; `stream` a torch.cuda.Stream (default: the current stream)."""
    import torch

    lib = _need_lib()
    prec = precision_code(backend)
    rq, keepalive, _ = paths_request(p, n_pairs, seed, first_pair, keep, windows)
    shape = (len(p.names), rq.n_cols, 2 * n_pairs)
    if out is None:
        out = torch.empty(shape, dtype=torch.float64, device="cuda")
    elif tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float64 CUDA tensor of shape {shape}")
    st = stream if stream is not None else torch.cuda.current_stream()
    _check(lib, lib.optlab_paths_dev(C.byref(rq), prec, C.c_void_p(out.data_ptr()), C.c_void_p(st.cuda_stream)))
    del keepalive
    return out, {"kernel_ms": rq.kernel_ms, "total_ms": rq.total_ms}


def simulate(model, market, times, n_paths: int, seed=None, dt_max=1 / 252, dt_min=1 / 2000, backend: str = "auto",
             first_pair: int = 0, threads: int = 0):
    """`models.simulate` (a {name: array} dict for a Basket) on the engine (native/__init__.py:407-426)."""
    from optlab.structured.multiasset import Basket

    p = plan(model, market, times, dt_max, dt_min)
    spots, _, _ = run(p, n_paths // 2, seed64(seed), first_pair, backend, threads)
    if isinstance(model, Basket):
        return dict(zip(p.names, spots))
    return spots[0]


@dataclass
class LSMResult:
    """american_lsm's result: MCResult's (price, stderr, n_paths) (products.py:16-24), the regression coefficients
    of each exercise date (NaN where the date exercises nothing) and the device timings."""

    price: float
    stderr: float
    n_paths: int
    coefs: np.ndarray
    timing: dict

    def __iter__(self):
        return iter((self.price, self.stderr))


def american_lsm(model, market, K, T, cp="put", n_exercise=50, n_paths=100_000, rng=None, dt_max=1 / 252,
                 backend: str = "auto", min_itm: int = 50) -> LSMResult:
    """`core.products.american_lsm` (products.py:112-169) on the engine: Longstaff-Schwartz with the basis
    (1, s, s^2, s^3) in s = S / K, plus (v, v^2, s v) when the model has a variance factor, regressing on the
    in-the-money paths only; the rule is estimated on the first n_paths / 2 pairs of the seed and applied to the next
    n_paths / 2 (an independent set: a lower bound up to Monte Carlo error). The regressions run on the device with
    FP64 normal equations (the reference: NumPy lstsq); the paths are the engine's (Philox), in the backend's
    precision."""
    lib = _need_lib()
    prec = precision_code(backend)
    times = np.linspace(T / n_exercise, T, n_exercise)
    p = plan(model, market, times, dt_max)
    if len(p.names) != 1:
        raise ValueError("american_lsm prices an option on one name")
    pr, keepalive, _ = paths_request(p, n_paths // 2, seed64(rng))
    rq = _abi.LsmRequest()
    rq.paths = pr
    disc = np.ascontiguousarray(market.D(times), dtype=float)
    coefs = np.empty((n_exercise, 7))
    rq.strike, rq.call = float(K), int(str(cp).lower().startswith("c"))
    rq.disc, rq.spot = _abi.ptr(disc), float(market.spot)
    rq.basis_var, rq.min_itm = int(bool(model.has_variance)), int(min_itm)
    rq.coefs = _abi.ptr(coefs)
    _check(lib, lib.optlab_lsm(C.byref(rq), prec))
    del keepalive
    return LSMResult(rq.price, rq.se, n_paths, coefs,
                     {"paths_ms": rq.paths_ms, "regression_ms": rq.regression_ms, "apply_ms": rq.apply_ms,
                      "total_ms": rq.total_ms, "precision": "fp64" if prec == 0 else "fp32"})


@dataclass
class RegressResult:
    """regress's result: the coefficients (None when the normal matrix is not positive definite), the fitted
    continuation on eval_state (None without one) and the device timings."""

    coef: np.ndarray | None
    values: np.ndarray | None
    timing: dict


def regress(state, y, mean, scale, powers, knots=(), eval_state=None, backend: str = "auto") -> RegressResult:
    """The issuer's call regression of exercise.py:191-238 on the device: the least-squares fit of y on the terms of
    the state (exercise.py:150-171, Rule.features): z = (state - mean) / scale, polynomial terms prod_i z_i^p_i for
    each exponent tuple of `powers` (0 to 3 each), then hinges max(z_v - k, 0) for each (v, k) of `knots`; FP64
    normal equations summed in a fixed order and a Cholesky solve (the reference: np.linalg.lstsq). With
    `eval_state`, the fitted continuation there (Rule.continuation). `backend` is accepted for symmetry: both
    compute in FP64."""
    lib = _need_lib()
    prec = precision_code(backend)
    x = np.ascontiguousarray(np.atleast_2d(np.asarray(state, dtype=float).T).T)
    n, d = x.shape
    yv = np.ascontiguousarray(y, dtype=float)
    if yv.shape != (n,):
        raise ValueError("y must have one value per state row")
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mean, dtype=float), (d,)))
    sc = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, dtype=float), (d,)))
    pw = np.ascontiguousarray(np.asarray(powers, dtype=np.int32).reshape(-1, d))
    kv = np.ascontiguousarray([int(v) for v, _ in knots], dtype=np.int32)
    kl = np.ascontiguousarray([float(k) for _, k in knots], dtype=float)
    coef = np.zeros(len(pw) + len(kl))
    rq = _abi.RegressRequest()
    rq.n_rows, rq.n_vars = n, d
    rq.state, rq.y, rq.mean, rq.scale = _abi.ptr(x), _abi.ptr(yv), _abi.ptr(mu), _abi.ptr(sc)
    rq.n_powers, rq.powers = len(pw), _abi.ptr(pw, _abi._PI)
    rq.n_knots, rq.knot_var, rq.knot = len(kl), _abi.ptr(kv, _abi._PI), _abi.ptr(kl)
    rq.coef = _abi.ptr(coef)
    ev = out = None
    if eval_state is not None:
        ev = np.ascontiguousarray(np.atleast_2d(np.asarray(eval_state, dtype=float).T).T)
        if ev.shape[1] != d:
            raise ValueError("eval_state must have the state's variables")
        out = np.empty(len(ev))
        rq.n_eval, rq.eval_state, rq.eval_out = len(ev), _abi.ptr(ev), _abi.ptr(out)
    _check(lib, lib.optlab_regress(C.byref(rq), prec))
    return RegressResult(coef if rq.ok else None, out if rq.ok else None,
                         {"fit_ms": rq.fit_ms, "eval_ms": rq.eval_ms, "total_ms": rq.total_ms})


def leverage_request(params: HestonParams, surface, T_max: float, mix: float, n_particles: int, dt_max, dt_min,
                     n_x, bandwidth_scale, rng, l_bounds, regression, positive_part, n_bins=400, n_fine=8192,
                     atm_vol=None, bandwidth=None):
    """(request, keep-alive arrays, t0) of the particle method, as native/__init__.py:460-495 fills it."""
    if regression not in ("nadaraya_watson", "local_linear"):
        raise ValueError(f"unknown regression {regression!r}, expected nadaraya_watson or local_linear")
    grid = time_grid([T_max], dt_max, dt_min)
    dt, t0 = np.diff(grid), grid[:-1].copy()
    lv_lo, lv_inv, lv_rows = local_var_rows(surface, t0, N_X)
    if atm_vol is None:
        T1 = min(1.0, float(surface.maturities[-1]))
        atm_vol = float(np.sqrt(surface.theta(T1) / T1))
    arrays = {
        "dt": np.ascontiguousarray(dt),
        "t0": np.ascontiguousarray(t0),
        "lv_lo": np.ascontiguousarray(lv_lo),
        "lv_inv_dx": np.ascontiguousarray(lv_inv),
        "lv_rows": np.ascontiguousarray(lv_rows),
        "x_grids": np.empty((len(t0), n_x)),
        "values": np.empty((len(t0), n_x)),
    }
    rq = _abi.LeverageRequest()
    rq.n_steps, rq.n_pairs, rq.seed = len(t0), int(n_particles) // 2, seed64(rng)
    rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho = params.v0, params.kappa, params.theta, mix * params.xi, params.rho
    rq.n_x, rq.n_bins, rq.n_fine = int(n_x), int(n_bins), int(n_fine)
    rq.atm_vol, rq.bandwidth_scale = float(atm_vol), float(bandwidth_scale)
    rq.l_min, rq.l_max, rq.q_lo, rq.q_hi = float(l_bounds[0]), float(l_bounds[1]), 0.0005, 0.9995
    rq.local_linear, rq.positive_part = int(regression == "local_linear"), int(bool(positive_part))
    rq.n_lv = N_X
    for name, a in arrays.items():
        setattr(rq, name, _abi.ptr(a))
    rq._keepalive = arrays
    rq._options = leverage_options(bandwidth)
    return rq, arrays, t0


def leverage_options(bandwidth=None) -> _abi.LeverageOptions:
    """None: the reference's bandwidth (bandwidth_scale atm_vol sqrt(max(t, 0.15)) N^-0.2); ("std", c): c std(ln S)
    N^-0.2 over the particles at each step."""
    opt = _abi.LeverageOptions()
    if bandwidth is None:
        opt.bandwidth_rule, opt.bandwidth_coef = 0, 0.0
    elif isinstance(bandwidth, tuple) and len(bandwidth) == 2 and bandwidth[0] == "std":
        opt.bandwidth_rule, opt.bandwidth_coef = 1, float(bandwidth[1])
    else:
        raise ValueError(f"bandwidth must be None or ('std', coef), not {bandwidth!r}")
    return opt


def calibrate_leverage(params, surface, market, T_max: float, mix: float = 1.0, n_particles: int = 100_000,
                       dt_max=1 / 100, dt_min=1 / 2000, n_x=81, bandwidth_scale=1.0, rng=None, l_bounds=(0.05, 8.0),
                       regression="nadaraya_watson", positive_part=False, backend: str = "auto", threads: int = 0,
                       *, n_bins: int = 400, bandwidth=None):
    """`slv.calibrate_leverage` on the engine (native/__init__.py:429-496): 400 regression bins, quantiles from an
    8192-bin histogram, local variance from uniform rows, Philox draws. Additions (keyword only): `n_bins` (at most
    512) and `bandwidth=("std", c)` for a bandwidth of c std(ln S) N^-0.2; dt_min=None gives a uniform grid."""
    lib = _need_lib()
    prec = precision_code(backend)
    rq, arrays, t0 = leverage_request(params, surface, T_max, mix, n_particles, dt_max, dt_min, n_x,
                                      bandwidth_scale, rng, l_bounds, regression, positive_part, n_bins=n_bins,
                                      bandwidth=bandwidth)
    if bandwidth is None:
        _check(lib, lib.optlab_leverage(C.byref(rq), prec, int(threads)))
    else:
        _check(lib, lib.optlab_leverage_ex(C.byref(rq), C.byref(rq._options), prec))
    return SLVModel(params, Leverage(t0, arrays["x_grids"], arrays["values"]), mix)


def summary() -> dict[str, Any]:
    lib = _lib()
    return {
        "engine": default_engine(),
        "compiled": lib is not None,
        "hip_build": bool(lib is not None and lib.optlab_has_hip()),
        "devices": [asdict(d) for d in devices()],
        "cpu_threads": cpu_threads(),
        "resolved": "numpy" if lib is None or default_engine() == "numpy" else ("gpu" if gpu_available() else "none"),
    }
