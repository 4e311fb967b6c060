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


# --- models to tables: the plan layer (native/__init__.py:112-302) ------------------------------------------------
#
# A plan is the arrays of include/optlab_mi.h's path request that do not depend on the pairs: the time grid, the
# observation steps, each name's kind and parameters, the forwards, and the rows of every tabulated name stacked in
# one table (a name's steps point at its own rows). Local-variance rows are the Dupire variance at each step start
# on n_x points spanning [-8, 6] ATM standard deviations; SLV rows are the squared leverage resampled on n_x points.

SD_RANGE = (-8.0, 6.0)
_NO_FACTOR = (0.0, 1.0, 1.0, 0.0, 0.0)  # Heston slot of a name without a variance factor


class _RowCache:
    """Tables built from a surface or a leverage, kept for the last few (object, grid) keys while the object lives:
    a plan rebuilt on the same model and grid reuses its rows."""

    def __init__(self, size: int = 16):
        self.size = size
        self._entries: OrderedDict = OrderedDict()

    def get(self, obj, key, build):
        key = (id(obj), *key)
        entry = self._entries.get(key)
        if entry is None or entry[0] is not obj:
            entry = (obj, build())
            self._entries[key] = entry
            while len(self._entries) > self.size:
                self._entries.popitem(last=False)
        self._entries.move_to_end(key)
        return entry[1]


_ROWS = _RowCache()


def _surface_rows_by_row(surface, t0, n_x: int, sd_range):
    """Rows of a surface that only offers w(k, T) and local_var(k, T) at one maturity at a time."""
    T = np.maximum(np.asarray(t0, dtype=float), 1.0 / 365)
    ends = np.empty((len(T), 2))
    rows = np.empty((len(T), n_x))
    for i, Ti in enumerate(T):
        sd = float(np.sqrt(surface.w(0.0, float(Ti))))
        ends[i] = sd_range[0] * sd, sd_range[1] * sd
        rows[i] = surface.local_var(np.linspace(ends[i, 0], ends[i, 1], n_x), float(Ti))
    return ends[:, 0].copy(), (n_x - 1) / (ends[:, 1] - ends[:, 0]), rows


def local_var_rows(surface, t0: np.ndarray, n_x: int, sd_range=SD_RANGE):
    """(lo, inv_dx, rows) of the Dupire local variance at each step start (native/__init__.py:129-144). A surface
    with its own `local_var_rows` (optlab.core's SVI and SSVI: every row in one vectorised pass) builds them
    itself; any other is asked row by row."""
    t0 = np.asarray(t0, dtype=float)
    own = getattr(surface, "local_var_rows", None)
    build = (lambda: own(t0, n_x, sd_range)) if own is not None else (
        lambda: _surface_rows_by_row(surface, t0, n_x, sd_range))
    return _ROWS.get(surface, ("lv", n_x, t0.tobytes(), tuple(sd_range)), build)


def _leverage_curves(leverage):
    """(x_first, x_last, x, L) of each time row of a leverage: calibrated rows (x_grids, values), or uniform rows
    (x0, inv_dx, values) as fastmodels.UniformLeverage holds them."""
    if hasattr(leverage, "rows"):
        for x0, inv_dx, y in leverage.rows:
            y = np.asarray(y, dtype=float)
            x1 = x0 + (len(y) - 1) / inv_dx
            yield x0, x1, np.linspace(x0, x1, len(y)), y
    else:
        for x, y in zip(leverage.x_grids, leverage.values):
            yield x[0], x[-1], x, y


def leverage_rows(leverage, t0: np.ndarray, n_x: int):
    """(lo, inv_dx, squared leverage rows, row of each step) (native/__init__.py:147-167): each time row resampled
    linearly on n_x points between its first and last x, squared; step i reads the last time row at or before t0[i]
    (+1e-12)."""

    def build():
        curves = list(_leverage_curves(leverage))
        lo = np.array([c[0] for c in curves], dtype=float)
        hi = np.array([c[1] for c in curves], dtype=float)
        table = np.stack([np.interp(np.linspace(a, b, n_x), x, y) ** 2 for a, b, x, y in curves])
        times = np.asarray(leverage.times, dtype=float)
        which = np.searchsorted(times, np.asarray(t0) + 1e-12, side="right") - 1
        return lo, (n_x - 1) / (hi - lo), table, np.clip(which, 0, len(times) - 1).astype(np.int32)

    return _ROWS.get(leverage, ("lev", n_x, np.asarray(t0).tobytes()), build)


@dataclass
class _Name:
    """One name of a plan: its kind code, constant variance, Heston slot, and its table rows (lo, inv_dx, rows) with
    the row each step reads, when it reads a table."""

    kind: int
    cvar: float = 0.0
    factor: tuple = _NO_FACTOR
    rows: tuple | None = None
    step_rows: np.ndarray | None = None


def _name_of(model, t0: np.ndarray, n_x: int) -> _Name:
    if isinstance(model, BlackScholesModel):
        return _Name(_KIND["const"], cvar=model.vol**2)
    if isinstance(model, LocalVolModel):
        rows = local_var_rows(model.surface, t0, n_x, getattr(model, "sd_range", SD_RANGE))
        return _Name(_KIND["table"], rows=rows, step_rows=np.arange(len(t0), dtype=np.int32))
    if isinstance(model, (HestonModel, SLVModel)):
        h = model.heston
        factor = (h.v0, h.kappa, h.theta, h.xi, h.rho)
        if isinstance(model, HestonModel):
            return _Name(_KIND["heston"], factor=factor)
        lo, inv, table, step = leverage_rows(model.leverage, t0, n_x)
        return _Name(_KIND["slv"], factor=factor, rows=(lo, inv, table), step_rows=step)
    raise NotImplementedError(f"{type(model).__name__} is not in the path engine")


_KNOWN = (BlackScholesModel, LocalVolModel, SLVModel, HestonModel)


def _model_problem(model) -> str | None:
    if not isinstance(model, _KNOWN):
        return f"{type(model).__name__} is not in the path engine"
    if isinstance(model, SLVModel) and not (isinstance(model.leverage, Leverage) or hasattr(model.leverage, "rows")):
        return "this leverage is not in the path engine"
    return None


def unsupported(model) -> str | None:
    """None when the engine can simulate the model or basket, else why not (native/__init__.py:190-212). An
    addition: the stochastic mixing weight (`StochasticMix`, localcorr.py:87-108), which the reference engine
    refuses, runs here as mixer 2."""
    from optlab.structured.localcorr import LocalCorrelation, StochasticMix
    from optlab.structured.multiasset import Basket

    if not isinstance(model, Basket):
        return _model_problem(model)
    if len(model.names) > MAX_NAMES:
        return f"more than {MAX_NAMES} names"
    foreign = [type(model.models[n]).__name__ for n in model.names if not isinstance(model.models[n], _KNOWN)]
    if foreign:
        return f"{', '.join(foreign)} not in the path engine"
    mixer = model.mixer
    if mixer is None or isinstance(mixer, (LocalCorrelation, StochasticMix)):
        return None
    return f"{type(mixer).__name__} is not in the path engine (local correlation and StochasticMix only)"


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


def _members(model, market):
    """(names, models, markets, correlation root, same-variance-driver flag) of a model or a Basket."""
    from optlab.structured.multiasset import Basket

    if not isinstance(model, Basket):
        return ("_",), [model], [market], np.eye(1), 1
    names = tuple(model.names)
    return (names, [model.models[n] for n in names], [model.markets[n] for n in names],
            np.asarray(model.root, dtype=float), int(model.vol_corr == "same"))


def _mixer_fields(model):
    """(mixer code, the 8 mixer values, low-matrix factors) of a basket's mixer (none: code 0). Code 1, local
    correlation: (w0, level, down, up, sigma, market sign, cos, sin of the common factor's turn); code 2, the
    stochastic weight: (w0, theta, nu, rho, kappa, market sign, cos, sin)."""
    from optlab.structured.localcorr import StochasticMix, aligned_factors

    mx = getattr(model, "mixer", None)
    if mx is None:
        return 0, (0.0,) * 8, None
    low, sign, turn = aligned_factors(model.corr, model.root, mx.w0)
    turn_cs = (np.cos(turn), np.sin(turn))
    if isinstance(mx, StochasticMix):
        code, head = 2, (mx.w0, mx.theta, mx.nu, mx.rho, mx.kappa)
    else:
        code, head = 1, (mx.w0, mx.level, mx.down, mx.up, mx.sigma)
    return code, tuple(float(v) for v in (*head, sign, *turn_cs)), np.ascontiguousarray(low, dtype=float)


def plan(model, market, times, dt_max=1 / 252, dt_min=1 / 2000, n_x: int = N_X) -> Plan:
    """The simulation of one model (`models.simulate`) or of a Basket (`simulate_basket`) as the engine's arrays
    (native/__init__.py:243-302). A grid of Black-Scholes names only steps straight from one observation to the
    next (their step is exact)."""
    problem = unsupported(model)
    if problem:
        raise NotImplementedError(problem)
    times = np.asarray(times, dtype=float)
    names, models, markets, root, same = _members(model, market)
    exact = all(isinstance(m, BlackScholesModel) for m in models)
    grid = time_grid(times, np.inf, None) if exact else time_grid(times, dt_max, dt_min)
    dt, t0 = np.diff(grid), grid[:-1].copy()
    parts = [_name_of(m, t0, n_x) for m in models]
    tabulated = [q for q in parts if q.rows is not None]
    step_row = None
    if tabulated:
        # a name's steps point at its own rows in the stacked table
        step_row = np.zeros((len(parts), len(dt)), np.int32)
        base = 0
        for a, q in enumerate(parts):
            if q.rows is not None:
                step_row[a] = q.step_rows + base
                base += len(q.rows[0])
    fwd = np.zeros((len(parts), 0))
    if len(times):
        fwd = np.stack([np.asarray(mk.F(times), dtype=float) for mk in markets])
    mixer, mix, low = _mixer_fields(model)
    return Plan(
        names=names,
        kind=np.array([q.kind for q in parts], np.int32),
        cvar=np.array([q.cvar for q in parts], dtype=float),
        heston=np.ascontiguousarray([q.factor for q in parts], dtype=float),
        root=np.ascontiguousarray(root),
        same=same,
        dt=dt,
        t0=t0,
        obs_step=np.searchsorted(grid, times).astype(np.int32),
        fwd=np.ascontiguousarray(fwd),
        step_row=step_row,
        lo=np.concatenate([q.rows[0] for q in tabulated]) if tabulated else None,
        inv_dx=np.concatenate([q.rows[1] for q in tabulated]) if tabulated else None,
        table=np.ascontiguousarray(np.concatenate([q.rows[2] for q in tabulated])) if tabulated else None,
        mixer=mixer,
        mix=mix,
        low_root=low,
        n_x=n_x,
    )


def paths_request(p: Plan, n_pairs: int, seed: int, first_pair: int = 0, keep=None, windows=()):
    """(request, keep-alive arrays, n_keep) of a run: the optlab_paths_request of include/optlab_mi.h over the
    plan's arrays. `out` is left null; the caller points it at host memory or passes a device buffer."""
    kept = np.arange(p.n_obs) if keep is None else np.asarray(keep, dtype=int)
    spans = np.asarray(list(windows), dtype=np.int32).reshape(-1, 2)  # (first, last) observation of each window
    if len(spans) > MAX_WINDOWS:
        raise ValueError(f"{len(spans)} windows asked: a run follows at most {MAX_WINDOWS}")
    col_of = -np.ones(p.n_obs, np.int32)  # output column of each observation, -1 where it is not kept
    col_of[kept] = np.arange(kept.size, dtype=np.int32)
    win_lo, win_hi = np.ascontiguousarray(spans[:, 0]), np.ascontiguousarray(spans[:, 1])
    n_cols = kept.size + 2 * len(spans)
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
    rq.n_cols, rq.n_keep, rq.n_windows = n_cols, kept.size, len(spans)
    rq._keepalive = arrays  # the request points into these arrays: they live as long as it does
    return rq, arrays, kept.size


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
    ext = out[:, m:].reshape(k, -1, 2, 2 * n_pairs)  # (name, window, lowest / highest, row)
    extremes = [[(ext[a, w, 0], ext[a, w, 1]) for w in range(ext.shape[1])] for a in range(k)]
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
                     atm_vol=None, bandwidth=None, lv_source="rows"):
    """(request, keep-alive arrays, t0) of the particle method, as native/__init__.py:460-495 fills it. lv_source
    "rows": the Dupire rows built on the host (the reference's request); "slices": one slice record a step
    (`row_records` of optlab.core's surfaces) for the device to evaluate, in rq._slices (lv_rows stays null);
    "auto": slices where the surface has them."""
    regressions = ("nadaraya_watson", "local_linear")
    if regression not in regressions:
        raise ValueError(f"regression must be one of {regressions}, not {regression!r}")
    if lv_source not in ("rows", "slices", "auto"):
        raise ValueError(f"lv_source must be rows, slices or auto, not {lv_source!r}")
    steps = time_grid([T_max], dt_max, dt_min)
    dt, t0 = np.diff(steps), steps[:-1].copy()
    use_slices = lv_source == "slices" or (lv_source == "auto" and hasattr(surface, "row_records"))
    slices = None
    if use_slices:
        lv_lo, lv_hi, lv_inv, records = surface.row_records(t0, N_X)
        lv_rows = None
        slices = _abi.LvSlices()
        slices.n_rows, slices.floor, slices.cap = len(t0), 1e-4, 4.0
    else:
        lv_lo, lv_inv, lv_rows = local_var_rows(surface, t0, N_X)
    if atm_vol is None:  # the ATM vol at one year, or at the last pillar when the surface ends sooner
        horizon = min(1.0, float(surface.maturities[-1]))
        atm_vol = float(np.sqrt(surface.theta(horizon) / horizon))
    arrays = {
        "dt": np.ascontiguousarray(dt),
        "t0": np.ascontiguousarray(t0),
        "lv_lo": np.ascontiguousarray(lv_lo),
        "lv_inv_dx": np.ascontiguousarray(lv_inv),
        "lv_rows": None if lv_rows is None else np.ascontiguousarray(lv_rows),
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
    if slices is not None:
        arrays["lv_hi"] = np.ascontiguousarray(lv_hi, dtype=float)
        arrays["lv_records"] = np.ascontiguousarray(records, dtype=float)
        slices.hi, slices.records = _abi.ptr(arrays["lv_hi"]), _abi.ptr(arrays["lv_records"])
    rq._keepalive = arrays
    rq._options = leverage_options(bandwidth)
    rq._slices = slices
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
                                      bandwidth=bandwidth, lv_source="auto")
    if rq._slices is not None:  # optlab.core's surfaces: the device evaluates the rows' entries from slice records
        _check(lib, lib.optlab_leverage_slices(C.byref(rq), C.byref(rq._options), C.byref(rq._slices), prec))
    elif bandwidth is None:
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
