"""The fixed products on the HIP engine (packages/native/src/optlab/native/gpu.py:1-223), plus the strip.

European calls and puts (Black-Scholes one exact step; Heston ceil(T / dt_max) uniform QE-M steps), Black-Scholes
Asians on n equally spaced fixings (arithmetic, geometric, and -- an addition -- the arithmetic with the geometric
control variate), single barriers monitored on n equally spaced dates (down-and-out, and up-and-out as an addition),
and `strip`: a grid of vanillas over strikes and maturities on one QE-M (or Black-Scholes) simulation.

`backend`: "gpu" (FP32 paths, compensated FP64 sums), "cpu" or "fp64" (the same kernels in FP64: the reference CPU
engine's arithmetic on Philox draws), "auto" ("gpu"). The reference's "numpy" fallback is not part of this engine:
asking for it, or for a case the engine does not cover, raises instead of silently pricing elsewhere."""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from optlab.core.engines import EngineError
from optlab.core.market import Market
from optlab.core.models import BlackScholesModel, HestonModel
from optlab.native import _abi

_LEG = {"call": 1, "put": 2}
_MODEL_BS, _MODEL_HESTON = 0, 1
_EUROPEAN, _ASIAN, _BARRIER_DOWN, _BARRIER_UP = 0, 1, 2, 3


@dataclass
class EngineResult:
    """price, stderr, n_paths, the backend that ran and its kernel time (native/gpu.py:36-41)."""

    price: float
    stderr: float
    n_paths: int
    backend: str = "gpu"
    compute_ms: float = math.nan


def _ext():
    try:
        return _abi.load()
    except OSError:
        return None


def engine_available() -> bool:
    return _ext() is not None


def gpu_available() -> bool:
    ext = _ext()
    return ext is not None and ext.optlab_gpu_count() > 0


def gpu_name() -> str:
    ext = _ext()
    if ext is None:
        return ""
    buf = C.create_string_buffer(512)
    ext.optlab_gpu_name(buf, 512)
    return buf.value.decode()


def _unsupported(model, market: Market, T: float, cp: str, why: str | None):
    if why:
        return why
    if not isinstance(model, (BlackScholesModel, HestonModel)):
        return f"{type(model).__name__} is not in the engine (Black-Scholes and Heston only)"
    if cp not in _LEG:
        return f"cp must be 'call' or 'put', not {cp!r}"
    if market.flat_rates(T) is None:
        return "the engine needs a flat market (constant r and q)"
    return None


def _choose(backend: str, reason: str | None) -> tuple[str, int]:
    """(backend name, precision code). Unlike the reference there is no NumPy fallback to choose."""
    if backend not in ("auto", "gpu", "cpu", "fp64", "fp32"):
        raise ValueError(f"backend must be 'auto', 'gpu', 'cpu', 'fp64' or 'fp32', not {backend!r}")
    if reason is not None:
        raise ValueError(f"backend {backend!r}: {reason}")
    if not engine_available():
        raise EngineError(f"the HIP engine ({_abi.library_path()}) is not built: run build.py")
    if not gpu_available():
        raise EngineError("no HIP device")
    name = "gpu" if backend == "auto" else backend
    return name, 0 if name in ("cpu", "fp64") else 1


def _check(ext, status: int):
    if status != _abi.OK:
        msg = (ext.optlab_last_error() or b"").decode(errors="replace")
        raise (ValueError if status == _abi.EINVAL else EngineError)(msg)


def request(model, market, T, cp, product, n_steps, n_paths, seed, K, H=0.0) -> _abi.Request:
    """The optlab_request of a fixed product (native/gpu.py:97-128)."""
    rates = market.flat_rates(T)
    assert rates is not None  # noqa: S101
    r, q = rates
    rq = _abi.Request()
    rq.model, rq.product = (_MODEL_HESTON if isinstance(model, HestonModel) else _MODEL_BS), int(product)
    rq.legs[0], rq.legs[1], rq.legs[2] = _LEG[cp], 0, 0
    rq.strikes[0] = rq.strikes[1] = rq.strikes[2] = float(K)
    rq.H, rq.n_steps = float(H), int(n_steps)
    rq.S0, rq.r, rq.q, rq.T = float(market.spot), float(r), float(q), float(T)
    if isinstance(model, HestonModel):
        p = model.params
        rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho = p.v0, p.kappa, p.theta, p.xi, p.rho
    else:
        rq.sigma = float(model.vol)
    rq.n_paths = max(4, n_paths + n_paths % 2)
    rq.seed = (seed if seed is not None else int(np.random.SeedSequence().entropy % 2**64)) & (2**64 - 1)
    return rq


def price(rq: _abi.Request, backend: str = "gpu") -> _abi.Result:
    """optlab_price on a filled request: the four outputs with their standard errors."""
    name, prec = _choose(backend, None)
    ext = _ext()
    out = _abi.Result()
    _check(ext, ext.optlab_price(C.byref(rq), prec, 0, C.byref(out)))
    return out


def _result(out: _abi.Result, k: int, backend: str) -> EngineResult:
    return EngineResult(out.price[k], out.se[k], int(out.n_paths), backend, out.kernel_ms)


def european(model, market: Market, K, T, cp="call", n_paths=1_000_000, seed: int | None = None, dt_max=1 / 252,
             backend="auto") -> EngineResult:
    name, _ = _choose(backend, _unsupported(model, market, T, cp, None))
    n_steps = 1 if isinstance(model, BlackScholesModel) else max(1, math.ceil(T / dt_max - 1e-9))
    return _result(price(request(model, market, T, cp, _EUROPEAN, n_steps, n_paths, seed, K), name), 0, name)


def asian(model, market: Market, K, T, n_fixings=12, cp="call", average="arithmetic", n_paths=1_000_000,
          seed: int | None = None, backend="auto") -> EngineResult:
    """average: "arithmetic", "geometric" or "control" (the arithmetic with the geometric control variate)."""
    if average not in ("arithmetic", "geometric", "control"):
        raise ValueError(f"average must be 'arithmetic', 'geometric' or 'control', not {average!r}")
    why = None if isinstance(model, BlackScholesModel) else "Asian: Black-Scholes only in the engine"
    name, _ = _choose(backend, _unsupported(model, market, T, cp, why))
    out = price(request(model, market, T, cp, _ASIAN, n_fixings, n_paths, seed, K), name)
    return _result(out, {"arithmetic": 0, "geometric": 1, "control": 2}[average], name)


def barrier(model, market: Market, K, H, T, kind="down-out", cp="call", n_monitor=252, n_paths=1_000_000,
            seed: int | None = None, backend="auto") -> EngineResult:
    why = None
    if kind not in ("down-out", "up-out"):
        why = f"barrier: down-out or up-out only in the engine, not {kind!r}"
    elif kind == "down-out" and not 0 < H < market.spot:
        why = "barrier: the engine needs 0 < H < spot"
    elif kind == "up-out" and not H > market.spot:
        why = "barrier: up-out needs H > spot"
    name, _ = _choose(backend, _unsupported(model, market, T, cp, why))
    product = _BARRIER_DOWN if kind == "down-out" else _BARRIER_UP
    return _result(price(request(model, market, T, cp, product, n_monitor, n_paths, seed, K, H), name), 0, name)


@dataclass
class Strip:
    """price[j, k] and se[j, k] at maturity j and strike k; forward[j] = E[S_{t_j}] with its se."""

    strikes: np.ndarray
    maturities: np.ndarray
    price: np.ndarray
    se: np.ndarray
    forward: np.ndarray
    forward_se: np.ndarray
    kernel_ms: float
    total_ms: float
    backend: str


def strip_request(model, market: Market, strikes, mat_steps, n_steps: int, T: float, n_paths: int, seed: int,
                  cp="call", psi_c=1.5):
    strikes = np.ascontiguousarray(strikes, dtype=float)
    mat_steps = np.ascontiguousarray(mat_steps, dtype=np.int32)
    rates = market.flat_rates(T)
    if rates is None:
        raise ValueError("the engine needs a flat market (constant r and q)")
    r, q = rates
    rq = _abi.StripRequest()
    rq.model = _MODEL_HESTON if isinstance(model, HestonModel) else _MODEL_BS
    rq.put = int(cp == "put")
    rq.n_steps, rq.S0, rq.r, rq.q, rq.T = int(n_steps), float(market.spot), float(r), float(q), float(T)
    if isinstance(model, HestonModel):
        p = model.params
        rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho = p.v0, p.kappa, p.theta, p.xi, p.rho
    else:
        rq.sigma = float(model.vol)
    rq.psi_c = float(psi_c)
    rq.n_strikes, rq.strikes = len(strikes), _abi.ptr(strikes)
    rq.n_mats, rq.mat_step = len(mat_steps), _abi.ptr(mat_steps, _abi._PI)
    rq.n_paths, rq.seed = int(n_paths) + int(n_paths) % 2, int(seed) & (2**64 - 1)
    width = len(strikes) + 1
    out = {"price": np.empty((len(mat_steps), width)), "se": np.empty((len(mat_steps), width)),
           "strikes": strikes, "mat_step": mat_steps}
    rq.price, rq.se = _abi.ptr(out["price"]), _abi.ptr(out["se"])
    rq._keepalive = out  # the request points into these arrays: they live as long as it does
    return rq, out


def _strip_steps(maturities, T, dt_max):
    """(maturities, T, n_steps, the step of each maturity) of a strip of uniform steps T / n_steps."""
    maturities = np.asarray(maturities, dtype=float)
    T = float(maturities.max() if T is None else T)
    n_steps = max(1, math.ceil(T / dt_max - 1e-9))
    steps = maturities / T * n_steps
    mat_steps = np.rint(steps).astype(np.int32)
    if np.any(np.abs(steps - mat_steps) > 1e-6):
        raise ValueError("every maturity must fall on a step of T / n_steps")
    return maturities, T, n_steps, mat_steps


def strip(model, market: Market, strikes, maturities, T: float | None = None, dt_max=1 / 252, n_paths=1_000_000,
          seed: int = 0, cp="call", psi_c=1.5, backend="auto") -> Strip:
    """Vanillas at every (maturity, strike) on one simulation of uniform steps T / n_steps; every maturity must fall
    on a step."""
    maturities, T, n_steps, mat_steps = _strip_steps(maturities, T, dt_max)
    name, prec = _choose(backend, _unsupported(model, market, T, cp, None))
    rq, out = strip_request(model, market, strikes, mat_steps, n_steps, T, n_paths, seed, cp, psi_c)
    ext = _ext()
    _check(ext, ext.optlab_strip(C.byref(rq), prec))
    k = len(out["strikes"])
    return Strip(out["strikes"], maturities, out["price"][:, :k], out["se"][:, :k], out["price"][:, k],
                 out["se"][:, k], rq.kernel_ms, rq.total_ms, name)


@dataclass
class StripGreeks:
    """Pathwise sensitivities of a strip's prices (index [maturity, strike]) on its own paths: delta = d price / d S0
    and vega = d price / d `parameter` ("v0" under Heston, "sigma" under Black-Scholes), each with its standard error;
    forward_* are those of E[S_{t_j}]."""

    strikes: np.ndarray
    maturities: np.ndarray
    parameter: str
    delta: np.ndarray
    delta_se: np.ndarray
    vega: np.ndarray
    vega_se: np.ndarray
    forward_delta: np.ndarray
    forward_vega: np.ndarray
    kernel_ms: float
    total_ms: float
    backend: str


def strip_greeks(model, market: Market, strikes, maturities, T: float | None = None, dt_max=1 / 252,
                 n_paths=1_000_000, seed: int = 0, cp="call", psi_c=1.5, backend="auto") -> StripGreeks:
    """`strip`'s sensitivities on the same paths (same seed, steps and draws): the spot's by scaling the paths, the
    variance parameter's by a tangent carried through every step (optlab_strip_greeks)."""
    maturities, T, n_steps, mat_steps = _strip_steps(maturities, T, dt_max)
    name, prec = _choose(backend, _unsupported(model, market, T, cp, None))
    rq, out = strip_request(model, market, strikes, mat_steps, n_steps, T, n_paths, seed, cp, psi_c)
    shape = out["price"].shape
    d, dse, g, gse = (np.empty(shape) for _ in range(4))
    ext = _ext()
    _check(ext, ext.optlab_strip_greeks(C.byref(rq), prec, _abi.ptr(d), _abi.ptr(dse), _abi.ptr(g), _abi.ptr(gse)))
    k = len(out["strikes"])
    return StripGreeks(out["strikes"], maturities, "v0" if isinstance(model, HestonModel) else "sigma", d[:, :k],
                       dse[:, :k], g[:, :k], gse[:, :k], d[:, k], g[:, k], rq.kernel_ms, rq.total_ms, name)
