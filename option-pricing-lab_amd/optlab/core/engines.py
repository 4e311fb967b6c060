"""Engine selection by name, as the reference's registry exposes it (packages/core/src/optlab/core/engines.py:1-137).

The interface kept from the reference: the names "numpy", "cpu", "gpu" and "auto"; the entry point group
`optlab.engines` under which a compiled engine registers as "native"; the `Engine` protocol; `EngineError`; the
functions `compiled`, `set_default`, `default_engine`, `choose`, `require`; and `seed64`, which turns a seed into
the engine's 64-bit Philox key.

How this build resolves them: the compiled engine is the HIP engine `optlab.native`, found through the entry point
when a distribution registers one, else imported from the source tree. Both of its modes run on the GPU -- "gpu" is
its FP32 mode and "cpu" its FP64 mode (the reference CPU engine's numbers) -- so any named engine other than
"numpy" needs a device. "numpy" stays a name only: the reference's NumPy pricers are not part of this build."""

from __future__ import annotations

import os
from importlib.metadata import entry_points
from typing import Any, Protocol, cast

import numpy as np

ENGINES = ("numpy", "cpu", "gpu", "auto")
GROUP = "optlab.engines"


class EngineError(RuntimeError):
    """The compiled engine failed, or cannot run what was asked of it by name."""


class Engine(Protocol):
    MAX_WINDOWS: int

    def available(self) -> bool: ...

    def gpu_available(self) -> bool: ...

    def cpu_threads(self) -> int: ...

    def unsupported(self, model) -> str | None: ...

    def plan(self, model, market, times, dt_max: float = ..., dt_min: float = ..., n_x: int = ...) -> Any: ...

    def run(self, p, n_pairs: int, seed: int, first_pair: int = 0, backend: str = "cpu", threads: int = 0,
            keep=None, windows=()) -> tuple: ...

    def simulate(self, model, market, times, n_paths: int, seed=None, dt_max=..., dt_min=..., backend: str = "auto") -> Any: ...

    def calibrate_leverage(self, params, surface, market, T_max: float, *args, **kwargs) -> Any: ...


def _engine_name(name: str) -> str:
    key = str(name).strip().lower()
    if key not in ENGINES:
        raise ValueError(f"unknown engine {name!r}: the engine names are {', '.join(ENGINES)}")
    return key


class _Registry:
    """The process's engine preference and the compiled engine, looked up once on first use."""

    def __init__(self) -> None:
        self.preferred: str | None = None
        self._looked_up = False
        self._engine: Engine | None = None

    def engine(self) -> Engine | None:
        if not self._looked_up:
            self._engine = self._find()
            self._looked_up = True
        return self._engine

    @staticmethod
    def _find() -> Engine | None:
        # A registered engine that fails to import raises here: a broken install must not pass for "none".
        for ep in entry_points(group=GROUP):
            if ep.name == "native":
                return cast(Engine, ep.load())
        try:
            import optlab.native as in_tree
        except ModuleNotFoundError:
            return None
        return cast(Engine, in_tree)

    def why_not(self, model) -> str | None:
        """None when the compiled engine can run `model` now, else the reason."""
        eng = self.engine()
        if eng is None:
            return "the compiled engine optlab.native is not importable"
        reason = eng.unsupported(model)
        if reason is not None:
            return reason
        if not eng.available():
            return "the compiled engine is not built"
        if not eng.gpu_available():
            return "no HIP device"
        return None


_REGISTRY = _Registry()


def compiled() -> Engine | None:
    """The engine registered as "native" in `optlab.engines`, else the in-tree `optlab.native`, else None."""
    return _REGISTRY.engine()


def set_default(engine: str | None) -> None:
    """The engine used where a call names none; None or "" restores $OPTLAB_ENGINE (else "numpy")."""
    _REGISTRY.preferred = _engine_name(engine) if engine else None


def default_engine() -> str:
    if _REGISTRY.preferred is not None:
        return _REGISTRY.preferred
    env = os.environ.get("OPTLAB_ENGINE", "").strip().lower()
    return env if env in ENGINES else "numpy"


def choose(engine: str | None, model) -> tuple[str, str | None]:
    """(backend, reason) for simulating `model` on the engine named `engine` (None: the default). "numpy" is
    returned as is; "auto" resolves to "gpu", or to "numpy" with the reason when the compiled engine cannot run the
    model; "cpu" and "gpu" that cannot run raise EngineError with the reason."""
    name = _engine_name(engine or default_engine())
    if name == "numpy":
        return "numpy", None
    reason = _REGISTRY.why_not(model)
    if name == "auto":
        return ("numpy", reason) if reason is not None else ("gpu", None)
    if reason is not None:
        raise EngineError(reason)
    return name, None


def require() -> Engine:
    """The compiled engine; EngineError when there is none."""
    eng = compiled()
    if eng is None:
        raise EngineError("the compiled engine optlab.native is not importable")
    return eng


def seed64(seed) -> int:
    """The engine's 64-bit Philox key: an int modulo 2^64, a SeedSequence's first 64-bit state word, or that of
    fresh entropy for None (the reference's key derivation, engines.py:131-137)."""
    if seed is None:
        seed = np.random.SeedSequence()
    if isinstance(seed, np.random.SeedSequence):
        return int(seed.generate_state(1, dtype=np.uint64)[0])
    return int(seed) % (1 << 64)
