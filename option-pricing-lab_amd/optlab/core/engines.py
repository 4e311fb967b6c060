"""Which engine simulates, asked for by name (packages/core/src/optlab/core/engines.py:1-137).

"cpu" and "gpu" are the compiled engine; "auto" takes it when it can run the model; "numpy" is the library's own
NumPy reference, which lives outside this build. In this build the compiled engine is the HIP engine of
`optlab.native`: "gpu" runs its float mode and "cpu" its double mode -- the reference CPU engine's numbers -- both
on the GPU. The registry first looks for an engine registered under the entry point group `optlab.engines` (name
"native"), then for the in-tree `optlab.native`."""

from __future__ import annotations

import os
from importlib.metadata import entry_points
from typing import Any, Protocol, cast

import numpy as np

ENGINES = ("numpy", "cpu", "gpu", "auto")
GROUP = "optlab.engines"
NOT_INSTALLED = "no compiled engine is installed (optlab-native)"


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


_DEFAULT: str | None = None
_compiled: list = []


def compiled() -> Engine | None:
    """The engine registered as "native" in `optlab.engines`, else the in-tree `optlab.native`, else None. An engine
    that is found but fails to import raises: a broken install is never silently replaced."""
    if not _compiled:
        found = next((ep for ep in entry_points(group=GROUP) if ep.name == "native"), None)
        if found is not None:
            _compiled.append(cast(Engine, found.load()))
        else:
            try:
                import optlab.native as native
            except ModuleNotFoundError:
                native = None
            _compiled.append(cast(Engine, native) if native is not None else None)
    return _compiled[0]


def set_default(engine: str | None) -> None:
    global _DEFAULT
    if engine is not None and engine.lower() not in ENGINES:
        raise ValueError(f"engine must be one of {', '.join(ENGINES)}, not {engine!r}")
    _DEFAULT = engine.lower() if engine else None


def default_engine() -> str:
    if _DEFAULT:
        return _DEFAULT
    e = os.environ.get("OPTLAB_ENGINE", "numpy").lower()
    return e if e in ENGINES else "numpy"


def choose(engine: str | None, model) -> tuple[str, str | None]:
    """(backend, reason) for this request, as engines.py:96-120: "auto" falls back to "numpy" with the reason, a
    named engine that cannot run raises EngineError."""
    engine = (engine or default_engine()).lower()
    if engine not in ENGINES:
        raise ValueError(f"engine must be one of {', '.join(ENGINES)}, not {engine!r}")
    if engine == "numpy":
        return "numpy", None
    eng = compiled()
    if eng is None:
        why = NOT_INSTALLED
    else:
        why = eng.unsupported(model)
        if why is None and not eng.available():
            why = "the compiled engine is not built"
        if why is None and not eng.gpu_available():  # every mode of this engine runs on the GPU
            why = "no HIP device"
    if why is not None:
        if engine == "auto":
            return "numpy", why
        raise EngineError(why)
    assert eng is not None  # noqa: S101
    if engine == "auto":
        return "gpu", None
    return engine, None


def require() -> Engine:
    eng = compiled()
    if eng is None:
        raise EngineError(NOT_INSTALLED)
    return eng


def seed64(seed) -> int:
    """The engine's 64-bit Philox key from an int, a SeedSequence or None (engines.py:131-137)."""
    if isinstance(seed, np.random.SeedSequence):
        return int(seed.generate_state(1, np.uint64)[0])
    if seed is None:
        return int(np.random.SeedSequence().generate_state(1, np.uint64)[0])
    return int(seed) & (2**64 - 1)
