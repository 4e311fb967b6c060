"""Shared helpers of the tests: the CPU oracle (liboracle.so, test infrastructure only) and the compiled reference
(oracle/_ref/libref.so, present only where the reference sources were built), the simulated pricing context of the
reference's golden files, and the statistical checks."""

from __future__ import annotations

import ctypes as C
import json
import subprocess
from functools import lru_cache
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "_build" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libref.so"
FIXTURE = ROOT / "tests" / "golden" / "reference_native.json"


@lru_cache(maxsize=1)
def oracle() -> C.CDLL:
    """The oracle library, built from oracle/ if it is not there yet."""
    if not ORACLE.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    from optlab.native import _abi

    lib = C.CDLL(str(ORACLE))
    sig = {
        "oracle_last_error": (C.c_char_p, []),
        "oracle_philox": (None, [C.c_uint] * 6 + [C.c_longlong, C.c_void_p]),
        "oracle_normals": (None, [C.c_longlong, C.c_longlong, C.c_int, C.c_uint, C.c_ulonglong, C.c_void_p]),
        "oracle_paths": (C.c_int, [C.POINTER(_abi.PathsRequest), C.c_int]),
        "oracle_paths_var": (C.c_int, [C.POINTER(_abi.PathsRequest), C.c_int, C.c_void_p]),
        "oracle_price": (C.c_int, [C.POINTER(_abi.Request), C.c_int, C.POINTER(_abi.Result)]),
        "oracle_strip": (C.c_int, [C.POINTER(_abi.StripRequest), C.c_int]),
        "oracle_leverage": (C.c_int, [C.POINTER(_abi.LeverageRequest), C.c_int, C.c_int]),
        "oracle_leverage_ex": (C.c_int, [C.POINTER(_abi.LeverageRequest), C.POINTER(_abi.LeverageOptions), C.c_int,
                                         C.c_int]),
        "oracle_leverage_trace": (C.c_int, [C.POINTER(_abi.LeverageRequest), C.POINTER(_abi.LeverageOptions), C.c_int,
                                            C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
        "oracle_geometric_asian": (C.c_double, [C.c_double] * 6 + [C.c_int, C.c_int]),
    }
    return _abi.declare(lib, sig)


@lru_cache(maxsize=1)
def reference():
    """The reference's own CPU engine compiled from its sources (oracle/build_ref.sh), or None where it is absent."""
    if not REF.exists():
        return None
    from optlab.native import _abi

    lib = C.CDLL(str(REF))
    sig = {
        "ref_last_error": (C.c_char_p, []),
        "ref_paths": (C.c_int, [C.POINTER(_abi.PathsRequest), C.c_int]),
        "ref_leverage": (C.c_int, [C.POINTER(_abi.LeverageRequest), C.c_int]),
        "ref_philox": (None, [C.c_uint] * 6 + [C.c_longlong, C.c_void_p]),
        "ref_normals": (None, [C.c_longlong, C.c_longlong, C.c_int, C.c_uint, C.c_ulonglong, C.c_void_p]),
        "ref_fixed_sums": (C.c_int, [C.POINTER(_abi.Request), C.c_void_p, C.c_void_p]),
        "ref_geometric_asian": (C.c_double, [C.c_double] * 6 + [C.c_int, C.c_int]),
        "ref_bs_price": (C.c_double, [C.c_double] * 6 + [C.c_int]),
        "ref_down_out_call_bgk": (C.c_double, [C.c_double] * 7 + [C.c_int]),
    }
    return _abi.declare(lib, sig)


@lru_cache(maxsize=1)
def fixture() -> dict:
    return json.loads(FIXTURE.read_text())


def simulated_context():
    """(market, SVI surface, Heston params) of the reference's golden 'simulated' context."""
    from optlab.core.market import Market
    from optlab.core.models import HestonParams
    from optlab.core.svi import SVISurface

    ctx = fixture()["simulated_context"]
    m = ctx["market"]
    return Market.flat(m["spot"], m["r"], m["q"]), SVISurface.from_dict(ctx["surface"]), HestonParams(**ctx["heston"])


def digest_samples(a: np.ndarray) -> tuple[np.ndarray, float]:
    """The 32 evenly spaced samples and the sum the reference's golden digest keeps (golden.py:57-70)."""
    flat = np.ascontiguousarray(a, dtype=np.float64).ravel()
    idx = np.linspace(0, flat.size - 1, 32).astype(int)
    return flat[idx], float(flat.sum())


def oracle_paths(p, n_pairs: int, seed: int, first_pair: int = 0, keep=None, windows=(), lib=None, threads: int = 0):
    """The oracle's (or the compiled reference's) output array (names, n_cols, 2 n_pairs) for a plan."""
    from optlab import native
    from optlab.native import _abi

    rq, keepalive, _ = native.paths_request(p, n_pairs, seed, first_pair, keep, windows)
    out = np.empty((len(p.names), rq.n_cols, 2 * n_pairs))
    rq.out = _abi.ptr(out)
    if lib is None:
        status = oracle().oracle_paths(C.byref(rq), threads)
    else:
        status = lib.ref_paths(C.byref(rq), threads)
    assert status == 0
    del keepalive
    return out


def oracle_paths_var(p, n_pairs: int, seed: int, first_pair: int = 0):
    """(S, V): the oracle's spots of the plan's one name and its variance factor at every observation, each
    (2 n_pairs, n_obs)."""
    from optlab import native
    from optlab.native import _abi

    rq, keepalive, m = native.paths_request(p, n_pairs, seed, first_pair)
    out = np.empty((1, rq.n_cols, 2 * n_pairs))
    var = np.empty((rq.n_keep, 2 * n_pairs))
    rq.out = _abi.ptr(out)
    assert oracle().oracle_paths_var(C.byref(rq), 0, var.ctypes.data) == 0
    del keepalive
    return out[0, :m].T.copy(), var.T.copy()


def calls(S, D, strikes):
    """Prices and standard errors of calls on pair averages (rows i and n/2 + i are a pair)."""
    h = len(S) // 2
    pay = D * np.maximum(S[:, None] - np.asarray(strikes, dtype=float), 0.0)
    pay = 0.5 * (pay[:h] + pay[h:])
    return pay.mean(axis=0), pay.std(axis=0, ddof=1) / np.sqrt(h)
