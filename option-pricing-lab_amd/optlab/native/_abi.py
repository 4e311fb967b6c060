"""ctypes mirror of include/optlab_mi.h: the request structs (the layouts of the reference's
packages/native/cuda/include/optlab_cuda/requests.hpp:12-93, field for field) and the library loader.

The library is liboptlab_mi.so, built in-tree next to this file by `build.py`. Loading fails loudly: there is no
fallback engine behind this module."""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_D = C.c_double
_I = C.c_int
_LL = C.c_longlong
_ULL = C.c_ulonglong
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)

OK, EINVAL, EDEVICE, ENOMEM, EINTERNAL = 0, 1, 2, 3, 4


class Request(C.Structure):
    """optlab_request (requests.hpp:12-23)."""

    _fields_ = [
        ("model", _I),
        ("product", _I),
        ("legs", _I * 3),
        ("strikes", _D * 3),
        ("H", _D),
        ("n_steps", _I),
        ("S0", _D),
        ("r", _D),
        ("q", _D),
        ("T", _D),
        ("sigma", _D),
        ("v0", _D),
        ("kappa", _D),
        ("theta", _D),
        ("xi", _D),
        ("rho", _D),
        ("n_paths", _LL),
        ("seed", _ULL),
    ]


class Result(C.Structure):
    """optlab_result (requests.hpp:25-29)."""

    _fields_ = [("price", _D * 4), ("se", _D * 4), ("n_paths", _LL), ("kernel_ms", _D), ("total_ms", _D)]


class PathsRequest(C.Structure):
    """optlab_paths_request (requests.hpp:36-63)."""

    _fields_ = [
        ("n_names", _I),
        ("n_steps", _I),
        ("n_obs", _I),
        ("n_x", _I),
        ("n_rows", _I),
        ("n_pairs", _LL),
        ("first_pair", _LL),
        ("seed", _ULL),
        ("dt", _PD),
        ("t0", _PD),
        ("obs_step", _PI),
        ("root", _PD),
        ("kind", _PI),
        ("cvar", _PD),
        ("heston", _PD),
        ("same", _I),
        ("step_row", _PI),
        ("row_lo", _PD),
        ("row_inv_dx", _PD),
        ("table", _PD),
        ("fwd", _PD),
        ("mixer", _I),
        ("mix", _D * 8),
        ("low_root", _PD),
        ("n_cols", _I),
        ("n_keep", _I),
        ("col_of", _PI),
        ("n_windows", _I),
        ("win_lo", _PI),
        ("win_hi", _PI),
        ("out", _PD),
        ("kernel_ms", _D),
        ("total_ms", _D),
    ]


class Device(C.Structure):
    """optlab_device (requests.hpp:66-71)."""

    _fields_ = [
        ("name", C.c_char * 256),
        ("major", _I),
        ("minor", _I),
        ("sm_count", _I),
        ("clock_mhz", _I),
        ("total_mem", _LL),
        ("free_mem", _LL),
        ("integrated", _I),
        ("kernel_timeout", _I),
    ]


class LeverageRequest(C.Structure):
    """optlab_leverage_request (requests.hpp:76-93)."""

    _fields_ = [
        ("n_steps", _I),
        ("dt", _PD),
        ("t0", _PD),
        ("n_pairs", _LL),
        ("seed", _ULL),
        ("v0", _D),
        ("kappa", _D),
        ("theta", _D),
        ("xi", _D),
        ("rho", _D),
        ("n_x", _I),
        ("n_bins", _I),
        ("n_fine", _I),
        ("atm_vol", _D),
        ("bandwidth_scale", _D),
        ("l_min", _D),
        ("l_max", _D),
        ("q_lo", _D),
        ("q_hi", _D),
        ("local_linear", _I),
        ("positive_part", _I),
        ("n_lv", _I),
        ("lv_lo", _PD),
        ("lv_inv_dx", _PD),
        ("lv_rows", _PD),
        ("x_grids", _PD),
        ("values", _PD),
        ("total_ms", _D),
    ]


class LeverageOptions(C.Structure):
    """optlab_leverage_options (addition: the bandwidth rule of the particle method)."""

    _fields_ = [("bandwidth_rule", _I), ("bandwidth_coef", _D)]


class AutocallModels(C.Structure):
    """optlab_autocall_models (addition: each name's model in the worst-of autocall)."""

    _fields_ = [("kind", _PI), ("cvar", _PD), ("heston", _PD), ("lev2", _PD), ("same", _I)]


class ScriptResult(C.Structure):
    """optlab_script_result (addition: a payoff script fused onto the device)."""

    _fields_ = [("price", _D), ("se", _D), ("paths_ms", _D), ("payoff_ms", _D), ("compile_ms", _D), ("total_ms", _D)]


class LvSlices(C.Structure):
    """optlab_lv_slices (addition: the particle method's local variance as surface slices)."""

    _fields_ = [("n_rows", _I), ("hi", _PD), ("records", _PD), ("floor", _D), ("cap", _D)]


class StripRequest(C.Structure):
    """optlab_strip_request (addition: a strip of vanillas on one simulation)."""

    _fields_ = [
        ("model", _I),
        ("put", _I),
        ("n_steps", _I),
        ("S0", _D),
        ("r", _D),
        ("q", _D),
        ("T", _D),
        ("sigma", _D),
        ("v0", _D),
        ("kappa", _D),
        ("theta", _D),
        ("xi", _D),
        ("rho", _D),
        ("psi_c", _D),
        ("n_strikes", _I),
        ("strikes", _PD),
        ("n_mats", _I),
        ("mat_step", _PI),
        ("n_paths", _LL),
        ("seed", _ULL),
        ("price", _PD),
        ("se", _PD),
        ("kernel_ms", _D),
        ("total_ms", _D),
    ]


class SurfaceRows(C.Structure):
    """optlab_surface_rows (addition: a surface's rows on the engine's table grid, risk/surfaces.py:180-205)."""

    _fields_ = [
        ("n_rows", _I),
        ("n_x", _I),
        ("T", _PD),
        ("lo", _PD),
        ("inv_dx", _PD),
        ("w", _PD),
        ("w_k", _PD),
        ("w_kk", _PD),
        ("w_T", _PD),
    ]


GRID_STRIKES, GRID_TENORS = 10, 8
GRID_NODES = GRID_STRIKES * GRID_TENORS


class AutocallRequest(C.Structure):
    """optlab_autocall_request (addition: the worst-of autocall and its adjoint Greeks)."""

    _fields_ = [
        ("n_names", _I),
        ("n_steps", _I),
        ("dt", _D),
        ("surface", _PI),
        ("root", _PD),
        ("fwd", _PD),
        ("fixing", _PD),
        ("n_obs", _I),
        ("obs_step", _PI),
        ("disc", _PD),
        ("autocall", _D),
        ("coupon_barrier", _D),
        ("coupon", _D),
        ("memory", _I),
        ("ki_barrier", _D),
        ("put_strike", _D),
        ("smoothing", _D),
        ("n_pairs", _LL),
        ("first_pair", _LL),
        ("seed", _ULL),
        ("nodes", _PD),
        ("spot", _PD),
        ("price", _D),
        ("se", _D),
        ("batch", _D * 16),
        ("delta", _D * 5),
        ("delta_se", _D * 5),
        ("vega", _PD),
        ("vega_se", _PD),
        ("forward_ms", _D),
        ("backward_ms", _D),
        ("total_ms", _D),
        ("checkpoint_every", _I),
        ("checkpoint_used", _I),
    ]


class Trade(C.Structure):
    """optlab_trade (addition: one trade of a book)."""

    _fields_ = [
        ("product", _I),
        ("n_names", _I),
        ("names", _I * 5),
        ("root", _D * 25),
        ("n_steps", _I),
        ("obs_every", _I),
        ("strike", _D),
        ("barrier", _D),
        ("coupon", _D),
        ("coupon_barrier", _D),
        ("autocall", _D),
        ("memory", _I),
        ("quantity", _D),
        ("n_pairs", _LL),
        ("seed", _ULL),
    ]


class BookRequest(C.Structure):
    """optlab_book_request (addition: a book of trades revalued under scenarios in one persistent launch)."""

    _fields_ = [
        ("n_names", _I),
        ("n_steps", _I),
        ("dt", _D),
        ("surface", _PI),
        ("fwd", _PD),
        ("disc", _PD),
        ("n_trades", _I),
        ("trades", C.POINTER(Trade)),
        ("n_scen", _I),
        ("spot_shock", _PD),
        ("vol_shift", _PD),
        ("skew", _PD),
        ("mode", _I),
        ("block", _I),
        ("pairs_per_lane", _I),
        ("value", _PD),
        ("se", _PD),
        ("finish_ms", _PD),
        ("tables_ms", _D),
        ("kernel_ms", _D),
        ("total_ms", _D),
        ("name_steps", _LL),
        ("block_used", _I),
        ("pairs_per_lane_used", _I),
    ]


class LsmRequest(C.Structure):
    """optlab_lsm_request (addition: an American option by Longstaff-Schwartz on the engine's paths)."""

    _fields_ = [
        ("paths", PathsRequest),
        ("strike", _D),
        ("call", _I),
        ("disc", _PD),
        ("spot", _D),
        ("basis_var", _I),
        ("min_itm", _I),
        ("price", _D),
        ("se", _D),
        ("coefs", _PD),
        ("paths_ms", _D),
        ("regression_ms", _D),
        ("apply_ms", _D),
        ("total_ms", _D),
    ]


class RegressRequest(C.Structure):
    """optlab_regress_request (addition: the issuer's call regression, exercise.py:191-238)."""

    _fields_ = [
        ("n_rows", _LL),
        ("n_vars", _I),
        ("state", _PD),
        ("y", _PD),
        ("mean", _PD),
        ("scale", _PD),
        ("n_powers", _I),
        ("powers", _PI),
        ("n_knots", _I),
        ("knot_var", _PI),
        ("knot", _PD),
        ("coef", _PD),
        ("ok", _I),
        ("n_eval", _LL),
        ("eval_state", _PD),
        ("eval_out", _PD),
        ("fit_ms", _D),
        ("eval_ms", _D),
        ("total_ms", _D),
    ]


def ptr(a, kind=_PD):
    """A ctypes pointer to a NumPy array's data (None for None or an empty array)."""
    if a is None or a.size == 0:
        return C.cast(None, kind)
    return a.ctypes.data_as(kind)


LIB_NAME = "liboptlab_mi.so"


def library_path() -> Path:
    return Path(os.environ.get("OPTLAB_MI_LIB", Path(__file__).with_name(LIB_NAME)))


_SIGNATURES = {
    "optlab_gpu_count": (_I, []),
    "optlab_has_hip": (_I, []),
    "optlab_cpu_threads": (_I, []),
    "optlab_gpu_name": (_I, [C.c_char_p, _I]),
    "optlab_device_info": (_I, [_I, C.POINTER(Device)]),
    "optlab_last_error": (C.c_char_p, []),
    "optlab_abi_version": (_I, []),
    "optlab_price": (_I, [C.POINTER(Request), _I, _I, C.POINTER(Result)]),
    "optlab_paths": (_I, [C.POINTER(PathsRequest), _I, _I]),
    "optlab_paths_dev": (_I, [C.POINTER(PathsRequest), _I, C.c_void_p, C.c_void_p]),
    "optlab_leverage": (_I, [C.POINTER(LeverageRequest), _I, _I]),
    "optlab_leverage_ex": (_I, [C.POINTER(LeverageRequest), C.POINTER(LeverageOptions), _I]),
    "optlab_leverage_slices": (_I, [C.POINTER(LeverageRequest), C.POINTER(LeverageOptions), C.POINTER(LvSlices), _I]),
    "optlab_strip": (_I, [C.POINTER(StripRequest), _I]),
    "optlab_strip_greeks": (_I, [C.POINTER(StripRequest), _I, _PD, _PD, _PD, _PD]),
    "optlab_strip_noalloc": (_I, [C.POINTER(StripRequest), _I, C.c_void_p, C.c_void_p, C.c_void_p]),
    "optlab_surface_upload": (_I, [C.POINTER(SurfaceRows), _PI]),
    "optlab_surface_free": (_I, [_I]),
    "optlab_surface_table": (_I, [_I, _PD, _D, _D, _PD]),
    "optlab_surface_node_grad": (_I, [_I, _PD, _PD]),
    "optlab_autocall": (_I, [C.POINTER(AutocallRequest), _I]),
    "optlab_adjoint": (_I, [C.POINTER(AutocallRequest), _I]),
    "optlab_script": (_I, [C.POINTER(PathsRequest), _I, C.c_char_p, C.POINTER(ScriptResult)]),
    "optlab_autocall_ex": (_I, [C.POINTER(AutocallRequest), C.POINTER(AutocallModels), _I, _I]),
    "optlab_book": (_I, [C.POINTER(BookRequest), _I]),
    "optlab_lsm": (_I, [C.POINTER(LsmRequest), _I]),
    "optlab_regress": (_I, [C.POINTER(RegressRequest), _I]),
    "optlab_philox_dev": (_I, [C.c_uint] * 6 + [_LL, C.c_void_p, C.c_void_p]),
    "optlab_normals_dev": (_I, [_LL, _LL, _I, C.c_uint, _ULL, _I, C.c_void_p, C.c_void_p]),
}


def declare(lib: C.CDLL, signatures: dict) -> C.CDLL:
    for name, (res, args) in signatures.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB: list = []


def _one_hip_runtime() -> None:
    """PyTorch-ROCm ships its own copy of the HIP runtime (libamdhip64.so.7, same soname as /opt/rocm's). Two
    copies in one process cannot both open the device, so when PyTorch is installed it is imported first and the
    engine binds to the runtime it loaded: the engine's device buffers and PyTorch's tensors then share one HIP
    context (the adjoint's autograd.Function and the device-resident outputs rely on it)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> C.CDLL:
    """The HIP engine. Raises OSError when the library is not built (run build.py): no engine runs without it."""
    if not _LIB:
        _one_hip_runtime()
        path = library_path()
        if not path.exists():
            raise OSError(f"{path} is not built: run `python __graft_entry__.py` (or build.py) to compile the HIP engine")
        _LIB.append(declare(C.CDLL(str(path)), _SIGNATURES))
    return _LIB[0]
