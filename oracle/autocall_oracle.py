"""CPU ORACLE -- test infrastructure only. The worst-of autocall (Phoenix) evaluated in double on paths of the oracle's
path engine (liboracle.so, oracle_paths), which the HIP autocall kernels are checked against.

    the script of a phoenix, worst of n names ........ packages/structured/src/optlab/structured/compiler.py:114-161,
                                                       270-295 (coupon with memory, call, knock-in put at maturity)
    smoothed comparisons ............................. packages/structured/src/optlab/structured/script/evaluator.py:93-108
    an `if` on a mask pays mask * then + (1 - mask) * else; `terminate` multiplies the live weight by 1 - mask
    local-vol paths of n names, one table row a step . packages/native/cuda/include/optlab_cuda/paths.hpp:274-361
    other kinds a name (constant, Heston, SLV) ......... paths.hpp:196-361 (the oracle path engine's own kinds)
"""

from __future__ import annotations

import ctypes as C

import numpy as np


def ramp_ge(a, b, eps):
    """a >= b, as the evaluator smooths it: clip((a - b) / (eps |b|) + 1/2, 0, 1); hard when eps |b| is 0."""
    scale = eps * abs(b)
    if scale <= 0:
        return (np.asarray(a) >= b).astype(float)
    return np.clip((np.asarray(a) - b) / scale + 0.5, 0.0, 1.0)


def ramp_lt(a, b, eps):
    scale = eps * abs(b)
    if scale <= 0:
        return (np.asarray(a) < b).astype(float)
    return np.clip((b - np.asarray(a)) / scale + 0.5, 0.0, 1.0)


def phoenix(closes, fixing, obs_step, disc, coupon, coupon_barrier, autocall, memory, ki_barrier, put_strike, eps):
    """Value of every path: closes (n_names, n_steps + 1, rows) of the spots at every step (t = 0 included)."""
    perf_all = np.min(closes / np.asarray(fixing, dtype=float)[:, None, None], axis=0)
    low = perf_all.min(axis=0)
    rows = closes.shape[2]
    alive, missed, value = np.ones(rows), np.zeros(rows), np.zeros(rows)
    last = len(obs_step) - 1
    for j, s in enumerate(obs_step):
        p, D = perf_all[s], disc[j]
        c = ramp_ge(p, coupon_barrier, eps)
        value += alive * c * coupon * ((1 + missed) if memory else 1.0) * D
        if memory:
            missed = (1 - c) * (missed + 1)
        if j < last:
            u = ramp_ge(p, autocall, eps)
            value += alive * u * D
            alive = alive * (1 - u)
        else:
            knocked = ramp_lt(low, ki_barrier, eps) if ki_barrier > 0 else np.zeros(rows)
            value += alive * D * (knocked * (1 - np.maximum(put_strike - p, 0.0)) + (1 - knocked))
    return value


def closes(lib, paths_request_type, n_pairs, seed, root, tables, lo, inv, fwd, dt, first_pair=0, threads=0,
           kinds=None, cvar=None, heston=None, same=0):
    """The spots at every step, (n_names, n_steps + 1, 2 n_pairs), from the oracle's path engine: table row s of name
    a read at step s, forwards fwd (n_names, n_steps + 1); every name local vol, or name a of kind kinds[a] (0
    constant variance cvar[a], 1 local vol, 2 Heston, 3 SLV with tables[a] its squared leverage) with Heston
    parameters heston[a] = (v0, kappa, theta, xi, rho) and the variance normals correlated by root when `same`."""
    n, n_steps, n_x = tables.shape
    rq = paths_request_type()
    keep = {}

    def put(name, a, ctype=C.c_double):
        arr = np.ascontiguousarray(a, dtype=np.float64 if ctype is C.c_double else np.int32)
        keep[name] = arr
        return arr.ctypes.data_as(C.POINTER(ctype))

    rq.n_names, rq.n_steps, rq.n_obs, rq.n_x, rq.n_rows = n, n_steps, n_steps + 1, n_x, n * n_steps
    rq.n_pairs, rq.first_pair, rq.seed = n_pairs, first_pair, seed
    rq.dt = put("dt", np.full(n_steps, dt))
    rq.t0 = put("t0", np.arange(n_steps) * dt)
    rq.obs_step = put("obs", np.arange(n_steps + 1), C.c_int)
    rq.root = put("root", root)
    rq.kind = put("kind", np.ones(n) if kinds is None else kinds, C.c_int)
    rq.cvar = put("cvar", np.zeros(n) if cvar is None else cvar)
    rq.heston = put("heston", np.zeros(5 * n) if heston is None else np.asarray(heston).reshape(-1))
    rq.same = int(same)
    rq.step_row = put("step_row", np.arange(n * n_steps), C.c_int)
    rq.row_lo = put("lo", np.asarray(lo).reshape(-1))
    rq.row_inv_dx = put("inv", np.asarray(inv).reshape(-1))
    rq.table = put("table", np.asarray(tables).reshape(-1))
    rq.fwd = put("fwd", fwd)
    rq.mixer = 0
    rq.n_cols = rq.n_keep = n_steps + 1
    rq.col_of = put("col", np.arange(n_steps + 1), C.c_int)
    rq.n_windows = 0
    out = np.empty((n, n_steps + 1, 2 * n_pairs))
    rq.out = out.ctypes.data_as(C.POINTER(C.c_double))
    status = lib.oracle_paths(C.byref(rq), threads)
    if status:
        raise RuntimeError(lib.oracle_last_error().decode())
    return out


def pair_mean(values):
    """Mean and standard error of pair averages (rows p and n/2 + p are a pair)."""
    h = len(values) // 2
    pairs = 0.5 * (values[:h] + values[h:])
    return float(pairs.mean()), float(pairs.std(ddof=1) / np.sqrt(h))
