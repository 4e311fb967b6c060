"""Implied-vol surfaces held on the device (packages/risk/src/optlab/risk/surfaces.py:95-205).

`base_rows` is the reference's `_base_rows`: the surface's total variance and its derivatives on the grid the
engine builds for its local-variance table (one row per step start, k in [-8, 6] ATM standard deviations, 801
points). `DeviceSurface` uploads them once; the engine then builds the local-variance table of the surface moved
by node tents, a parallel shift or a skew tilt on the device (TentBumpedSurface.local_var_rows), in FP64."""

from __future__ import annotations

import ctypes as C

import numpy as np

from optlab.native import _abi


def base_rows(surface, t0, n_x: int = 801, sd_range=(-8.0, 6.0)):
    """(lo, inv_dx, T, w, w_k, w_kk, w_T) with T = max(t0, 1/365) per row (surfaces.py:180-205)."""
    t0 = np.asarray(t0, dtype=float)
    if hasattr(surface, "row_grid"):  # every row in one vectorised pass (equal to the loop to the last bit)
        T, lo, hi, k = surface.row_grid(t0, n_x, sd_range)
        return lo, (n_x - 1) / (hi - lo), T, *surface.derivatives(k, T[:, None])
    n = len(t0)
    lo, inv = np.empty(n), np.empty(n)
    T = np.maximum(t0, 1.0 / 365)
    out = [np.empty((n, n_x)) for _ in range(4)]
    for i, t in enumerate(T):
        sd = float(np.sqrt(surface.w(0.0, float(t))))
        a, b = sd_range[0] * sd, sd_range[1] * sd
        lo[i], inv[i] = a, (n_x - 1) / (b - a)
        k = np.linspace(a, b, n_x)
        for o, v in zip(out, surface.derivatives(k, float(t))):
            o[i] = v
    return lo, inv, T, *out


class DeviceSurface:
    """A surface's rows uploaded to the current HIP device; `handle` names it in engine requests."""

    def __init__(self, surface, t0, n_x: int = 801, sd_range=(-8.0, 6.0)):
        self.surface = surface
        lo, inv, T, w, wk, wkk, wT = base_rows(surface, t0, n_x, sd_range)
        self.lo, self.inv_dx, self.T = lo, inv, T
        self.n_rows, self.n_x = len(T), n_x
        rows = _abi.SurfaceRows()
        rows.n_rows, rows.n_x = self.n_rows, n_x
        keep = [np.ascontiguousarray(a, dtype=float) for a in (T, lo, inv, w, wk, wkk, wT)]
        for name, a in zip(("T", "lo", "inv_dx", "w", "w_k", "w_kk", "w_T"), keep):
            setattr(rows, name, _abi.ptr(a))
        self._lib = _abi.load()
        handle = C.c_int(0)
        _check(self._lib, self._lib.optlab_surface_upload(C.byref(rows), C.byref(handle)))
        self.handle = handle.value

    def k(self) -> np.ndarray:
        """The table grid in log forward moneyness, (n_rows, n_x)."""
        return self.lo[:, None] + np.arange(self.n_x)[None, :] / self.inv_dx[:, None]

    def local_var(self, nodes=None, parallel: float = 0.0, skew: float = 0.0) -> np.ndarray:
        """The local-variance table of the surface moved by node bumps (n_strikes x n_tenors, vol units), a parallel
        shift and a skew tilt, built on the device in FP64 (n_rows x n_x)."""
        out = np.empty((self.n_rows, self.n_x))
        h = None if nodes is None else np.ascontiguousarray(nodes, dtype=float).reshape(-1)
        if h is not None and h.size != _abi.GRID_NODES:
            raise ValueError(f"nodes must hold {_abi.GRID_NODES} values")
        _check(self._lib, self._lib.optlab_surface_table(self.handle, _abi.ptr(h), parallel, skew, _abi.ptr(out)))
        return out

    def node_gradient(self, tbar) -> np.ndarray:
        """sum over entries of tbar * dLV / dH_node at the unbumped surface, (n_strikes, n_tenors)."""
        tb = np.ascontiguousarray(tbar, dtype=float)
        if tb.shape != (self.n_rows, self.n_x):
            raise ValueError("tbar must be n_rows x n_x")
        out = np.empty(_abi.GRID_NODES)
        _check(self._lib, self._lib.optlab_surface_node_grad(self.handle, _abi.ptr(tb), _abi.ptr(out)))
        return out.reshape(_abi.GRID_STRIKES, _abi.GRID_TENORS)

    def close(self) -> None:
        if getattr(self, "handle", 0):
            self._lib.optlab_surface_free(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check(lib, code: int) -> None:
    if code:
        from optlab.core.engines import EngineError

        raise EngineError(lib.optlab_last_error().decode())
