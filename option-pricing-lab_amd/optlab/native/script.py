"""Payoff scripts fused onto the device: a program of the payoff language (optlab.structured.script) lowered to HIP
source, which the engine compiles by hipRTC and runs over the path engine's output in HBM (optlab_script).

The evaluation follows packages/structured/src/optlab/structured/script/evaluator.py:166-330, one leg a call of
`leg_value`: every variable a double per leg (0 before its first assignment); a condition a number in [0, 1] -- hard
(0 or 1, NaN when an operand is NaN) or, with smoothing eps > 0, the ramp clip((a - b) / (eps |b|) + 1/2, 0, 1) of
`a >= b` (b - a for `<`, `<=`; `==` stays hard); inside an `if` of mask m an assignment is m e + (1 - m) x (e where
m = 1, x where m = 0), `pay e` adds D(t_pay) w e with w = m alive (nothing where w = 0), `terminate` multiplies alive
by 1 - m, a NaN condition counts as 0; `and` is a b, `or` a + b - a b, `not` 1 - a; min, max, worst and best
propagate NaN as NumPy's minimum and maximum do. The leg's value is the discounted sum of its payments (the
reference sums flows by payment date and discounts them afterwards; the same sum). The evaluator's checks that
raise on undefined values are not made on the device: a NaN payment makes a NaN price.

Observation times are every event's time and every window's times (n equally spaced in (start, end]); the path
engine keeps the events' columns and follows each distinct window's lowest and highest spot (at most two windows,
every name)."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from optlab.core.engines import EngineError, seed64
from optlab.native import _abi
from optlab.structured import script as sc

BLOCK = 256

_PREAMBLE = r"""// Generated from a payoff script (optlab.native.script): the script's events in order, one leg a call.
namespace {
__device__ __forceinline__ bool s_nan(double x) { return x != x; }
__device__ __forceinline__ double s_hard(bool r, double a, double b) {
    return (s_nan(a) || s_nan(b)) ? __builtin_nan("") : (r ? 1.0 : 0.0);
}
__device__ __forceinline__ double s_ramp(double d, double b, double eps, double hard) {
    if (eps == 0.0) return hard;
    const double scale = eps * fabs(b);
    double x = d / scale + 0.5;
    x = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);  // NaN stays NaN
    return scale > 0.0 ? x : hard;
}
__device__ __forceinline__ double s_ge(double a, double b, double eps) { return s_ramp(a - b, b, eps, s_hard(a >= b, a, b)); }
__device__ __forceinline__ double s_gt(double a, double b, double eps) { return s_ramp(a - b, b, eps, s_hard(a > b, a, b)); }
__device__ __forceinline__ double s_le(double a, double b, double eps) { return s_ramp(b - a, b, eps, s_hard(a <= b, a, b)); }
__device__ __forceinline__ double s_lt(double a, double b, double eps) { return s_ramp(b - a, b, eps, s_hard(a < b, a, b)); }
__device__ __forceinline__ double s_eq(double a, double b, double) { return s_hard(a == b, a, b); }
__device__ __forceinline__ double s_and(double a, double b) {
    return (s_nan(a) || s_nan(b)) ? ((a == 0.0 || b == 0.0) ? 0.0 : a * b) : a * b;
}
__device__ __forceinline__ double s_or(double a, double b) {
    return (s_nan(a) || s_nan(b)) ? ((a == 1.0 || b == 1.0) ? 1.0 : a + b - a * b) : a + b - a * b;
}
__device__ __forceinline__ double s_min(double a, double b) { return (s_nan(a) || s_nan(b)) ? a + b : (b < a ? b : a); }
__device__ __forceinline__ double s_max(double a, double b) { return (s_nan(a) || s_nan(b)) ? a + b : (b > a ? b : a); }
__device__ __forceinline__ double s_assign(double m, double e, double old) {
    return m == 1.0 ? e : (m == 0.0 ? old : m * e + (1.0 - m) * old);
}
__device__ __forceinline__ double s_cond(double c) { return s_nan(c) ? 0.0 : c; }
}  // namespace
"""

_KERNEL = r"""
extern "C" __global__ void __launch_bounds__(256) script_payoff(const double* __restrict__ out, long long n_pairs,
                                                                double* __restrict__ rows) {
    __shared__ double part[2][4];
    const long long p = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    double d = 0.0;
    if (p < n_pairs) {
        const long long legs = 2 * n_pairs;
        d = 0.5 * (leg_value(out, p, legs) + leg_value(out, n_pairs + p, legs));
    }
    double s = d, q = d * d;
    for (int o = 32; o >= 1; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        q += __shfl_xor(q, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = s;
        part[1][threadIdx.x >> 6] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rows[2 * blockIdx.x] = ((part[0][0] + part[0][1]) + part[0][2]) + part[0][3];
        rows[2 * blockIdx.x + 1] = ((part[1][0] + part[1][1]) + part[1][2]) + part[1][3];
    }
}
"""


def _lit(x: float) -> str:
    x = float(x)
    if math.isnan(x):
        return '__builtin_nan("")'
    if math.isinf(x):
        return "__builtin_huge_val()" if x > 0 else "(-__builtin_huge_val())"
    return f"({float.hex(x)})"  # exact


class _Lowering:
    def __init__(self, prog, params, eps, spot_col, window_col, name_index, n_cols, disc):
        self.prog, self.params, self.eps = prog, params, float(eps)
        self.spot_col, self.window_col, self.name_index, self.n_cols = spot_col, window_col, name_index, n_cols
        self.disc = disc
        self.assigned: set = set()
        self.vars: list = []
        self.lines: list = []
        self.counter = 0
        self.k = 0

    def fresh(self, stem):
        self.counter += 1
        return f"{stem}{self.counter}"

    def read(self, a: int, col: int) -> str:
        return f"out[{a * self.n_cols + col}ll * legs + row]"

    def expr(self, e) -> str:
        if isinstance(e, sc.Num):
            return _lit(e.value)
        if isinstance(e, sc.Name):
            if e.id in self.assigned:
                return f"v_{e.id}"
            if e.id in self.params:
                return _lit(self.params[e.id])
            return self.read(self.name_index[e.id], self.spot_col[self.k])
        if isinstance(e, sc.Unary):
            v = self.expr(e.operand)
            return f"(-{v})" if e.op == "-" else f"(1.0 - {v})"
        if isinstance(e, sc.Binary):
            a, b = self.expr(e.left), self.expr(e.right)
            if e.op in "+-*/":
                return f"({a} {e.op} {b})"
            if e.op == "and":
                return f"s_and({a}, {b})"
            if e.op == "or":
                return f"s_or({a}, {b})"
            fn = {">=": "s_ge", ">": "s_gt", "<=": "s_le", "<": "s_lt", "==": "s_eq"}[e.op]
            return f"{fn}({a}, {b}, {_lit(self.eps)})"
        if isinstance(e, sc.Call):
            args = [self.expr(x) for x in e.args]
            if e.func in ("worst", "best", "min", "max"):
                fn = "s_min" if e.func in ("worst", "min") else "s_max"
                acc = args[0]
                for x in args[1:]:
                    acc = f"{fn}({acc}, {x})"
                return acc
            return f"{e.func}({args[0]})"
        if isinstance(e, sc.Window):
            lo, hi = self.window_col[(e.start, e.end, e.n)]
            return self.read(self.name_index[e.underlying], lo if e.func == "lowest" else hi)
        raise TypeError(f"cannot lower {e!r}")

    def block(self, body, m: str, indent: str):
        for s in body:
            if isinstance(s, sc.Assign):
                e = self.expr(s.expr)  # read before the name is assigned (x = x + 1 reads the old x)
                if s.name not in self.assigned:
                    self.assigned.add(s.name)
                    if s.name not in self.vars:
                        self.vars.append(s.name)
                self.lines.append(f"{indent}v_{s.name} = s_assign({m}, {e}, v_{s.name});")
            elif isinstance(s, sc.Pay):
                w = self.fresh("w")
                self.lines.append(f"{indent}{{ const double {w} = {m} * alive;")
                self.lines.append(f"{indent}  if ({w} != 0.0) value += {_lit(self.disc[self.k])} * ({w} * "
                                  f"{self.expr(s.expr)}); }}")
            elif isinstance(s, sc.Terminate):
                self.lines.append(f"{indent}alive = alive * (1.0 - {m});")
            elif isinstance(s, sc.If):
                c = self.fresh("c")
                self.lines.append(f"{indent}const double {c} = s_cond({self.expr(s.cond)});")
                mb = self.fresh("m")
                self.lines.append(f"{indent}const double {mb} = {m} * {c};")
                self.block(s.body, mb, indent)
                if s.orelse:
                    me = self.fresh("m")
                    self.lines.append(f"{indent}const double {me} = {m} * (1.0 - {c});")
                    self.block(s.orelse, me, indent)

    def source(self) -> str:
        for k, ev in enumerate(self.prog.events):
            self.k = k
            self.lines.append(f"    {{  // event t={ev.t_obs!r} pay t={ev.t_pay!r}")
            self.block(ev.body, "1.0", "        ")
            self.lines.append("    }")
        head = ["__device__ __forceinline__ double leg_value(const double* __restrict__ out, long long row, "
                "long long legs) {", "    double alive = 1.0, value = 0.0;"]
        head += [f"    double v_{v} = 0.0;" for v in self.vars]
        return _PREAMBLE + "\n".join(head + self.lines + ["    return value;", "}"]) + "\n" + _KERNEL


class ScriptPricer:
    """A script priced on the engine's paths of `model` (a model, or a Basket whose names are the script's
    underlyings) with `market` (the discount curve; for a Basket, `market` or else its first name's market), with
    `params` over the script's own and comparisons smoothed by `smoothing` (0: hard)."""

    def __init__(self, script, model, market=None, params=None, smoothing: float = 0.0, dt_max=1 / 252,
                 dt_min=1 / 2000):
        from optlab import native
        from optlab.structured.multiasset import Basket

        prog = sc.parse(script) if isinstance(script, str) else script
        unknown = set(params or {}) - set(prog.params)
        if unknown:
            raise KeyError(f"unknown parameter(s): {', '.join(sorted(unknown))}")
        self.program, self.params, self.smoothing = prog, {**prog.params, **(params or {})}, float(smoothing)
        if isinstance(model, Basket):
            names = list(model.names)
            curve = market if market is not None else model.markets[names[0]]
        else:
            if len(prog.underlyings) != 1:
                raise ValueError("a script on several underlyings needs a Basket of them")
            names = list(prog.underlyings)
            curve = market
        missing = set(prog.underlyings) - set(names)
        if missing:
            raise ValueError(f"the model has no {', '.join(sorted(missing))}")
        key = lambda t: round(float(t), 12)  # noqa: E731
        spans = []
        for w in prog.windows():
            if (w.start, w.end, w.n) not in spans:
                spans.append((w.start, w.end, w.n))
        if len(spans) > native.MAX_WINDOWS:
            raise EngineError(f"{len(spans)} distinct windows: the engine follows at most {native.MAX_WINDOWS}")
        times = {key(ev.t_obs) for ev in prog.events}
        for start, end, n in spans:
            times |= {key(start + (end - start) * j / n) for j in range(1, n + 1)}
        self.times = np.array(sorted(times))
        index = {t: i for i, t in enumerate(self.times)}
        event_obs = sorted({index[key(ev.t_obs)] for ev in prog.events})
        self.keep = np.array(event_obs, dtype=int)
        spot_col = [event_obs.index(index[key(ev.t_obs)]) for ev in prog.events]
        self.windows = [(index[key(s + (e - s) / n)], index[key(e)]) for s, e, n in spans]
        window_col = {sp: (len(event_obs) + 2 * i, len(event_obs) + 2 * i + 1) for i, sp in enumerate(spans)}
        self.plan = native.plan(model, market, self.times, dt_max, dt_min)
        disc = [float(curve.D(ev.t_pay)) for ev in prog.events]
        n_cols = len(event_obs) + 2 * len(spans)
        self.source = _Lowering(prog, self.params, self.smoothing, spot_col, window_col,
                                {n: i for i, n in enumerate(names)}, n_cols, disc).source()
        self.names, self.spot_col, self.spans, self.n_cols = names, spot_col, spans, n_cols
        self._lib = _abi.load()

    def price(self, n_paths: int, seed=None, backend: str = "fp64", first_pair: int = 0) -> dict:
        """{"price", "se", "paths_ms", "payoff_ms", "compile_ms", "total_ms"}: the mean over pairs of the pair-averaged
        discounted payoff, the paths simulated and valued on the device (`backend` sets the paths' precision)."""
        from optlab import native

        n_pairs = int(n_paths) // 2
        rq, keep, _ = native.paths_request(self.plan, n_pairs, seed64(seed), first_pair, self.keep, self.windows)
        res = _abi.ScriptResult()
        code = self._lib.optlab_script(C.byref(rq), native.precision_code(backend), self.source.encode(), C.byref(res))
        del keep
        if code:
            raise EngineError(self._lib.optlab_last_error().decode())
        return {f: getattr(res, f) for f, _ in _abi.ScriptResult._fields_}

    def paths(self, n_paths: int, seed=None, backend: str = "fp64", first_pair: int = 0):
        """The same paths on the host, as the script reads them: ({name: (legs, events)}, {name: {(start, end, n):
        (lowest, highest)}}), for checking the device's payoff against an evaluation of one's own."""
        from optlab import native

        spots, extremes, _ = native.run(self.plan, int(n_paths) // 2, seed64(seed), first_pair, backend,
                                        keep=self.keep, windows=self.windows)
        by_event = {n: np.asarray(spots[a])[:, self.spot_col] for a, n in enumerate(self.names)}
        win = {n: {sp: extremes[a][i] for i, sp in enumerate(self.spans)} for a, n in enumerate(self.names)}
        return by_event, win
