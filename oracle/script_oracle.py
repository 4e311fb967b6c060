"""CPU ORACLE -- test infrastructure only. The payoff language's evaluator restated in NumPy over whole arrays of
legs, which the fused device payoff (optlab_script) is checked against on the same paths.

    statements, masks, alive, flows ...... packages/structured/src/optlab/structured/script/evaluator.py:136-220
                                           (Evaluator.run, Evaluator.block)
    expressions, names ................... evaluator.py:277-318 (Evaluator.expr): a variable, else a parameter, else
                                           the underlying's spot at the event's observation
    comparisons and their smoothing ...... evaluator.py:93-108 (_compare)
    windows .............................. evaluator.py:320-334 (lowest / highest of the window's fixings)

Flows are summed by (payment time, tag) and discounted afterwards; the price is the mean over antithetic pairs of
the pair average (rows p and n/2 + p)."""

from __future__ import annotations

import functools

import numpy as np

HARD = {"<": np.less, "<=": np.less_equal, ">": np.greater, ">=": np.greater_equal, "==": np.equal}
FUNCS = {"min": np.minimum, "max": np.maximum, "abs": np.abs, "exp": np.exp, "log": np.log, "sqrt": np.sqrt}


def _nan_in(a, b):
    return bool(np.any(np.isnan(a)) or np.any(np.isnan(b)))


def _compare(op, a, b, eps):
    hard = HARD[op](a, b).astype(float)
    if _nan_in(a, b):
        hard = np.where(np.isnan(a) | np.isnan(b), np.nan, hard)
    if eps == 0 or op == "==":
        return hard
    scale = eps * np.abs(b)
    d = np.subtract(a, b) if op in (">", ">=") else np.subtract(b, a)
    with np.errstate(divide="ignore", invalid="ignore"):
        ramp = np.clip(d / scale + 0.5, 0.0, 1.0)
    return np.where(scale > 0, ramp, hard)


class Evaluation:
    """One run of a program over n legs: spot(name, k) -> the name's spots at event k, window(w) -> the window's
    lowest or highest spot of its underlying, disc(t) -> the discount factor to t."""

    def __init__(self, prog, params, eps, n, spot, window, disc):
        self.prog, self.params, self.eps, self.n = prog, params, float(eps), n
        self.spot, self.window, self.disc = spot, window, disc
        self.vars = {}
        self.alive = np.ones(n)
        self.flows = {}

    def run(self):
        from optlab.structured import script as sc

        self.sc = sc
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            for k, ev in enumerate(self.prog.events):
                self.k, self.t_pay = k, ev.t_pay
                self.block(ev.body, 1.0)
        total = np.zeros(self.n)
        for (t, _tag), f in self.flows.items():
            total = total + self.disc(t) * f
        return total

    def block(self, body, m):
        sc = self.sc
        for s in body:
            if isinstance(s, sc.Assign):
                e = np.broadcast_to(np.asarray(self.expr(s.expr), dtype=float), (self.n,))
                if isinstance(m, float) and m == 1.0:
                    self.vars[s.name] = e
                else:
                    old = self.vars.get(s.name, 0.0)
                    self.vars[s.name] = np.where(m == 1.0, e, np.where(m == 0.0, old, m * e + (1 - m) * old))
            elif isinstance(s, sc.Pay):
                w = m * self.alive
                e = self.expr(s.expr)
                paid = np.where(w != 0, w * e, 0.0)
                key = (self.t_pay, s.tag)
                self.flows[key] = self.flows.get(key, 0.0) + paid
            elif isinstance(s, sc.Terminate):
                self.alive = self.alive * (1.0 - m)
            elif isinstance(s, sc.If):
                c = np.broadcast_to(np.asarray(self.expr(s.cond), dtype=float), (self.n,))
                c = np.where(np.isnan(c), 0.0, c)
                self.block(s.body, m * c)
                if s.orelse:
                    self.block(s.orelse, m * (1.0 - c))

    def expr(self, e):
        sc = self.sc
        if isinstance(e, sc.Num):
            return e.value
        if isinstance(e, sc.Name):
            if e.id in self.vars:
                return self.vars[e.id]
            if e.id in self.params:
                return self.params[e.id]
            return self.spot(e.id, self.k)
        if isinstance(e, sc.Binary):
            a, b = self.expr(e.left), self.expr(e.right)
            if e.op == "+":
                return np.add(a, b)
            if e.op == "-":
                return np.subtract(a, b)
            if e.op == "*":
                return np.multiply(a, b)
            if e.op == "/":
                return np.divide(a, b)
            if e.op == "and":
                return np.where((a == 0) | (b == 0), 0.0, a * b) if _nan_in(a, b) else a * b
            if e.op == "or":
                return np.where((a == 1) | (b == 1), 1.0, a + b - a * b) if _nan_in(a, b) else a + b - a * b
            return _compare(e.op, a, b, self.eps)
        if isinstance(e, sc.Unary):
            v = self.expr(e.operand)
            return -v if e.op == "-" else 1.0 - v
        if isinstance(e, sc.Call):
            args = [self.expr(a) for a in e.args]
            if e.func in ("worst", "best"):
                return functools.reduce(np.minimum if e.func == "worst" else np.maximum, args)
            return FUNCS[e.func](*args)
        if isinstance(e, sc.Window):
            return self.window(e)
        raise TypeError(f"cannot evaluate {e!r}")


def leg_values(prog, params, eps, spots, windows, disc):
    """The discounted payoff of every leg: spots {name: (legs, events)}, windows {name: {(start, end, n): (lowest,
    highest)}}, disc(t)."""
    n = len(next(iter(spots.values())))

    def spot(name, k):
        return spots[name][:, k]

    def window(w):
        lo, hi = windows[w.underlying][(w.start, w.end, w.n)]
        return lo if w.func == "lowest" else hi

    return Evaluation(prog, {**prog.params, **(params or {})}, eps, n, spot, window, disc).run()


def pair_mean(values):
    h = len(values) // 2
    pairs = 0.5 * (values[:h] + values[h:])
    return float(pairs.mean()), float(pairs.std(ddof=1) / np.sqrt(h))
