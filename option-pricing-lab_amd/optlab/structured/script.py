"""The payoff language of the structured products (packages/structured/src/optlab/structured/script/nodes.py, lexer.py,
parser.py), in time mode: observation and payment dates are times in years (`t=0.25`), windows read n equally spaced
times in (start, end]. What the fused engine payoff runs (optlab.native.script).

    program  := { underlying NAME | param NAME = [-]NUMBER } { event }
    event    := event t=NUMBER [pay t=NUMBER] NEWLINE { stmt } end
    stmt     := NAME = expr | pay expr as TAG | terminate | if expr then NEWLINE { stmt } [else NEWLINE { stmt }] end
    expr     := or-chains of and-chains of `not`, comparisons (< <= > >= ==), + -, * /, unary -, and primaries:
                NUMBER | NAME | ( expr ) | FUNC ( expr, ... ) | lowest|highest ( NAME , t=a , t=b , n = INT )
    FUNC     := min max (2 args), abs exp log sqrt (1), worst best (2 or more)

A name reads, in this order: a variable assigned earlier in the script, a parameter, the underlying's spot at the
event's observation time. Calendar dates, the issuer's call (`state`, `call pay`) and realised-variance windows are
not part of this subset."""

from __future__ import annotations

import re
from dataclasses import dataclass, field

KEYWORDS = {"underlying", "param", "event", "pay", "as", "end", "if", "then", "else", "terminate", "and", "or", "not"}
FUNCTIONS = {"min": 2, "max": 2, "abs": 1, "exp": 1, "log": 1, "sqrt": 1, "worst": None, "best": None}
WINDOWS = ("lowest", "highest")
TAGS = ("coupon", "upside", "redemption")
COMPARISONS = ("<", "<=", ">", ">=", "==")


class ScriptError(ValueError):
    def __init__(self, message: str, line: int = 0):
        self.message, self.line = message, line
        super().__init__(f"line {line}: {message}" if line else message)


# ---- the tree ------------------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class Num:
    value: float


@dataclass(frozen=True)
class Name:
    id: str
    line: int = field(default=0, compare=False)


@dataclass(frozen=True)
class Unary:
    op: str  # "-" or "not"
    operand: object


@dataclass(frozen=True)
class Binary:
    op: str  # + - * / < <= > >= == and or
    left: object
    right: object
    line: int = field(default=0, compare=False)


@dataclass(frozen=True)
class Call:
    func: str
    args: tuple


@dataclass(frozen=True)
class Window:
    """The lowest or highest spot of an underlying at n equally spaced times in (start, end]."""

    func: str
    underlying: str
    start: float
    end: float
    n: int

    def times(self):
        return [self.start + (self.end - self.start) * k / self.n for k in range(1, self.n + 1)]


@dataclass(frozen=True)
class Assign:
    name: str
    expr: object


@dataclass(frozen=True)
class If:
    cond: object
    body: tuple
    orelse: tuple = ()


@dataclass(frozen=True)
class Pay:
    expr: object
    tag: str


@dataclass(frozen=True)
class Terminate:
    pass


@dataclass(frozen=True)
class Event:
    t_obs: float
    t_pay: float
    body: tuple


@dataclass(frozen=True)
class Program:
    underlyings: tuple
    params: dict
    events: tuple

    def windows(self) -> list:
        """Every distinct window the script reads, in order of first appearance."""
        out = []

        def walk(node):
            if isinstance(node, Window):
                if node not in out:
                    out.append(node)
            elif isinstance(node, (Unary,)):
                walk(node.operand)
            elif isinstance(node, Binary):
                walk(node.left), walk(node.right)
            elif isinstance(node, Call):
                for a in node.args:
                    walk(a)
            elif isinstance(node, Assign):
                walk(node.expr)
            elif isinstance(node, Pay):
                walk(node.expr)
            elif isinstance(node, If):
                walk(node.cond)
                for s in node.body + node.orelse:
                    walk(s)

        for ev in self.events:
            for s in ev.body:
                walk(s)
        return out


# ---- tokens --------------------------------------------------------------------------------------------------------
_TOKENS = re.compile(r"""
    (?P<skip>[ \t\r]+|\#[^\n]*)
  | (?P<nl>\n)
  | (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_]\w*)
  | (?P<op><=|>=|==|[-+*/()<>=,])
""", re.VERBOSE)


def _tokens(text: str) -> list:
    out, line, pos = [], 1, 0
    while pos < len(text):
        m = _TOKENS.match(text, pos)
        if m is None:
            raise ScriptError(f"unexpected character {text[pos]!r}", line)
        kind = m.lastgroup
        if kind == "nl":
            out.append(("nl", "\n", line))
            line += 1
        elif kind == "name":
            out.append(("kw" if m.group() in KEYWORDS else "name", m.group(), line))
        elif kind != "skip":
            out.append((kind, m.group(), line))
        pos = m.end()
    out += [("nl", "\n", line), ("eof", "", line)]
    return out


# ---- the parser ----------------------------------------------------------------------------------------------------
class _Parser:
    def __init__(self, text: str):
        self.toks = _tokens(text)
        self.i = 0

    def peek(self, kind=None, text=None, ahead=0):
        k, t, _ = self.toks[min(self.i + ahead, len(self.toks) - 1)]
        return (kind is None or k == kind) and (text is None or t == text)

    def take(self, kind=None, text=None, what=None):
        if not self.peek(kind, text):
            k, t, line = self.toks[self.i]
            found = {"nl": "end of line", "eof": "end of script"}.get(k, repr(t))
            raise ScriptError(f"expected {what or text or kind}, found {found}", line)
        tok = self.toks[self.i]
        self.i += 1
        return tok

    def line(self):
        return self.toks[self.i][2]

    def newlines(self):
        while self.peek("nl"):
            self.i += 1

    def end_of_line(self):
        if not self.peek("eof"):
            self.take("nl", what="end of line")
        self.newlines()

    def number(self):
        sign = -1.0 if self.peek("op", "-") and self.take() else 1.0
        return sign * float(self.take("num", what="a number")[1])

    def when(self):
        if not (self.peek("name", "t") and self.peek("op", "=", 1)):
            raise ScriptError("a date is written t=<years> here (calendar dates are not in this subset)", self.line())
        self.i += 2
        return self.number()

    def program(self) -> Program:
        unders, params, events = [], {}, []
        self.newlines()
        while self.peek("kw", "underlying") or self.peek("kw", "param"):
            if self.take()[1] == "underlying":
                unders.append(self.take("name", what="an underlying's name")[1])
            else:
                name = self.take("name", what="a parameter's name")[1]
                self.take("op", "=")
                params[name] = self.number()
            self.end_of_line()
        while not self.peek("eof"):
            events.append(self.event())
        if not events:
            raise ScriptError("a script needs at least one event", self.line())
        times = [e.t_obs for e in events]
        if any(b < a for a, b in zip(times, times[1:])):
            raise ScriptError("events must come in order of their times")
        return Program(tuple(unders), params, tuple(events))

    def event(self) -> Event:
        self.take("kw", "event")
        t_obs = self.when()
        t_pay = self.when() if self.peek("kw", "pay") and self.take() else t_obs
        self.end_of_line()
        body = self.block(("end",))
        self.take("kw", "end")
        self.end_of_line()
        return Event(t_obs, t_pay, tuple(body))

    def block(self, stops) -> list:
        out = []
        while not any(self.peek("kw", s) for s in stops):
            if self.peek("eof"):
                raise ScriptError("missing 'end'", self.line())
            out.append(self.statement())
        return out

    def statement(self):
        line = self.line()
        if self.peek("kw", "pay"):
            self.take()
            e = self.expr()
            self.take("kw", "as")
            tag = self.take("name", what="a tag")[1]
            if tag not in TAGS:
                raise ScriptError(f"unknown tag {tag!r}: {', '.join(TAGS)}", line)
            self.end_of_line()
            return Pay(e, tag)
        if self.peek("kw", "terminate"):
            self.take()
            self.end_of_line()
            return Terminate()
        if self.peek("kw", "if"):
            self.take()
            cond = self.expr()
            self.take("kw", "then")
            self.end_of_line()
            body = self.block(("else", "end"))
            orelse = []
            if self.peek("kw", "else"):
                self.take()
                self.end_of_line()
                orelse = self.block(("end",))
            self.take("kw", "end")
            self.end_of_line()
            return If(cond, tuple(body), tuple(orelse))
        name = self.take("name", what="a statement")[1]
        self.take("op", "=")
        e = self.expr()
        self.end_of_line()
        return Assign(name, e)

    def expr(self):
        left = self.conj()
        while self.peek("kw", "or"):
            self.take()
            left = Binary("or", left, self.conj())
        return left

    def conj(self):
        left = self.neg()
        while self.peek("kw", "and"):
            self.take()
            left = Binary("and", left, self.neg())
        return left

    def neg(self):
        if self.peek("kw", "not"):
            self.take()
            return Unary("not", self.neg())
        return self.comparison()

    def comparison(self):
        left = self.arith()
        if self.peek("op") and self.toks[self.i][1] in COMPARISONS:
            op, line = self.take()[1], self.line()
            return Binary(op, left, self.arith(), line)
        return left

    def arith(self):
        left = self.term()
        while self.peek("op", "+") or self.peek("op", "-"):
            left = Binary(self.take()[1], left, self.term())
        return left

    def term(self):
        left = self.unary()
        while self.peek("op", "*") or self.peek("op", "/"):
            left = Binary(self.take()[1], left, self.unary())
        return left

    def unary(self):
        if self.peek("op", "-"):
            self.take()
            return Unary("-", self.unary())
        return self.primary()

    def primary(self):
        line = self.line()
        if self.peek("num"):
            return Num(float(self.take()[1]))
        if self.peek("op", "("):
            self.take()
            e = self.expr()
            self.take("op", ")")
            return e
        name = self.take("name", what="an expression")[1]
        if not self.peek("op", "("):
            return Name(name, line)
        self.take()
        if name in WINDOWS:
            under = self.take("name", what="an underlying")[1]
            self.take("op", ",")
            start = self.when()
            self.take("op", ",")
            end = self.when()
            self.take("op", ",")
            if not (self.peek("name", "n") and self.peek("op", "=", 1)):
                raise ScriptError("a window in time mode needs n=<count>", line)
            self.i += 2
            n = int(self.take("num", what="a count")[1])
            self.take("op", ")")
            if n < 1 or not end > start:
                raise ScriptError("a window needs n >= 1 and end after start", line)
            return Window(name, under, start, end, n)
        if name not in FUNCTIONS:
            raise ScriptError(f"unknown function {name!r}", line)
        args = [self.expr()]
        while self.peek("op", ","):
            self.take()
            args.append(self.expr())
        self.take("op", ")")
        want = FUNCTIONS[name]
        if (want is not None and len(args) != want) or (want is None and len(args) < 2):
            raise ScriptError(f"{name} takes {want or '2 or more'} arguments, not {len(args)}", line)
        return Call(name, tuple(args))


def parse(text: str) -> Program:
    """The program of a script in time mode; names resolve as the module docstring says, and a name that is neither
    assigned before, nor a parameter, nor an underlying is an error."""
    prog = _Parser(text).program()
    _check_names(prog)
    return prog


def _check_names(prog: Program) -> None:
    assigned: set = set()

    def read(node, line=0):
        if isinstance(node, Name):
            if node.id not in assigned and node.id not in prog.params and node.id not in prog.underlyings:
                raise ScriptError(f"{node.id!r} is not assigned before, a parameter or an underlying", node.line)
        elif isinstance(node, Window):
            if node.underlying not in prog.underlyings:
                raise ScriptError(f"{node.underlying!r} is not an underlying")
        elif isinstance(node, Unary):
            read(node.operand)
        elif isinstance(node, Binary):
            read(node.left), read(node.right)
        elif isinstance(node, Call):
            for a in node.args:
                read(a)

    def walk(body):
        for s in body:
            if isinstance(s, Assign):
                read(s.expr)
                assigned.add(s.name)
            elif isinstance(s, Pay):
                read(s.expr)
            elif isinstance(s, If):
                read(s.cond)
                walk(s.body)
                walk(s.orelse)

    for ev in prog.events:
        walk(ev.body)


# ---- two products of the preset list (presets.py:7-31), written in time mode ----------------------------------------
def _num(x: float) -> str:
    return repr(float(x))


def phoenix_script(names, maturity: float = 2.0, frequency: float = 0.25, coupon_rate: float = 0.08,
                   coupon_barrier: float = 0.7, trigger: float = 1.0, put_strike: float = 1.0, ki_barrier: float = 0.6,
                   daily_ki: int = 0) -> str:
    """A worst-of phoenix (compiler.py:270-295 with presets.py:24-31): conditional coupon with memory every
    `frequency` years, autocall at `trigger` on every date before the last, at maturity a put knocked in below
    ki_barrier -- on the final performance, or (daily_ki = n) on the lowest of n equally spaced closes."""
    names = list(names)
    basket = len(names) > 1
    lines = [f"underlying {n}" for n in names]
    lines += [f"param N = 1.0", f"param cpn = {_num(coupon_rate * frequency)}",
              f"param coupon_barrier = {_num(coupon_barrier)}", f"param trigger = {_num(trigger)}",
              f"param put_strike = {_num(put_strike)}", f"param ki_barrier = {_num(ki_barrier)}", "", "event t=0"]
    lines += [f"  S0_{n} = {n}" for n in names] if basket else [f"  S0 = {names[0]}"]
    lines += ["  missed = 0", "end"]
    perf = f"worst({', '.join(f'{n} / S0_{n}' for n in names)})" if basket else f"{names[0]} / S0"
    n_dates = int(round(maturity / frequency))
    for j in range(1, n_dates + 1):
        t = _num(maturity * j / n_dates)
        body = [f"perf = {perf}", "if perf >= coupon_barrier then", "  pay N * cpn * (1 + missed) as coupon",
                "  missed = 0", "else", "  missed = missed + 1", "end"]
        if j < n_dates:
            body += ["if perf >= trigger then", "  pay N as redemption", "  terminate", "end"]
        else:
            if daily_ki and basket:
                lows = ", ".join(f"lowest({n}, t=0, t={t}, n={daily_ki}) / S0_{n}" for n in names)
                body.append(f"knocked_in = worst({lows}) < ki_barrier")
            elif daily_ki:
                body.append(f"knocked_in = lowest({names[0]}, t=0, t={t}, n={daily_ki}) < ki_barrier * S0")
            else:
                body.append("knocked_in = perf < ki_barrier")
            body += ["if knocked_in then", "  pay N - N * max(put_strike - perf, 0) as redemption", "else",
                     "  pay N as redemption", "end"]
        lines += ["", f"event t={t}", *("  " + b for b in body), "end"]
    return "\n".join(lines) + "\n"


def barrier_reverse_convertible_script(name: str, maturity: float = 1.0, frequency: float = 0.25,
                                       coupon_rate: float = 0.08, put_strike: float = 1.0, ki_barrier: float = 0.7,
                                       daily_ki: int = 252) -> str:
    """A barrier reverse convertible (presets.py:14-21): a fixed coupon every `frequency` years, at maturity the
    notional less a put at put_strike knocked in when a close of the daily window fell below ki_barrier."""
    lines = [f"underlying {name}", "param N = 1.0", f"param cpn = {_num(coupon_rate * frequency)}",
             f"param put_strike = {_num(put_strike)}", f"param ki_barrier = {_num(ki_barrier)}", "", "event t=0",
             f"  S0 = {name}", "end"]
    n_dates = int(round(maturity / frequency))
    for j in range(1, n_dates + 1):
        t = _num(maturity * j / n_dates)
        body = ["pay N * cpn as coupon"]
        if j == n_dates:
            body = [f"perf = {name} / S0", *body,
                    f"knocked_in = lowest({name}, t=0, t={t}, n={daily_ki}) < ki_barrier * S0",
                    "if knocked_in then", "  pay N - N * max(put_strike - perf, 0) as redemption", "else",
                    "  pay N as redemption", "end"]
        lines += ["", f"event t={t}", *("  " + b for b in body), "end"]
    return "\n".join(lines) + "\n"
