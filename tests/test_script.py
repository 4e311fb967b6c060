"""The payoff language on the host (no GPU): the parser on the two products of the preset list; the oracle's evaluator
(evaluator.py restated) against the hand-written phoenix of the autocall oracle and a barrier reverse convertible
written out in NumPy; the lowered source's leg valuation compiled for the CPU (g++) against the oracle's evaluator on
random path arrays, hard and smoothed; and the whole lowered source compiled by hipRTC for gfx950."""

from __future__ import annotations

import ctypes as C
import subprocess
import sys

import numpy as np
import pytest
from _support import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import autocall_oracle as ao  # noqa: E402
import script_oracle as so  # noqa: E402

from optlab.core.market import Market  # noqa: E402
from optlab.core.models import BlackScholesModel  # noqa: E402
from optlab.structured import script as sc  # noqa: E402
from optlab.structured.multiasset import Basket  # noqa: E402

NAMES = ("A", "B", "C")


def phoenix_pricer(daily=24, smoothing=0.0):
    from optlab.native.script import ScriptPricer

    basket = Basket(NAMES, {n: BlackScholesModel(v) for n, v in zip(NAMES, (0.2, 0.25, 0.3))},
                    {n: Market.flat(100.0, 0.03, 0.0) for n in NAMES}, np.full((3, 3), 0.5) + 0.5 * np.eye(3))
    text = sc.phoenix_script(NAMES, 2.0, 0.25, daily_ki=daily)
    return ScriptPricer(text, basket, smoothing=smoothing), basket


def test_the_parser_reads_the_presets_and_locates_errors():
    p = sc.parse(sc.phoenix_script(NAMES, 3.0, daily_ki=756))
    assert p.underlyings == NAMES and len(p.events) == 13 and p.params["cpn"] == pytest.approx(0.02)
    assert [e.t_obs for e in p.events][:3] == [0.0, 0.25, 0.5]
    assert len(p.windows()) == 3 and {(w.start, w.end, w.n) for w in p.windows()} == {(0.0, 3.0, 756)}
    q = sc.parse(sc.barrier_reverse_convertible_script("X"))
    assert len(q.events) == 5 and q.windows()[0].times()[-1] == pytest.approx(1.0)
    for bad, what in (("underlying X\nevent t=1\n  pay Y as coupon\nend\n", "'Y'"),
                      ("underlying X\nevent t=1\n  pay X as bonus\nend\n", "unknown tag"),
                      ("underlying X\nevent 2024-01-01\nend\n", "t=<years>"),
                      ("underlying X\nevent t=1\n  a = min(X)\nend\n", "min takes 2"),
                      ("underlying X\nevent t=1\n  if X > 1 then\n  pay X as coupon\nend\n", "missing 'end'")):
        with pytest.raises(sc.ScriptError, match=what):
            sc.parse(bad)


def _closes(n_paths, n_steps, seed=3):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((3, n_steps, n_paths))
    x = np.cumsum(-0.5 * 0.25**2 / 252 * 3 + 0.25 * np.sqrt(3 / 252) * z, axis=1)
    return 100.0 * np.exp(np.concatenate([np.zeros((3, 1, n_paths)), x], axis=1))


def test_the_oracle_evaluator_on_the_phoenix_script_is_the_hand_written_phoenix():
    """Hard comparisons (the script's memory coupon and the hand-written one part under smoothing: the script's else
    branch reads the missed count its then branch just moved)."""
    n_steps = 168  # two years of steps of 3/252, the window every step
    c = _closes(4000, n_steps)
    prog = sc.parse(sc.phoenix_script(NAMES, 2.0, 0.25, daily_ki=n_steps))
    obs = [round(j * n_steps / 8) for j in range(1, 9)]
    D = lambda t: float(np.exp(-0.03 * t))  # noqa: E731
    spots = {n: c[a][[0] + obs].T for a, n in enumerate(NAMES)}
    wins = {n: {(0.0, 2.0, n_steps): (c[a, 1:].min(axis=0), c[a, 1:].max(axis=0))} for a, n in enumerate(NAMES)}
    got = so.leg_values(prog, None, 0.0, spots, wins, D)
    want = ao.phoenix(c, c[:, 0, 0], obs, [D(t) for t in np.arange(1, 9) * 0.25], 0.02, 0.7, 1.0, True, 0.6, 1.0, 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    assert 0.05 < np.mean(got < 0.99) < 0.95  # both outcomes are on the paths


def test_the_oracle_evaluator_on_the_brc_script_is_the_brc():
    c = _closes(3000, 84)[0]  # one name, a year of steps of 3/252
    prog = sc.parse(sc.barrier_reverse_convertible_script("X", 1.0, daily_ki=84))
    D = lambda t: float(np.exp(-0.03 * t))  # noqa: E731
    obs = [0, 21, 42, 63, 84]
    got = so.leg_values(prog, None, 0.0, {"X": c[obs].T}, {"X": {(0.0, 1.0, 84): (c[1:].min(0), c[1:].max(0))}}, D)
    perf, knocked = c[84] / c[0], c[1:].min(0) < 0.7 * c[0]
    want = sum(0.02 * D(t) for t in (0.25, 0.5, 0.75, 1.0)) + D(1.0) * np.where(knocked, 1 - np.maximum(1 - perf, 0), 1)
    np.testing.assert_allclose(got, want, rtol=1e-14)
    assert 0.05 < knocked.mean() < 0.95


@pytest.fixture(scope="module")
def host_leg(tmp_path_factory):
    """The lowered sources' leg valuation built for the CPU: __device__ code compiled by g++ with its qualifiers
    defined away, called through an extern "C" wrapper."""

    def build(source, tag):
        body = source.split('extern "C" __global__')[0]
        text = ("#include <cmath>\nusing std::exp; using std::log; using std::sqrt; using std::fabs; "
                "using std::abs;\n#define __device__\n#define __forceinline__ inline\n#define __restrict__\n" + body +
                '\nextern "C" double host_leg(const double* out, long long row, long long legs) '
                "{ return leg_value(out, row, legs); }\n")
        d = tmp_path_factory.mktemp(tag)
        (d / "leg.cpp").write_text(text)
        subprocess.run(["g++", "-O1", "-ffp-contract=off", "-shared", "-fPIC", str(d / "leg.cpp"), "-o",
                        str(d / "leg.so")], check=True)
        lib = C.CDLL(str(d / "leg.so"))
        lib.host_leg.restype = C.c_double
        lib.host_leg.argtypes = [C.POINTER(C.c_double), C.c_longlong, C.c_longlong]
        return lib

    return build


@pytest.mark.parametrize("smoothing", [0.0, 0.02])
def test_the_lowered_leg_valuation_is_the_oracle_evaluator(host_leg, smoothing):
    pricer, basket = phoenix_pricer(daily=24, smoothing=smoothing)
    lib = host_leg(pricer.source, f"phoenix{int(smoothing * 100)}")
    legs = 4000
    rng = np.random.default_rng(7)
    n_cols = pricer.n_cols
    out = 100.0 * np.exp(rng.normal(0.0, 0.2, (3, n_cols, legs)))
    out[:, 0] = 100.0  # the spots at t = 0
    for a in range(3):  # a window's lowest below, its highest above, every kept spot
        out[a, n_cols - 2] = np.minimum(out[a, n_cols - 2], out[a, :n_cols - 2].min(axis=0))
        out[a, n_cols - 1] = np.maximum(out[a, n_cols - 1], out[a, :n_cols - 2].max(axis=0))
    flat = np.ascontiguousarray(out).ravel()
    ptr = flat.ctypes.data_as(C.POINTER(C.c_double))
    got = np.array([lib.host_leg(ptr, i, legs) for i in range(legs)])
    spots = {n: out[a][pricer.spot_col].T for a, n in enumerate(NAMES)}
    wins = {n: {pricer.spans[0]: (out[a, n_cols - 2], out[a, n_cols - 1])} for a, n in enumerate(NAMES)}
    D = basket.markets["A"].D
    want = so.leg_values(pricer.program, None, smoothing, spots, wins, lambda t: float(D(t)))
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    if smoothing:
        assert np.any((want % 0.02 != 0) & (want < 0.99))  # ramps were actually crossed


def test_the_lowered_source_compiles_with_hiprtc_for_gfx950():
    lib = C.CDLL("libhiprtc.so")
    pricer, _ = phoenix_pricer(daily=24, smoothing=0.005)
    prog = C.c_void_p()
    src = pricer.source.encode()
    assert lib.hiprtcCreateProgram(C.byref(prog), src, b"script.hip", 0, None, None) == 0
    opts = (C.c_char_p * 3)(b"--offload-arch=gfx950", b"-O3", b"-std=c++17")
    rc = lib.hiprtcCompileProgram(prog, 3, opts)
    size = C.c_size_t()
    lib.hiprtcGetProgramLogSize(prog, C.byref(size))
    log = C.create_string_buffer(size.value + 1)
    lib.hiprtcGetProgramLog(prog, log)
    assert rc == 0, log.value.decode(errors="replace")
    lib.hiprtcGetCodeSize(prog, C.byref(size))
    assert size.value > 1000
    lib.hiprtcDestroyProgram(C.byref(prog))
