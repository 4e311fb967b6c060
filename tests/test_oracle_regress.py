"""The issuer's call regression oracle (oracle/regress_oracle.py) pinned on the CPU: the monomial exponents against a
brute-force enumeration; the terms against hand values; the normal equations against lstsq and against an exact
polynomial; the choice of terms against exercise.py:126-140's rules."""

from __future__ import annotations

import itertools
import sys

import numpy as np
from _support import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import regress_oracle as ro  # noqa: E402


def test_powers_are_every_monomial_up_to_the_degree_constant_first():
    for d, degree in ((1, 3), (2, 2), (3, 2), (5, 1), (4, 0)):
        got = ro.powers(d, degree)
        want = {p for p in itertools.product(range(degree + 1), repeat=d) if sum(p) <= degree}
        assert set(got) == want and len(got) == len(want)
        assert got[0] == (0,) * d and [sum(p) for p in got] == sorted(sum(p) for p in got)


def test_terms_by_hand():
    state = np.array([[1.0, 2.0], [3.0, -1.0]])
    mean, scale = np.array([1.0, 0.5]), np.array([2.0, 0.5])
    z = (state - mean) / scale  # [[0, 3], [1, -3]]
    pw = [(0, 0), (1, 0), (0, 2), (1, 1), (3, 0), (0, 3)]
    knots = [(1, 0.5), (0, -0.25)]
    X = ro.features(state, mean, scale, pw, knots)
    want = np.column_stack([np.ones(2), z[:, 0], z[:, 1] ** 2, z[:, 0] * z[:, 1], z[:, 0] ** 3, z[:, 1] ** 3,
                            np.maximum(z[:, 1] - 0.5, 0), np.maximum(z[:, 0] + 0.25, 0)])
    np.testing.assert_array_equal(X, want)


def test_normal_equations_are_lstsq_and_recover_a_polynomial():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4000, 2))
    pw = ro.powers(2, 2)
    X = ro.features(x, 0.0, 1.0, pw, [(0, 0.3), (1, -0.4)])
    beta = rng.standard_normal(X.shape[1])
    np.testing.assert_allclose(ro.fit(X, X @ beta), beta, rtol=1e-9, atol=1e-10)
    y = X @ beta + 0.1 * rng.standard_normal(len(X))
    np.testing.assert_allclose(ro.fit(X, y), ro.fit(X, y, "lstsq"), rtol=1e-9, atol=1e-11)


def test_basis_follows_the_room_the_paths_allow():
    rng = np.random.default_rng(2)
    pw, kn = ro.basis(rng.standard_normal((100, 1)))  # room 5: the cubic (4 terms) and one hinge
    assert len(pw) == 4 and len(kn) == 1
    pw, kn = ro.basis(rng.standard_normal((5000, 3)))  # room 25: the quadratic (10) and 5 hinges a variable
    assert len(pw) == 10 and len(kn) == 15
    pw, kn = ro.basis(rng.standard_normal((60, 3)))  # room 3: degree 2 (10) and 1 (4) do not fit, the constant
    assert pw == [(0, 0, 0)] and len(kn) == 0


def test_the_product_s_term_choice_is_the_oracle_s():
    from optlab.structured.exercise import powers, terms

    rng = np.random.default_rng(3)
    for n, d in ((100, 1), (5000, 3), (60, 3), (20_000, 1), (50_000, 5), (800, 2)):
        z = rng.standard_normal((n, d)) * rng.uniform(0.5, 2.0, d)
        assert terms(z) == ro.basis(z)
        assert powers(d, 2) == ro.powers(d, 2)
