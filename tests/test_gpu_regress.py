"""The issuer's call regression on the device (optlab_regress; exercise.py:191-238): FP64 coefficients against the
oracle's normal equations on the same rows (their backward error on the oracle's system), the fitted values against
the reference's lstsq, the continuation on a second set of states against the oracle's terms times the device's
coefficients; a singular design reported, not solved; determinism and checked requests."""

from __future__ import annotations

import sys

import numpy as np
import pytest
from _support import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import regress_oracle as ro  # noqa: E402

from optlab import native  # noqa: E402

pytestmark = pytest.mark.gpu


def pilot(n, d, seed):
    """Spots of d names at a call date (lognormal, correlated) and a note's continuation seen from the issuer: a
    coupon stream plus the worst-of's kinks at the knock-in and the call level, with noise."""
    rng = np.random.default_rng(seed)
    c = np.full((d, d), 0.6) + 0.4 * np.eye(d)
    s = np.exp(0.25 * rng.standard_normal((n, d)) @ np.linalg.cholesky(c).T - 0.03)
    w = s.min(axis=1)
    y = 0.9 + 0.05 * np.minimum(w, 1.2) - 0.8 * np.maximum(0.6 - w, 0.0) + 0.02 * rng.standard_normal(n)
    return s, y


@pytest.mark.parametrize("d,n", [(1, 20_000), (3, 100_000), (5, 50_000)])
def test_fp64_fit_is_the_oracle_s_normal_equations(d, n):
    x, y = pilot(n, d, d)
    mean, scale = x.mean(axis=0), x.std(axis=0)
    pw, kn = ro.basis((x - mean) / scale)
    ev, _ = pilot(3 * n, d, 100 + d)
    print(f"regress d={d} n={n}: {len(pw)} powers, {len(kn)} hinges", flush=True)
    got = native.regress(x, y, mean, scale, pw, kn, eval_state=ev, backend="fp64")
    assert got.coef is not None and got.coef.shape == (len(pw) + len(kn),)
    X = ro.features(x, mean, scale, pw, kn)
    want = ro.fit(X, y)
    A = X.T @ X
    # another summation order moves the coefficients by the conditioning times the rounding: what a correct solve
    # fixes is its backward error on the oracle's system (about 1e-16 for another summation order)
    back = np.linalg.norm(A @ (got.coef - want)) / (np.linalg.norm(A) * np.linalg.norm(want))
    assert back < 1e-12, back
    # the fitted continuation is the reference lstsq's (condition numbers 3e3 to 6e5 here: 1e-12 apart on the CPU)
    ref = ro.fit(X, y, "lstsq")
    assert np.max(np.abs(X @ got.coef - X @ ref)) < 1e-10
    # the continuation on the second set: the oracle's terms times the device's coefficients
    Xe = ro.features(ev, mean, scale, pw, kn)
    np.testing.assert_allclose(got.values, Xe @ got.coef, rtol=1e-12, atol=1e-13)
    again = native.regress(x, y, mean, scale, pw, kn, eval_state=ev, backend="fp64")
    assert np.array_equal(again.coef, got.coef) and np.array_equal(again.values, got.values)


def test_an_exact_polynomial_is_recovered():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((30_000, 2))
    pw = ro.powers(2, 2)
    kn = [(0, 0.25), (1, -0.5)]
    beta = rng.standard_normal(len(pw) + len(kn))
    y = ro.features(x, 0.0, 1.0, pw, kn) @ beta
    got = native.regress(x, y, [0.0, 0.0], [1.0, 1.0], pw, kn)
    np.testing.assert_allclose(got.coef, beta, rtol=1e-9, atol=1e-10)


def test_a_singular_design_is_reported():
    x, y = pilot(5000, 2, 4)
    pw = [(0, 0), (1, 0), (1, 0)]  # a term twice: X'X is singular
    got = native.regress(x, y, x.mean(axis=0), x.std(axis=0), pw, [], eval_state=x[:10])
    assert got.coef is None and got.values is None


def test_requests_are_checked():
    x, y = pilot(100, 2, 5)
    with pytest.raises(ValueError, match="powers"):
        native.regress(x, y, 0.0, 1.0, [(4, 0)], [])
    with pytest.raises(ValueError, match="knot_var"):
        native.regress(x, y, 0.0, 1.0, [(0, 0)], [(2, 0.0)])
    with pytest.raises(ValueError, match="scale"):
        native.regress(x, y, 0.0, [1.0, 0.0], [(0, 0)], [])
    with pytest.raises(ValueError, match="1 to 25"):
        native.regress(x, y, 0.0, 1.0, ro.powers(2, 6), [])
