"""Pins of the risk oracles without a GPU: the tents of the risk grid and the tent-bumped Dupire table
(oracle/surface_oracle.py) against what the mathematics fixes (partitions of unity, finite differences, the
unbumped surface's own Dupire), and the Phoenix script (oracle/autocall_oracle.py) on paths whose value is known by
hand."""

from __future__ import annotations

import sys

import numpy as np
import pytest
from _support import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import autocall_oracle as ao  # noqa: E402
import surface_oracle as so  # noqa: E402

from optlab.core.ssvi import SSVI  # noqa: E402

PILLARS = np.array([0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0])


def _surface(scale=0.04, rho=-0.6):
    return SSVI.power_law(rho, 1.2, 0.5, PILLARS, scale * PILLARS)


def test_tenor_tents_are_a_smooth_partition_of_unity():
    T = np.concatenate([np.linspace(0.0, 6.0, 2001), np.asarray(so.TENORS)])
    psi, dpsi = so.tenor_tents(T)
    np.testing.assert_allclose(psi.sum(axis=0), 1.0, atol=1e-15)
    np.testing.assert_allclose(dpsi.sum(axis=0), 0.0, atol=1e-12)
    assert np.all(psi >= 0) and np.all(psi <= 1)
    for j, t in enumerate(so.TENORS):  # each tent is one at its own node
        p, _ = so.tenor_tents(np.array([t]))
        assert p[j, 0] == 1.0 and p.sum() == 1.0
    h = 1e-6  # the derivative is that of the tents, away from the nodes
    T = np.array([0.1, 0.3, 0.7, 1.2, 1.8, 2.4, 4.0])
    fd = (so.tenor_tents(T + h)[0] - so.tenor_tents(T - h)[0]) / (2 * h)
    np.testing.assert_allclose(so.tenor_tents(T)[1], fd, atol=1e-7)


def test_strike_tents_sum_to_one_and_their_derivatives_are_exact():
    k = np.linspace(-1.5, 1.0, 301)
    for T in (0.01, 0.3, 1.0, 2.7):
        phi = so.strike_tents(k, np.full_like(k, T))
        np.testing.assert_allclose(phi[0].sum(axis=0), 1.0, atol=1e-13)
        for d in (1, 2, 3):
            np.testing.assert_allclose(phi[d].sum(axis=0), 0.0, atol=1e-10)
        h = 1e-6
        up, dn = so.strike_tents(k + h, np.full_like(k, T)), so.strike_tents(k - h, np.full_like(k, T))
        np.testing.assert_allclose(phi[1], (up[0] - dn[0]) / (2 * h), atol=1e-6 * np.abs(phi[1]).max())
        np.testing.assert_allclose(phi[2], (up[1] - dn[1]) / (2 * h), atol=1e-6 * np.abs(phi[2]).max())
        if T > so.T_FLOOR + 1e-4:
            up, dn = so.strike_tents(k, np.full_like(k, T + h)), so.strike_tents(k, np.full_like(k, T - h))
            np.testing.assert_allclose(phi[3], (up[0] - dn[0]) / (2 * h), atol=1e-6 * np.abs(phi[3]).max())
    # far from every node the smoothing has died away: the hats themselves
    far = so.strike_tents(np.array([-2.0, 2.0]), np.array([1.0, 1.0]))[0]
    np.testing.assert_allclose(far[:, 0], np.eye(10)[0], atol=1e-12)
    np.testing.assert_allclose(far[:, 1], np.eye(10)[-1], atol=1e-12)


def test_unbumped_table_is_the_surface_dupire_and_a_bump_of_every_node_is_the_parallel_bump():
    surface = _surface()
    t0 = np.arange(0, 756, 37) / 252
    rows = so.base_rows(surface, t0, n_x=201)
    lo, inv, k, T = rows[:4]
    base = so.bumped_local_var(rows, np.zeros((10, 8)))
    np.testing.assert_allclose(base, surface.local_var(k, T[:, None]), rtol=1e-12)
    all_nodes = so.bumped_local_var(rows, np.full((10, 8), 0.01))
    np.testing.assert_allclose(all_nodes, so.bumped_local_var(rows, np.zeros((10, 8)), parallel=0.01), rtol=1e-12)
    assert not np.allclose(all_nodes, base)


def test_bumped_table_matches_dupire_by_finite_differences_of_the_bumped_vols():
    """sigma' = sigma + sum H phi psi + skew tilt, w' = T sigma'^2, local variance from w' by central differences in
    k and T: agrees with the analytic table (a dropped term of w'_k, w'_kk or w'_T would not)."""
    surface = _surface(0.05, -0.7)
    rng = np.random.default_rng(3)
    H = rng.normal(0.0, 0.01, (10, 8))
    skew = 0.02
    t0 = np.array([0.2, 0.6, 1.1, 1.7, 2.6])
    rows = so.base_rows(surface, t0, n_x=61, sd_range=(-3.0, 2.5))
    lo, inv, k, T = rows[:4]
    got = so.bumped_local_var(rows, H, skew=skew, cap=1e9, floor=-1e9)

    def wprime(kk, TT):
        phi = so.strike_tents(kk, TT)[0]
        psi, _ = so.tenor_tents(TT)
        a = np.einsum("ij,i...,j...->...", H, phi, psi) + so.skew_bump(skew, kk, TT)[0]
        return TT * (np.sqrt(surface.w(kk, TT) / TT) + a) ** 2

    Tm = np.broadcast_to(T[:, None], k.shape)
    hk, hT = 1e-4, 1e-5
    w = wprime(k, Tm)
    wk = (wprime(k + hk, Tm) - wprime(k - hk, Tm)) / (2 * hk)
    wkk = (wprime(k + hk, Tm) - 2 * w + wprime(k - hk, Tm)) / hk**2
    wT = (wprime(k, Tm + hT) - wprime(k, Tm - hT)) / (2 * hT)
    g = (1 - k * wk / (2 * w)) ** 2 - wk**2 / 4 * (1 / w + 0.25) + wkk / 2
    sound = g > 0.05  # away from butterfly arbitrage, where wT / g is well conditioned
    assert sound.mean() > 0.9
    np.testing.assert_allclose(got[sound], (wT / g)[sound], rtol=2e-5)


# ---- the Phoenix script ----------------------------------------------------------------------------------------


def _closes(perf_path):
    """One name, closes = 100 * perf path (fixing 100)."""
    return 100.0 * np.asarray(perf_path, dtype=float)[None, :, None]


TERMS = dict(coupon=0.02, coupon_barrier=0.7, autocall=1.0, memory=True, ki_barrier=0.6, put_strike=1.0)


def _value(path, eps=0.0, **over):
    t = {**TERMS, **over}
    obs = np.array([1, 2, 3, 4])
    disc = np.array([0.99, 0.98, 0.97, 0.96])
    return ao.phoenix(_closes(path), [100.0], obs, disc, t["coupon"], t["coupon_barrier"], t["autocall"], t["memory"],
                      t["ki_barrier"], t["put_strike"], eps)[0]


def test_phoenix_script_on_paths_known_by_hand():
    # called at the first date: its coupon and par
    assert _value([1.0, 1.05, 0.5, 0.5, 0.5]) == pytest.approx(0.99 * (0.02 + 1.0), abs=1e-15)
    # below the coupon barrier three times, then at 0.8 at maturity: four coupons paid then (memory), par, no KI
    assert _value([1.0, 0.65, 0.66, 0.67, 0.8]) == pytest.approx(0.96 * (4 * 0.02 + 1.0), abs=1e-15)
    # the same without memory: one coupon
    assert _value([1.0, 0.65, 0.66, 0.67, 0.8], memory=False) == pytest.approx(0.96 * (0.02 + 1.0), abs=1e-15)
    # a close at 0.55 knocks in; maturity at 0.5: 1 - (1 - 0.5), no coupon
    assert _value([1.0, 0.55, 0.9, 0.9, 0.5]) == pytest.approx(0.98 * 0.04 + 0.97 * 0.02 + 0.96 * 0.5, abs=1e-15)
    # knocked in, 1 % under the strike at maturity: the coupon and 99 %; at the strike: par
    assert _value([1.0, 0.55, 0.9, 0.9, 0.99]) == pytest.approx(0.98 * 0.04 + 0.97 * 0.02 + 0.96 * 1.01, abs=1e-15)
    assert _value([1.0, 0.55, 0.9, 0.9, 1.0]) == pytest.approx(0.98 * 0.04 + 0.97 * 0.02 + 0.96 * 1.02, abs=1e-15)
    # worst of two names: the lower one decides
    c = np.stack([_closes([1.0, 1.2, 1.2, 1.2, 1.2])[0], _closes([1.0, 0.95, 1.01, 1.0, 1.0])[0]])
    v = ao.phoenix(c, [100.0, 100.0], np.array([1, 2, 3, 4]), np.array([0.99, 0.98, 0.97, 0.96]), 0.02, 0.7, 1.0,
                   True, 0.6, 1.0, 0.0)[0]
    assert v == pytest.approx(0.99 * 0.02 + 0.98 * 1.02, abs=1e-15)


def test_phoenix_smoothing_is_the_evaluator_ramp():
    eps = 0.005
    # exactly on the call trigger: half called
    called_half = _value([1.0, 1.0, 1.1, 0.5, 0.5], eps=eps)
    assert called_half == pytest.approx(0.99 * (0.02 + 0.5) + 0.5 * 0.98 * 1.02, abs=1e-13)
    # a quarter of the ramp above the trigger: 3/4 called
    up = 1.0 + 0.25 * eps
    v = _value([1.0, up, 1.1, 0.5, 0.5], eps=eps)
    assert v == pytest.approx(0.99 * (0.02 + 0.75) + 0.25 * 0.98 * 1.02, abs=1e-12)
    # outside the ramp the smoothed script is the hard one
    for path in ([1.0, 0.65, 0.66, 0.67, 0.8], [1.0, 0.55, 0.9, 0.9, 0.5]):
        assert _value(path, eps=eps) == pytest.approx(_value(path), abs=1e-15)


def test_var_es_reads_an_actual_scenario():
    from optlab.risk.book import var_es

    x = np.arange(-50.0, 50.0)  # 100 scenarios: losses down to -50
    r = var_es(x)
    assert r["var"] == 50.0  # the 1 % "lower" quantile of 100 points is the worst one
    assert r["es"] == pytest.approx(np.mean([50.0, 49.0, 48.0]))  # beyond the 2.5 % quantile: 3 scenarios
    assert x[r["var_index"]] == -50.0


def test_tents_equal_the_reference_grid_vectors():
    """The reference's own strike tents, their smoothing width and the tenor tents on a (k, T) set that crosses every
    node, the width's floor and the ends (tests/golden/make_tent_fixtures.py wrote them from risk/grid.py)."""
    g = np.load(ROOT / "tests" / "golden" / "risk_tents.npz")
    k, T = g["k"], g["T"]
    np.testing.assert_allclose(np.stack(so.width(k, T)), g["smoothing_width"], rtol=1e-14, atol=1e-15)
    got = so.strike_tents(k, T)
    for d in range(4):
        scale = np.abs(g["strike_tents"][d]).max()
        np.testing.assert_allclose(got[d], g["strike_tents"][d], rtol=1e-14, atol=1e-14 * scale)
    psi, dpsi = so.tenor_tents(g["tenor_T"])
    np.testing.assert_allclose(psi, g["tenor_tents"][0], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(dpsi, g["tenor_tents"][1], rtol=1e-14, atol=1e-14)
