"""CPU: the credibility statistics and report (strouhal_amd/credibility.py),
pinned by closed forms and designed series rather than by restating them."""
from __future__ import annotations

import math

import numpy as np
import pytest

from strouhal_amd import credibility as cr


def test_batch_means_of_a_designed_series():
    # 16 blocks of 300 equal values, block values +1 +1 -1 -1 ...: mean 0,
    # batch means sd sqrt(16/15), lag-1 sum of products +1 over sum of squares 16
    blocks = np.array([1, 1, -1, -1] * 4, dtype=float)
    v = np.repeat(blocks, 300)
    s = cr.StatUncertainty.of(v)
    assert s.batches == 16
    assert s.mean == 0.0
    assert s.lag1 == pytest.approx(1.0 / 16.0, abs=1e-15)
    assert s.half_width95 == pytest.approx(1.96 * math.sqrt(16.0 / 15.0) / 4.0, rel=1e-14)
    assert s.trustworthy


def test_batch_means_of_white_noise_and_of_a_slow_drift():
    rng = np.random.default_rng(3)
    v = rng.normal(2.0, 1.0, 48000)
    s = cr.StatUncertainty.of(v)
    assert s.trustworthy and s.batches == 16
    assert abs(s.mean - 2.0) < 4 * 1.0 / math.sqrt(v.size)
    # the half-width of the mean of n iid samples: 1.96 sigma / sqrt(n)
    assert s.half_width95 == pytest.approx(1.96 / math.sqrt(v.size), rel=0.4)
    # a slow sine drift keeps neighbouring batch means correlated at any count
    t = np.arange(4800)
    slow = cr.StatUncertainty.of(np.sin(2 * math.pi * t / 4800.0))
    assert slow.batches == 6 and not slow.trustworthy and slow.lag1 > 0.3
    short = cr.StatUncertainty.of(np.ones(23))
    assert short.batches == 0 and math.isnan(short.half_width95)


def test_observed_order_and_finest_pair_spread():
    # phi(n) = phi* + C / n^2 on n = 16, 32, 64: order exactly 2
    rungs = [cr.Rung(n, 3.2 + 40.0 / n ** 2, 1e-4) for n in (64, 16, 32)]
    num = cr.NumUncertainty.of(rungs)
    assert [r.cells_per_feature for r in num.rungs] == [16, 32, 64]
    assert num.observed_order == pytest.approx(2.0, rel=1e-12)
    assert num.u_num == pytest.approx(40.0 / 32 ** 2 - 40.0 / 64 ** 2, rel=1e-12)
    # agreement within the finest rung's noise: u_num is the noise
    quiet = cr.NumUncertainty.of([cr.Rung(32, 1.0, 0.01), cr.Rung(64, 1.004, 0.01)])
    assert quiet.u_num == 0.01 and "within statistical noise" in quiet.note
    # non-monotone and non-converging ladders give no order
    assert cr.NumUncertainty.of([cr.Rung(16, 1.0, 0), cr.Rung(32, 1.1, 0),
                                 cr.Rung(64, 1.05, 0)]).observed_order is None
    assert cr.NumUncertainty.of([cr.Rung(16, 1.0, 0), cr.Rung(32, 1.5, 0),
                                 cr.Rung(64, 2.0, 0)]).observed_order is None
    # first-order convergence on a ratio-2 ladder
    first = cr.NumUncertainty.of([cr.Rung(n, 1.0 + 1.0 / n, 0) for n in (16, 32, 64)])
    assert first.observed_order == pytest.approx(1.0, rel=1e-12)
    assert math.isnan(cr.NumUncertainty.of([cr.Rung(16, 1.0, 0)]).u_num)


def test_mach_residual_of_a_quadratic_compressibility_error():
    # phi(Ma) = phi0 + a Ma^2: the shift to half Ma is 3/4 of the baseline error
    phi0, a, ma = 3.1, 0.8, 0.1
    err = a * ma ** 2
    assert cr.mach_uncertainty(phi0 + err, phi0 + a * (ma / 2) ** 2) == pytest.approx(err, rel=1e-12)


def test_validation_domain_verdicts():
    g = "bluff body in channel"
    v = cr.classify(100.0, 0.1, 64, g)
    assert v.kind == "inside" and "2D-2" in v.anchor
    v = cr.classify(100.0, 0.1, 32, g)          # (40 - 32) / 40 = 0.2
    assert v.kind == "near_edge" and v.factor == pytest.approx(1.4)
    v = cr.classify(100.0, 0.1, 100, g)         # (100 - 80) / 40 = 0.5
    assert v.kind == "near_edge" and v.factor == pytest.approx(2.0)
    v = cr.classify(100.0, 0.2, 64, g)          # (0.2 - 0.14) / 0.14
    assert v.kind == "near_edge" and v.factor == pytest.approx(1.0 + 2.0 * 0.06 / 0.14)
    v = cr.classify(300.0, 0.1, 64, g)          # nearest: 2D-2, (300 - 100) / 50 = 4
    assert v.kind == "outside" and v.reasons == ["Re 300 outside validated 100–100"]
    assert cr.classify(100.0, 0.1, 64, "wing").kind == "outside"
    # the reference's adversarial gate (Cases.swift:1146-1158): Re 250,000 at
    # 48 cells per feature is outside, and its budget withholds the bar
    far = cr.classify(250_000.0, 0.09, 48, g)
    assert far.kind == "outside"
    assert math.isnan(cr.Budget("C_D", 1.0, 0.01, 0.01, 0.01, far).combined)


def test_budget_and_report():
    rng = np.random.default_rng(5)
    ladder = {n: 3.2 + 30.0 / n ** 2 + 0.01 * rng.standard_normal(4000) for n in (16, 24, 32, 48, 64)}
    half = 3.2 + 30.0 / 64 ** 2 - 0.004 + 0.01 * rng.standard_normal(4000)
    r = cr.credibility("mean C_D", ladder, half, "ab" * 32, re=100.0, finest_mach=0.1,
                       geometry="bluff body in channel", setup="test ladder", wall_seconds=12.5)
    b = r.budget
    assert b.verdict.kind == "inside"
    assert b.u_num == pytest.approx(abs(r.rungs[-1].value - r.rungs[-2].value)) or b.u_num == b.u_stat
    assert b.combined == pytest.approx(2.0 * math.sqrt(b.u_num ** 2 + b.u_stat ** 2 + b.u_ma ** 2))
    md = cr.markdown(r, ["gate one: yes"])
    for head in ("# Credibility report — mean C_D", "## Result", "## Uncertainty budget",
                 "## Resolution ladder", "## Validation domain", "## Solver verification (this solver)",
                 "## Reproducibility", "## What this report does not claim"):
        assert head in md
    assert b.headline in md and "gate one: yes" in md and "ab" * 32 in md
    assert md.count("\n| ") >= 5 + 4
    out = cr.credibility("mean C_D", ladder, half, "", re=300.0, finest_mach=0.1,
                         geometry="bluff body in channel", setup="", wall_seconds=1.0)
    assert math.isnan(out.budget.combined) and "OUTSIDE" in out.budget.headline
    assert "not published" in cr.markdown(out, [])
