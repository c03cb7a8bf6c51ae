"""The credibility run around the resolution ladder: the uncertainty budget of
a time-averaged quantity of interest and its markdown report.

Host-side statistics over the force series the batched ladder produces, in
the reference's definitions:

* ``stationary_tail``   QoISeries.stationaryTail (Sources/StrouhalCore/Uncertainty.swift:37-43)
* ``StatUncertainty``   non-overlapping batch means with a lag-1 check (Uncertainty.swift:51-103)
* ``NumUncertainty``    finest-pair spread, observed order as a diagnostic (Uncertainty.swift:123-165)
* ``mach_uncertainty``  half-Mach anchor, O(Ma^2) residual (Uncertainty.swift:170-182)
* ``classify``          validation-domain verdict against the gate anchors (Uncertainty.swift:188-260)
* ``Budget``            U = k sqrt(u_num^2 + u_stat^2 + u_Ma^2), k = 2 (Uncertainty.swift:263-297)
* ``credibility``       Ladder.credibility (Sources/StrouhalCore/Ladder.swift:74-121)
* ``markdown``          CredibilityReport.markdown (Ladder.swift:126-202)
"""
from __future__ import annotations

import datetime
import math
from dataclasses import dataclass, field

import numpy as np


def stationary_tail(values, fraction: float = 0.5) -> np.ndarray:
    """The record without its first `fraction` (the start-up transient)."""
    v = np.asarray(values, dtype=np.float64)
    return v[int(len(v) * fraction):]


@dataclass
class StatUncertainty:
    """95% half-width of the mean of a stationary series by non-overlapping
    batch means. From max_batches down to min_batches (step 2), the first
    batch count whose batch means have |lag-1 autocorrelation| <= 0.3 is
    taken: the tightest interval whose batches are long enough to be
    independent; if none qualifies, min_batches, marked untrustworthy."""
    mean: float
    half_width95: float
    batches: int
    lag1: float

    @property
    def trustworthy(self) -> bool:
        return self.batches >= 6 and abs(self.lag1) <= 0.3

    @classmethod
    def of(cls, values, max_batches: int = 16, min_batches: int = 6) -> "StatUncertainty":
        v = np.asarray(values, dtype=np.float64)
        if v.size < min_batches * 4:
            return cls(float(v.mean()) if v.size else math.nan, math.nan, 0, math.nan)

        def stats(M):
            per = v.size // M
            means = np.array([v[b * per:(b + 1) * per].mean() for b in range(M)])
            m = means.mean()
            d = means - m
            sd = math.sqrt(float((d * d).sum()) / (M - 1))
            den = float((d * d).sum())
            num = float((d[1:] * d[:-1]).sum())
            return float(m), 1.96 * sd / math.sqrt(M), (num / den if den > 0 else math.nan)

        for M in range(max_batches, min_batches - 1, -2):
            m, half, rho = stats(M)
            if abs(rho) <= 0.3:
                return cls(m, half, M, rho)
        m, half, rho = stats(min_batches)
        return cls(m, half, min_batches, rho)


@dataclass
class Rung:
    cells_per_feature: int
    value: float
    stat: float          # u_stat of the rung (95% half-width)


@dataclass
class NumUncertainty:
    """Discretisation uncertainty from the ladder: |finest - next|, never below
    the finest rung's statistical half-width. With three rungs converging
    monotonically (d1 d2 > 0, |d2| < |d1|) the observed order
    ln|d1/d2| / ln(r) is reported as a diagnostic only."""
    rungs: list
    u_num: float
    observed_order: float | None
    note: str

    @classmethod
    def of(cls, rungs) -> "NumUncertainty":
        rs = sorted(rungs, key=lambda r: r.cells_per_feature)
        if len(rs) < 2:
            return cls(rs, math.nan, None, "single resolution — u_num not estimated (run a ladder)")
        fine, mid = rs[-1], rs[-2]
        noise = 0.0 if math.isnan(fine.stat) else fine.stat
        delta = abs(fine.value - mid.value)
        u = max(delta, noise)
        order = None
        if len(rs) >= 3:
            c = rs[-3]
            d1, d2 = mid.value - c.value, fine.value - mid.value
            if d1 * d2 > 0 and abs(d2) < abs(d1) and abs(d2) > 0:
                r = fine.cells_per_feature / mid.cells_per_feature
                if r > 1.05:
                    order = math.log(abs(d1 / d2)) / math.log(r)
        if delta <= noise:
            note = "finest two rungs agree within statistical noise — resolution-insensitive over the tested range"
        else:
            note = "u_num = |finest − next| (no Richardson/GCI: invalid for scale-resolving runs)"
        return cls(rs, u, order, note)


def mach_uncertainty(baseline: float, half_mach: float) -> float:
    """The compressibility error is O(Ma^2): halving Ma removes 3/4 of it, so
    the residual at the baseline is 4/3 of the measured shift."""
    return abs(baseline - half_mach) * 4.0 / 3.0


@dataclass(frozen=True)
class Anchor:
    name: str
    re: tuple
    mach: tuple
    cells_per_feature: tuple
    geometry: str


# The gates the solver is verified on (the reference's anchors, Uncertainty.swift:205-218).
ANCHORS = (
    Anchor("Ghia lid-driven cavity (RMS 0.004 vs published)", (100.0, 1000.0), (0.0, 0.18), (128, 256),
           "enclosed"),
    Anchor("Schäfer–Turek DFG 2D-1 (C_D within 0.16%)", (20.0, 20.0), (0.0, 0.14), (40, 64),
           "bluff body in channel"),
    Anchor("Schäfer–Turek DFG 2D-2 (St 0.2996, max C_L 0.9998)", (100.0, 100.0), (0.0, 0.14), (40, 80),
           "bluff body in channel"),
    Anchor("Poiseuille (exact) / Taylor–Green (order 1.96)", (0.0, 100.0), (0.0, 0.35), (32, 128),
           "periodic / channel"),
)


def _distance(x: float, lo: float, hi: float) -> float:
    """Fractional distance of x outside [lo, hi] (0 inside)."""
    if lo <= x <= hi:
        return 0.0
    span = max(hi - lo, abs(hi) * 0.5, 1e-9)
    return (lo - x) / span if x < lo else (x - hi) / span


@dataclass
class Verdict:
    kind: str                    # "inside", "near_edge", "outside"
    anchor: str = ""
    factor: float = 1.0          # inflation of the bar near the edge
    reasons: list = field(default_factory=list)


def classify(re: float, mach: float, cells_per_feature: int, geometry: str) -> Verdict:
    """The nearest anchor of the same geometry class by the largest fractional
    distance over (Re, Ma, cells per feature): inside at 0, near the edge up to
    0.5 (bar inflated by 1 + 2d), outside beyond."""
    best = None
    for a in ANCHORS:
        if a.geometry != geometry:
            continue
        d = max(_distance(re, *a.re), _distance(mach, *a.mach),
                _distance(float(cells_per_feature), float(a.cells_per_feature[0]),
                          float(a.cells_per_feature[1])))
        if best is None or d < best[1]:
            best = (a, d)
    if best is None:
        return Verdict("outside", reasons=[f"no validated case of geometry class “{geometry}”"])
    a, d = best
    if d == 0.0:
        return Verdict("inside", a.name)
    if d <= 0.5:
        return Verdict("near_edge", a.name, 1.0 + 2.0 * d)
    reasons = []
    if _distance(re, *a.re) > 0:
        reasons.append(f"Re {re:.0f} outside validated {a.re[0]:.0f}–{a.re[1]:.0f}")
    if _distance(mach, *a.mach) > 0:
        reasons.append(f"Ma {mach:.3f} outside validated ≤{a.mach[1]:.2f}")
    if _distance(float(cells_per_feature), float(a.cells_per_feature[0]), float(a.cells_per_feature[1])) > 0:
        reasons.append(f"{cells_per_feature} cells/feature outside validated "
                       f"{a.cells_per_feature[0]}–{a.cells_per_feature[1]}")
    return Verdict("outside", a.name, reasons=reasons or ["outside the validated domain"])


@dataclass
class Budget:
    qoi: str
    value: float
    u_num: float
    u_stat: float
    u_ma: float
    verdict: Verdict
    notes: list = field(default_factory=list)

    @property
    def combined(self) -> float:
        """k = 2 root-sum-square of the estimated parts; NaN outside the domain."""
        parts = [u for u in (self.u_num, self.u_stat, self.u_ma) if not math.isnan(u)]
        rss = math.sqrt(sum(u * u for u in parts))
        if self.verdict.kind == "inside":
            return 2.0 * rss
        if self.verdict.kind == "near_edge":
            return 2.0 * rss * self.verdict.factor
        return math.nan

    @property
    def headline(self) -> str:
        if self.verdict.kind == "outside":
            return f"{self.qoi} = {self.value:.4f} — OUTSIDE the validated domain: no calibrated uncertainty"
        return f"{self.qoi} = {self.value:.4f} ± {self.combined:.4f} (k=2)"


@dataclass
class CredibilityRun:
    qoi: str
    budget: Budget
    rungs: list
    num_note: str
    observed_order: float | None
    stat_batches: int
    stat_lag1: float
    mach_baseline: float
    mach_half: float
    wall_seconds: float
    digest: str
    setup: str


def credibility(qoi: str, ladder: dict, half_mach_series, finest_digest: str, re: float,
                finest_mach: float, geometry: str, setup: str, wall_seconds: float) -> CredibilityRun:
    """Ladder.credibility's assembly. `ladder` maps cells per feature to the
    sampled QoI series of that rung at the baseline Mach number;
    `half_mach_series` is the finest rung's series at half the lattice
    velocity. Every series keeps its stationary tail (the second half)."""
    rungs, finest_stat = [], None
    finest = max(ladder)
    for n in sorted(ladder):
        tail = stationary_tail(ladder[n])
        st = StatUncertainty.of(tail)
        rungs.append(Rung(n, float(tail.mean()), st.half_width95))
        if n == finest:
            finest_stat = st
    half = float(stationary_tail(half_mach_series).mean())
    num = NumUncertainty.of(rungs)
    u_ma = mach_uncertainty(rungs[-1].value, half)
    verdict = classify(re, finest_mach, finest, geometry)
    notes = [num.note]
    if not finest_stat.trustworthy:
        notes.append(f"statistics weak: {finest_stat.batches} batches, lag-1 ρ = {finest_stat.lag1:.2f} — "
                     "average over more shedding cycles")
    budget = Budget(qoi, rungs[-1].value, num.u_num, finest_stat.half_width95, u_ma, verdict, notes)
    return CredibilityRun(qoi, budget, rungs, num.note, num.observed_order, finest_stat.batches,
                          finest_stat.lag1, rungs[-1].value, half, wall_seconds, finest_digest, setup)


def markdown(r: CredibilityRun, build_gates, now: datetime.datetime | None = None) -> str:
    """The credibility report in ASME V&V 20 vocabulary, with the parts that
    are not estimated named as such (the sections of Ladder.swift:126-202)."""
    b = r.budget
    stamp = (now or datetime.datetime.now(datetime.timezone.utc)).strftime("%Y-%m-%dT%H:%M:%SZ")
    out = [f"# Credibility report — {r.qoi}\n", f"_Generated by Strouhal · {stamp}_\n",
           "## Result\n", f"**{b.headline}**\n", f"{r.setup}\n",
           "## Uncertainty budget\n", "U(φ) = k·√(u_num² + u_stat² + u_Ma²), k = 2 (≈95%)\n",
           "| Component | Value | Basis |", "|---|---|---|",
           f"| u_num (discretization) | {b.u_num:.4f} | {r.num_note} |",
           f"| u_stat (time average) | {b.u_stat:.4f} | non-overlapping batch means, {r.stat_batches} batches, "
           f"lag-1 ρ = {r.stat_lag1:.2f} |",
           f"| u_Ma (compressibility) | {b.u_ma:.4f} | half-Mach anchor: {r.mach_baseline:.4f} → "
           f"{r.mach_half:.4f}, O(Ma²) extrapolation |"]
    v = b.verdict
    if v.kind == "inside":
        out.append(f"| **U(φ) (k=2)** | **{b.combined:.4f}** | inside the validated domain ({v.anchor}) |")
    elif v.kind == "near_edge":
        out.append(f"| **U(φ) (k=2)** | **{b.combined:.4f}** | near the edge of the validated domain "
                   f"({v.anchor}); inflated ×{v.factor:.2f} |")
    else:
        out.append("| **U(φ)** | **not published** | outside the validated domain — see below |")
    out += ["", "## Resolution ladder\n", "| cells/feature | φ | u_stat |", "|---|---|---|"]
    out += [f"| {g.cells_per_feature} | {g.value:.4f} | {g.stat:.4f} |" for g in r.rungs]
    if r.observed_order is not None:
        out.append(f"\nObserved order of convergence: **{r.observed_order:.2f}** (diagnostic only). The "
                   "error bar does not use it: a Richardson extrapolation needs monotone convergence at a "
                   "fitted order, and a scale-resolving run couples discretization and subgrid error, so "
                   "that is not guaranteed. No GCI is given.")
    else:
        out.append("\nNo observed order: the ladder does not converge monotonically, so no order can be "
                   "fitted.")
    out += ["", "## Validation domain\n"]
    if v.kind == "inside":
        out.append(f"This run lies **inside** the domain validated by: {v.anchor}.")
    elif v.kind == "near_edge":
        out.append(f"This run lies **near the edge** of the domain validated by: {v.anchor}. The bar grows "
                   f"by ×{v.factor:.2f} with the extrapolation distance (Oberkampf & Roy).")
    else:
        out.append("**This run lies OUTSIDE the validated domain.** No calibrated uncertainty is given: "
                   "none could be defended.\n")
        out += [f"- {x}" for x in v.reasons]
        out.append("\nThe values above are what the solver computed, not measurements with a stated "
                   "accuracy.")
    out.append("")
    notes = [n for n in b.notes if n]
    if notes:
        out += ["## Notes\n"] + [f"- {n}" for n in notes] + [""]
    out += ["## Solver verification (this solver)\n"] + [f"- {g}" for g in build_gates] + [""]
    out += ["## Reproducibility\n", f"- Final state digest (SHA-256): `{r.digest}`",
            f"- Total wall time for the full credibility run: {r.wall_seconds:.1f} s",
            "- The same build, settings and machine give the same state bit for bit and the same QoIs.\n",
            "## What this report does not claim\n",
            "- **No model-form uncertainty.** The error of the collision and subgrid closure itself is not "
            "bounded; that takes several models (Klein 2005), which this run does not vary.",
            "- **No Grid Convergence Index**, for the reason given with the ladder above.",
            "- A small numerical uncertainty does **not** make the answer right: in the FDA nozzle "
            "round-robin u_num stayed under 1% while the model missed the experiment by about a third. "
            "Only validation against data settles that, and the validation-domain section states whether "
            "it applies here."]
    return "\n".join(out) + "\n"
