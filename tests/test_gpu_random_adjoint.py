"""GPU adjoint over seeded random configurations, against the FP64 oracle.

Each case draws a lattice (shape, every wall kind, SRT or TRT, body force,
moving lid, ramped inflow, outlet sponge), one medium (Brinkman porosity,
Noble-Torczynski partial cells with or without Smagorinsky LES, or plain LES)
and an objective (pressure columns, window, weights), then differentiates it
with strh_vjp from a random state, a random checkpoint length, and a
two-level plan. Checked for every case:

* the recording forward leaves the lattice on the plain run's digest;
* J, dJ/du_in and dJ/dalpha or dJ/deps against the FP64 oracle's discrete
  adjoint: J within 1e-6 of the scale of its terms, the gradients within
  1e-4 of their largest component (FP32 forward, FP32 adjoint state, FP64
  accumulation);
* the two-level plan (strh_vjp_nested) gives the one-level plan's bits.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import OracleSim
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

SEEDS = list(range(32))


def draw(seed):
    """Lattice, medium and objective of one case."""
    rng = np.random.default_rng(7000 + seed)
    nx = int(rng.integers(8, 25))
    ny = int(rng.integers(5, 13))
    nz = int(rng.choice([1, 1, 2, 4]))
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = 1
    fl[:, ny - 1, :] = int(rng.choice([1, 2]))
    fl[:, :, 0] = 3
    fl[:, :, nx - 1] = int(rng.choice([3, 4]))
    if rng.random() < 0.5:                                  # scattered cells of every kind
        inner = rng.choice([0] * 14 + [1, 2, 3, 4], size=(nz, ny - 2, nx - 2)).astype(np.uint8)
        fl[:, 1:ny - 1, 1:nx - 1] = inner
    tau = float(rng.uniform(0.55, 0.9))
    wp = 1.0 / tau
    wm = wp if rng.random() < 0.5 else 1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5))
    kw = dict(uin=float(rng.uniform(0.02, 0.06)), ramp_steps=int(rng.choice([0, 3, 6])),
              lid=tuple(float(v) for v in rng.uniform(-0.04, 0.04, 3)),
              force=tuple(float(v) for v in rng.uniform(-1e-5, 1e-5, 3)) if rng.random() < 0.5
              else (0.0, 0.0, 0.0))
    medium = str(rng.choice(["porous", "partial", "partial_les", "les"]))
    field = None
    if medium == "porous":
        field = np.where(rng.random((nz, ny, nx)) < 0.4, rng.uniform(0.0, 0.2, (nz, ny, nx)), 0.0)
    elif medium.startswith("partial"):
        field = np.where(rng.random((nz, ny, nx)) < 0.3, rng.random((nz, ny, nx)), 0.0)
    if field is not None:
        field = field.astype(np.float32)
    if medium.endswith("les"):
        kw["c_smago"] = float(rng.uniform(0.1, 0.3))
    sponge = None
    if rng.random() < 0.5:
        x0 = float(rng.uniform(0.4, 0.8) * nx)
        sponge = (x0, float(rng.uniform(1.0, nx - x0 + 1.0)), float(rng.uniform(0.6, 1.2)))
    obj = dict(inlet_x=int(rng.integers(-1, nx // 2)), outlet_x=int(rng.integers(nx // 2, nx)),
               spinup=int(rng.integers(0, 6)), avg=int(rng.integers(1, 10)),
               w_p=float(rng.uniform(0.5, 2.0)),
               w_d=float(rng.uniform(0.5, 2.0)) if medium == "porous" else 0.0,
               w_f=float(rng.uniform(0.5, 3.0)) if medium.startswith("partial") else 0.0)
    every = int(rng.choice([0, -1, 1, 2, 3, 5]))
    outer, inner = int(rng.integers(2, 7)), int(rng.integers(1, 4))
    return (nx, ny, nz), fl, wp, wm, kw, medium, field, sponge, obj, every, (outer, inner)


def make(gpu, seed, with_oracle=True):
    from strouhal_amd import Simulation
    (nx, ny, nz), fl, wp, wm, kw, medium, field, sponge, obj, every, plan = draw(seed)
    g = Simulation(gpu, "fp32", nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, flags=fl, **kw)
    if medium == "porous":
        g.set_porosity(field)
    elif field is not None:
        g.set_solid_fractions(field)
    g.sponge = sponge
    img = random_state(g.cells, 13 * seed + 1).astype(np.float32)
    g.upload_ddf(img)
    r = None
    if with_oracle:
        r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f64", storage="f64", lid=g.lid_vel,
                      uin=g.uin_target, ramp_steps=g.ramp_steps, force=g.force, c_smago=g.c_smago,
                      flags=fl, threads=8)
        r.sponge = g.sponge
        if medium == "porous":
            r.set_porosity(field)
        elif field is not None:
            r.set_solid_fractions(field)
        r.upload_ddf(img.astype(np.float64))
    return g, r


def _within(a, b, rel):
    a, b = np.asarray(a, dtype=np.float64).reshape(-1), np.asarray(b, dtype=np.float64).reshape(-1)
    scale = np.max(np.abs(b))
    err = np.max(np.abs(a - b)) / scale if scale > 0 else np.max(np.abs(a))
    return err <= rel, err


@pytest.mark.parametrize("seed", SEEDS)
def test_random_vjp_matches_fp64_oracle(gpu, seed):
    (nx, ny, nz), fl, wp, wm, kw, medium, field, sponge, obj, every, (outer, inner) = draw(seed)
    print(f"seed {seed}: {(nx, ny, nz)} {medium} omega {wp:.3f}/{wm:.3f} {kw} sponge {sponge} "
          f"obj {obj} every {every} plan {outer}/{inner}", flush=True)
    porous, eps = medium == "porous", field is not None and medium != "porous"
    g, r = make(gpu, seed)
    J, ga, gu, info = g.vjp(**obj, checkpoint_every=every, want_alpha=porous, want_eps=eps)
    Jr, gar, ger, gur = r.vjp(**obj, with_eps=True)
    # J's scale: its own size, or the pressure weight where the two column
    # pressures (about 1/3 each) nearly cancel
    jscale = max(abs(Jr), obj["w_p"])
    assert abs(J - Jr) <= 1e-6 * jscale, (J, Jr)
    assert gu == pytest.approx(gur, rel=1e-4, abs=1e-9), (gu, gur)
    ok, err = True, 0.0
    if porous:
        ok, err = _within(ga, gar, 1e-4)
        assert ok, ("dJ/dalpha", err)
    if eps:
        ge = info["grad_eps"]
        ok, err = _within(ge, ger, 1e-4)
        assert ok, ("dJ/deps", err)
        assert np.all(ge.reshape(-1)[field.reshape(-1) == 0] == 0.0)
    print(f"  J {abs(J - Jr) / jscale:.2e}  du_in {abs(gu - gur) / max(abs(gur), 1e-300):.2e}  "
          f"field gradient {err:.2e}", flush=True)
    # the recording forward ends on the plain run's digest
    p, _ = make(gpu, seed, with_oracle=False)
    p.run(obj["spinup"] + obj["avg"])
    assert g.state_digest == p.state_digest
    # two checkpoint levels: the one-level plan's bits
    t, _ = make(gpu, seed, with_oracle=False)
    J2, ga2, gu2, info2 = t.vjp(**obj, checkpoint_every=outer, inner_every=inner, want_alpha=porous,
                                want_eps=eps)
    assert J2 == J and gu2 == gu
    if porous:
        assert np.array_equal(ga2, ga)
    if eps:
        assert np.array_equal(info2["grad_eps"], info["grad_eps"])
    assert t.state_digest == p.state_digest
