"""GPU parity over seeded random configurations, bit for bit against the oracle.

The targeted parity tests pin each feature of the step on its own lattice.
These cases draw a lattice shape (odd and power-of-two extents, so the
blocked layout and the row-staged odd pass are both reached), a flag layout
(periodic, cavity, channel or every flag scattered), SRT or TRT, Smagorinsky
LES, a body force, the inflow ramp and its uniform profile, an outlet sponge,
and either Noble-Torczynski solid fractions or Brinkman porosity, all from one
seed, so feature combinations no targeted test names are exercised too. The
step follows Sources/StrouhalCore/Kernels.metal:105-355; each case compares
the raw DDF bits after odd and even step counts, then the per-cell force and
the box sum of a momentum-exchange probe (Sim.swift:359-372), then one more
step without force output.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import assert_bit_exact, make_pair, random_state, to_storage

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp32", "fp16s", "bf16"]
SEEDS = list(range(48))


def _draw(seed):
    """One configuration: constructor keywords, flags, solid fractions or
    porosity, sponge and inflow profile."""
    rng = np.random.default_rng(1000 + seed)
    nx = int(rng.choice([5, 9, 13, 17, 24, 32, 64, 128]))
    ny = int(rng.integers(3, 12))
    nz = int(rng.integers(1, 7))
    kind = str(rng.choice(["periodic", "cavity", "channel", "mixed"]))
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    if kind == "cavity":
        fl[:, 0, :] = 1
        fl[:, :, 0] = 1
        fl[:, :, nx - 1] = 1
        fl[:, ny - 1, :] = 2
    elif kind == "channel":
        fl[:, 0, :] = 1
        fl[:, ny - 1, :] = 1
        fl[:, :, 0] = 3
        fl[:, :, nx - 1] = 3
    elif kind == "mixed":
        fl = rng.choice([0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4], size=(nz, ny, nx)).astype(np.uint8)
    omega = float(rng.uniform(0.7, 1.95))
    omega_minus = float(rng.uniform(0.8, 1.9)) if rng.random() < 0.5 else None
    kw = dict(omega=omega, omega_minus=omega_minus, flags=fl, wants_forces=True,
              lid=tuple(float(v) for v in rng.uniform(-0.05, 0.05, 3)),
              uin=float(rng.uniform(0.0, 0.06)), ramp_steps=int(rng.choice([0, 0, 3, 7])))
    if rng.random() < 0.5:
        kw["force"] = tuple(float(v) for v in rng.uniform(-2e-5, 2e-5, 3))
    medium = str(rng.choice(["none", "partial", "porous"]))
    if medium != "porous" and rng.random() < 0.5:
        kw["c_smago"] = float(rng.uniform(0.02, 0.17))
    field = None
    if medium == "partial":
        field = np.where(rng.random((nz, ny, nx)) < 0.3, rng.random((nz, ny, nx)), 0.0)
        field = field.astype(np.float32)
        field.reshape(-1)[int(rng.integers(0, field.size))] = 1.0
    elif medium == "porous":
        field = np.where(rng.random((nz, ny, nx)) < 0.4, rng.uniform(0.0, 0.3, (nz, ny, nx)), 0.0)
        field = field.astype(np.float32)
    sponge = None
    if rng.random() < 0.4:
        x0 = float(rng.uniform(0.3, 0.8) * nx)
        sponge = (x0, float(rng.uniform(1.0, nx - x0 + 1.0)), float(rng.uniform(0.5, 1.5)))
    uniform = bool(rng.random() < 0.5)
    return (nx, ny, nz), kind, kw, medium, field, sponge, uniform


def _describe(seed):
    shape, kind, kw, medium, field, sponge, uniform = _draw(seed)
    return (f"seed {seed}: {shape} {kind} omega {kw['omega']:.3f} omega_minus {kw['omega_minus']} "
            f"les {kw.get('c_smago', 0.0):.3f} force {kw.get('force')} {medium} sponge {sponge} "
            f"uniform {uniform} ramp {kw['ramp_steps']}")


def _make(gpu, precision, seed, medium_ok=("none", "partial")):
    """The pair (GPU lattice, oracle) of a drawn configuration with its seeded
    state uploaded; a drawn porosity is left out unless medium_ok names it
    (batches run lattices without porosity)."""
    (nx, ny, nz), kind, kw, medium, field, sponge, uniform = _draw(seed)
    om, omm = kw.pop("omega"), kw.pop("omega_minus")
    g, r = make_pair(gpu, precision, nx, ny, nz, om, omm, **kw)
    for s in (g, r):
        if medium == "partial":
            s.set_solid_fractions(field)
        elif medium == "porous" and "porous" in medium_ok:
            s.set_porosity(field)
        s.inflow_uniform = uniform
    g.sponge = sponge
    r.sponge = g.sponge
    img = to_storage(random_state(g.cells, 31 * seed + 5), precision)
    g.upload_ddf(img)
    r.upload_ddf(img if precision != "fp16s" else img.view(np.uint16))
    return g, r, (nx, ny)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("seed", SEEDS)
def test_random_configuration(gpu, precision, seed):
    print(_describe(seed), precision, flush=True)
    g, r, (nx, ny) = _make(gpu, precision, seed, medium_ok=("none", "partial", "porous"))
    for steps in (1, 2, 3):
        g.run(steps)
        r.run(steps)
        assert g.steps_done == r.steps_done
        assert_bit_exact(g.download_ddf(), r.download_ddf(), f"DDFs after {g.steps_done} steps")
    box = ((1, max(1, nx - 2)), (1, max(1, ny - 2)))
    fg = g.probe_force(*box)
    fr = r.probe_force(*box)
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    field_r = r.force_field()
    assert_bit_exact(g.download_force(), field_r.astype(np.float32), "per-cell force")
    scale = np.abs(field_r[:, :3]).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)
    g.run(1)
    r.run(1)
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after the step past the probe")


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", range(8))
def test_random_batch(gpu, precision, case):
    """Three drawn lattices advanced together by one batch (its own kernels),
    each with its own step count and probe schedule, against the oracle run
    lattice by lattice: DDF bits and the FP64 force samples
    (Ladder.swift:54-121 advances the rungs in turn)."""
    from strouhal_amd.batch import Batch
    rng = np.random.default_rng(500 + case)
    seeds = [int(s) for s in rng.choice(64, size=3, replace=False)]
    print(f"batch case {case}: seeds {seeds} {precision}", flush=True)
    pairs = [_make(gpu, precision, s) for s in seeds]
    b = Batch([g for g, _, _ in pairs])
    b.set_block_size(int(rng.choice([128, 256, 512])))
    b.set_graph_steps(int(rng.choice([1, 3, 16])))
    plans = []
    for i, (g, r, (nx, ny)) in enumerate(pairs):
        transient = 2 * int(rng.integers(0, 3))
        chunk = 2 * int(rng.integers(0, 3))
        count = int(rng.integers(1, 4))
        tail = int(rng.integers(0, 4))
        box = ((1, max(1, nx - 2)), (1, max(1, ny - 2)))
        b.set_probe(i, transient + chunk + 1, chunk + 2, count, *box)
        plans.append((transient, chunk, count, tail, box))
    b.run([t + n * (c + 2) + tail for t, c, n, tail, _ in plans])
    for i, ((g, r, _), (transient, chunk, count, tail, box)) in enumerate(zip(pairs, plans)):
        r.run(transient)
        ref, scale = [], []
        for _ in range(count):
            r.run(chunk)
            ref.append(r.probe_force(*box))
            scale.append(np.abs(r.force_field()[:, :3]).sum() + 1e-300)
        r.run(tail)
        assert g.steps_done == r.steps_done, (i, g.steps_done, r.steps_done)
        assert_bit_exact(g.download_ddf(), r.download_ddf(), f"lattice {i} (seed {seeds[i]}) DDFs")
        got = b.probe_series(i)
        ref = np.array(ref)
        assert got.shape == ref.shape == (count, 3)
        assert np.all(np.abs(got - ref) <= 1e-13 * np.array(scale)[:, None]), (i, got, ref)
    b.close()
