"""GPU: multi-lattice batches evolve bit for bit as the lattices do alone.

The reference advances ladder rungs one after another (Ladder.swift:54-121);
the batch advances them together. Every lattice's DDFs and its FP64 force
series must be the same bits either way, for any workgroup size, graph
length and mix of lattice kinds.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from strouhal_amd.batch import Batch
from tests.helpers import assert_bit_exact
from strouhal_amd.ladder import (build, cycle_stats, coefficients, probe_schedule, run_batched,
                                 run_sequential)

pytestmark = pytest.mark.gpu

BASE_UMAX = 1.5 * 0.1 / math.sqrt(3.0)


@pytest.mark.parametrize("precision", ["fp32", "fp16s"])
def test_ladder_batched_equals_sequential(gpu, precision):
    vb = build(gpu, precision, [8, 12], 0.01, BASE_UMAX)
    vs = build(gpu, precision, [8, 12], 0.01, BASE_UMAX)
    sb, _, bs, _ = run_batched(vb, autotune=True, graph_steps=16)
    ss, _ = run_sequential(vs)
    for a, b, v in zip(sb, ss, vb):
        assert a.shape == b.shape == (v.probes, 3)
        assert np.array_equal(a, b, equal_nan=True), (v.D, v.mach_scale)
    for a, b in zip(vb, vs):
        assert a.case.sim.steps_done == b.case.sim.steps_done
        assert a.case.sim.state_digest == b.case.sim.state_digest
    # the credibility report of the short ladder (Ladder.swift:74-202)
    from strouhal_amd.ladder import credibility_report
    rungs = []
    for v, s in zip(vb, sb):
        cd, cl, t = coefficients(v, s)
        rungs.append({"D": v.D, "mach_scale": v.mach_scale, "finite": bool(np.all(np.isfinite(s))),
                      **cycle_stats(cl, cd, t, v.D, v.case.u_mean)})
    md = credibility_report(vb, sb, rungs, {"precision": precision, "bitwise_identical": True}, 1.0,
                            BASE_UMAX)
    assert "# Credibility report — mean C_D" in md and "## Resolution ladder" in md
    assert vb[-2].case.sim.state_digest in md          # the finest Mach-1 rung's digest


@pytest.mark.parametrize("precision,resident,graph", [("fp32", 1, 64), ("fp16s", 1, 256), ("fp32", 2, 256),
                                                      ("fp16s", 2, 64), ("fp32", 3, 64), ("fp16s", 3, 64)])
def test_full_size_ladder_batched_equals_sequential(gpu, precision, resident, graph):
    # The ladder at its full size (D in {16, 24, 32, 48, 64}, each at Ma 0.1
    # and half Mach: 10 lattices, 1,504,680 cells) on 1% of the reference
    # schedule: the batch ends every lattice on the sequential run's bits,
    # with the same FP64 force samples, and no rung goes non-finite (FP16
    # rungs collide with SRT).
    res = [16, 24, 32, 48, 64]
    vb = build(gpu, precision, res, 0.01, BASE_UMAX)
    vs = build(gpu, precision, res, 0.01, BASE_UMAX)
    assert len(vb) == 10 and sum(v.case.sim.cells for v in vb) == 1_504_680
    # resident 1: LDS-resident launches for the segments whose rungs fit on
    # chip together (the tail); 2: for every segment, in time-sliced groups;
    # 3: as 1, and the segments that need groups switch to the faster kind
    # after timing both (segments of 8 graphs and more). 256-step graphs hold
    # resident launches of 256 steps, with the force samples of each reduced
    # after it.
    sb, _, _, _ = run_batched(vb, autotune=True, graph_steps=graph, resident=resident)
    ss, _ = run_sequential(vs)
    for a, b, v in zip(sb, ss, vb):
        assert a.shape == b.shape == (v.probes, 3)
        assert np.all(np.isfinite(a)), (v.D, v.mach_scale)
        assert np.array_equal(a, b), (v.D, v.mach_scale)
    for a, b in zip(vb, vs):
        assert a.case.sim.steps_done == b.case.sim.steps_done == a.total_steps
        assert a.case.sim.state_digest == b.case.sim.state_digest, (a.D, a.mach_scale)


def _mixed(gpu, precision):
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.cases import cavity
    a = Simulation(gpu, precision, nx=64, ny=64, nz=4, omega=1.8)      # wall-free, row-staged alone
    a.init_field(2, 0.05)
    b = cavity(gpu, precision, n=40, re=1e4, lam=0.25, ramp_steps=30)
    b.c_smago = 0.04                                                    # LES
    wp, wm = trt_omegas(0.7, 3.0 / 16.0)
    c = Simulation(gpu, precision, nx=33, ny=18, nz=3, omega=wp, omega_minus=wm,
                   force=(1e-5, 0.0, 0.0), flags=lambda x, y, z: np.where((y == 0) | (y == 17), 1, 0))
    return [a, b, c]


@pytest.mark.parametrize("bs,graph", [(128, 0), (256, 8), (512, 64)])
def test_mixed_batch_matches_solo_runs(gpu, bs, graph):
    steps = [37, 50, 21]
    batched = _mixed(gpu, "fp32")
    solo = _mixed(gpu, "fp32")
    bt = Batch(batched)
    bt.set_block_size(bs)
    bt.set_graph_steps(graph)
    bt.run(steps)            # every lattice starts even: one pass per launch
    bt.run([3, 0, 8])        # the two active lattices start odd
    bt.run([4, 0, 19])       # 40 (even) with 29 (odd): launches branch per lattice
    for s, k in zip(solo, steps):
        s.run(k)
    solo[0].run(3)
    solo[2].run(8)
    solo[0].run(4)
    solo[2].run(19)
    for x, y in zip(batched, solo):
        assert x.steps_done == y.steps_done
        assert x.state_digest == y.state_digest


def test_probe_rings_too_large_for_the_long_span(gpu):
    """A probed lattice whose ring of force buffers for 256 steps would pass
    1 GiB (600,000 cells, a sample every second step) takes 64-step spans: its
    resident launches and probe reductions then cover 64 steps inside the
    256-step graphs. Resident and streaming batches give the same samples, and
    the state is the solo run's."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import channel_flags

    def make():
        nx, ny = 1000, 600
        s = Simulation(gpu, "fp16s", nx=nx, ny=ny, nz=1, omega=1.7, uin=0.05, ramp_steps=20,
                       flags=channel_flags(nx, ny), wants_forces=True)
        s.init_field(0, 0.0)
        return s

    out = []
    for mode in (2, 0):
        sim = make()
        bt = Batch([sim])
        bt.set_resident(mode)
        bt.set_graph_steps(256)
        bt.set_probe(0, 1, 2, 300, (1, 40), (1, 598))     # odd steps 1, 3, ..., 599
        bt.run([600])
        out.append((bt.probe_series(0), sim.state_digest))
        bt.close()
        sim.close()
    solo = make()
    solo.run(600)
    (s2, d2), (s0, d0) = out
    assert s2.shape == s0.shape == (300, 3)
    assert np.all(np.isfinite(s2)) and s2[-1, 0] != 0.0
    assert np.array_equal(s2, s0)
    assert d2 == d0 == solo.state_digest


def test_autotune_keeps_state(gpu):
    sims = _mixed(gpu, "fp16s")
    bt = Batch(sims)
    before = [s.state_digest for s in sims]
    bs = bt.autotune()
    assert bs in (128, 256, 512)
    assert [s.state_digest for s in sims] == before
    assert all(s.steps_done == 0 for s in sims)


def _planar_specs():
    """Lattices one cell deep: a periodic wall-free lattice, a lid-driven
    cavity (moving wall, ramp) and two vortex-street ladder rungs, D = 8 and
    12 (NT cylinder, ramped parabolic inflow, outflow walls, sponge, force
    probes). Each spec builds a GPU lattice and its FP32-oracle twin."""
    from strouhal_amd.cases import cavity_flags, channel_flags, disk_solid_fractions
    specs = [dict(nx=100, ny=40, omega=1.7, init=(1, 0.05)),
             dict(nx=68, ny=68, omega=1.6, omega_minus=1.3, lid=(0.08, 0.0, 0.0), ramp=30,
                  flags=cavity_flags(66))]
    for D in (8, 12):
        nx, ny = 22 * D + 2, int(4.1 * D) + 2
        c = 0.5 + 2.0 * D
        specs.append(dict(nx=nx, ny=ny, omega=1.85, omega_minus=1.4, uin=0.1, ramp=40,
                          flags=channel_flags(nx, ny), eps=disk_solid_fractions(nx, ny, c, c, D / 2.0),
                          sponge=(float(nx - 1 - 3 * D), float(3 * D), 1.0), forces=True,
                          box=((int(c) - D // 2 - 3, int(c) + D // 2 + 3),
                               (int(c) - D // 2 - 3, int(c) + D // 2 + 3))))
    return specs


def _make(gpu, precision, spec, oracle=False):
    from oracle.oracle import OracleSim
    from strouhal_amd import Simulation
    from tests.helpers import STORAGE
    g = Simulation(gpu, precision, nx=spec["nx"], ny=spec["ny"], nz=1, omega=spec["omega"],
                   omega_minus=spec.get("omega_minus"), lid=spec.get("lid", (0.0, 0.0, 0.0)),
                   uin=spec.get("uin", 0.0), ramp_steps=spec.get("ramp", 0),
                   wants_forces=spec.get("forces", False), flags=spec.get("flags"))
    if "eps" in spec:
        g.set_solid_fractions(spec["eps"])
    if "sponge" in spec:
        g.sponge = spec["sponge"]
    if "init" in spec:
        g.init_field(*spec["init"])
    if not oracle:
        return g
    r = OracleSim(g.nx, g.ny, 1, g.omega, g.omega_minus, arith="f32", storage=STORAGE[precision],
                  lid=g.lid_vel, uin=g.uin_target, ramp_steps=g.ramp_steps, wants_forces=g.wants_forces,
                  flags=g.flags, threads=4)
    if "eps" in spec:
        r.set_solid_fractions(spec["eps"])
    r.sponge = g.sponge
    img = g.download_ddf()          # the GPU's initial field (its sin / cos)
    r.upload_ddf(img if precision != "fp16s" else img.view(np.uint16))
    return g, r


def _planar(gpu, precision):
    specs = _planar_specs()
    sims = [_make(gpu, precision, sp) for sp in specs]
    return sims[:3], specs[2]["box"]


# Batch runs of the oracle comparison: lock-step chunks (steps per lattice).
# The first run starts every lattice at step 0 (segments where lattices finish);
# the second starts the streets alone at odd steps; the third mixes parities
# (a lattice at an even step beside one at an odd step).
_RUNS = [[36, 50, 44, 30], [3, 0, 17, 9], [22, 9, 0, 14], [40, 24, 16, 31]]
# Probe schedules (first odd step, even period, count) of the two streets.
_PROBES = {2: (7, 6, 12), 3: (5, 2, 30)}


def _oracle_series(r, box, first, period, count, total):
    """The oracle run of `total` steps with Ladder.measure's probes: a force
    sample on odd step t is probe_force from step t - 1 (zero the field, two
    steps, force on the second, FP64 box sum)."""
    out = []
    for j in range(count):
        t = first + j * period
        if t + 1 > total:
            break
        r.run(t - 1 - r.steps_done)
        out.append(r.probe_force(*box))
    r.run(total - r.steps_done)
    return np.array(out).reshape(-1, 3)


@pytest.mark.parametrize("precision", ["fp32", "fp16s", "bf16"])
@pytest.mark.parametrize("graph,resident", [(16, 0), (0, 0), (16, 2), (0, 2)])
def test_batch_matches_oracle(gpu, precision, graph, resident):
    """Every lattice of a batch (wall-free, cavity, two street rungs with
    probes) against the FP32 oracle run alone: DDF images bit for bit, and
    each FP64 force sample within 1e-13 of the oracle's probe_force. resident
    0 streams every step through memory; 2 holds the lattices in LDS (tiles
    exchanging their edges) for every segment of every run."""
    specs = _planar_specs()
    pairs = [_make(gpu, precision, sp, oracle=True) for sp in specs]
    sims = [g for g, _ in pairs]
    bt = Batch(sims)
    bt.set_graph_steps(graph)
    bt.set_resident(resident)
    for i, (first, period, count) in _PROBES.items():
        bt.set_probe(i, first, period, count, *specs[i]["box"])
    for run in _RUNS:
        bt.run(run)
    totals = [sum(r[i] for r in _RUNS) for i in range(len(specs))]
    for i, ((g, r), sp) in enumerate(zip(pairs, specs)):
        assert g.steps_done == totals[i]
        if i in _PROBES:
            first, period, count = _PROBES[i]
            ref = _oracle_series(r, sp["box"], first, period, count, totals[i])
            got = bt.probe_series(i)
            assert got.shape == ref.shape and len(ref) > 0, (i, got.shape, ref.shape)
            assert np.all(np.isfinite(got))
            scale = np.abs(ref).max() + 1e-300
            assert np.all(np.abs(got - ref) <= 1e-13 * scale), (i, np.abs(got - ref).max(), scale)
            assert ref[-1, 0] != 0.0            # the ramped flow has reached the cylinder
        else:
            r.run(totals[i])
        assert r.steps_done == totals[i]
        assert_bit_exact(g.download_ddf(), r.download_ddf(), f"lattice {i} DDFs after {totals[i]} steps")


def test_autotune_keeps_state_of_planar_lattices(gpu):
    sims, _ = _planar(gpu, "fp16s")
    solo, _ = _planar(gpu, "fp16s")
    bt = Batch(sims)
    before = [s.state_digest for s in sims]
    assert bt.autotune() in (128, 256, 512)
    assert [s.state_digest for s in sims] == before
    bt.run([24, 24, 24])
    for s in solo:
        s.run(24)
    assert [s.state_digest for s in sims] == [s.state_digest for s in solo]


def test_probe_series_do_not_depend_on_how_runs_are_split(gpu):
    # the force ring is reduced once per 64 steps and at the end of each run:
    # splitting a run at odd places (spans that start mid-ring, runs that end
    # mid-span) must give the same samples and states as one long run
    def run(splits):
        sims, box = _planar(gpu, "fp32")
        bt = Batch(sims)
        bt.set_graph_steps(16)
        bt.set_probe(2, 3, 4, 40, *box)          # odd steps 3, 7, ..., 159
        for k in splits:
            bt.run([k, k, k])
        return [s.state_digest for s in sims], bt.probe_series(2)
    d0, s0 = run([160])
    d1, s1 = run([13, 29, 40, 6, 72])
    assert d1 == d0
    assert s0.shape == s1.shape == (40, 3)
    assert np.array_equal(s1, s0)


def test_probed_lattice_stepped_outside_the_batch_is_refused(gpu):
    from strouhal_amd._lib import StrouhalError
    sims, box = _planar(gpu, "fp32")
    bt = Batch(sims)
    bt.set_probe(2, 3, 4, 10, *box)
    bt.run([8, 8, 8])
    sims[0].run(2)                     # no probes on lattice 0: allowed
    bt.run([4, 4, 4])
    sims[2].run(2)                     # the probed street steps alone
    with pytest.raises(StrouhalError):
        bt.run([4, 4, 4])
    assert sims[2].steps_done == 14
    bt.set_probe(2, 15, 4, 2, *box)    # a new schedule from the current step
    bt.run([6, 6, 6])
    assert bt.probe_series(2).shape == (2, 3)


def test_resident_launches_refuse_lattices_that_do_not_qualify(gpu):
    from strouhal_amd._lib import StrouhalError
    sims = _mixed(gpu, "fp32")                 # two of them are 3-D
    bt = Batch(sims)
    bt.set_resident(2)
    with pytest.raises(StrouhalError):
        bt.run([4, 4, 4])
    assert all(s.steps_done == 0 for s in sims)
    bt.set_resident(1)                         # falls back to the streaming kernel
    bt.run([4, 4, 4])
    assert all(s.steps_done == 4 for s in sims)


@pytest.mark.parametrize("precision", ["fp32", "fp16s"])
def test_resident_tiles_of_every_shape(gpu, precision):
    """Lattices whose tiles wrap onto themselves (one tile across x or y),
    ragged tile grids, walls through tile edges and corners, a moving lid,
    partial cells: resident launches (mode 2) end on the streaming kernel's
    bits and force samples."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import cavity_flags, disk_solid_fractions
    rng = np.random.default_rng(7)

    def lattices():
        out = []
        for nx, ny in ((3, 3), (5, 40), (47, 4), (130, 61), (61, 130)):
            fl = (rng.random((1, ny, nx)) < 0.15).astype(np.uint8)      # scattered solid cells
            s = Simulation(gpu, precision, nx=nx, ny=ny, nz=1, omega=1.7, omega_minus=1.2,
                           flags=fl, wants_forces=True)
            s.init_field(1, 0.04)
            out.append(s)
        c = Simulation(gpu, precision, nx=90, ny=90, nz=1, omega=1.6, lid=(0.06, 0.0, 0.0), ramp_steps=20,
                       flags=cavity_flags(88), wants_forces=True)
        c.set_solid_fractions(disk_solid_fractions(90, 90, 40.5, 50.5, 9.0))
        out.append(c)
        return out

    seed_state = rng.bit_generator.state
    runs = []
    for mode in (0, 2):
        rng.bit_generator.state = seed_state
        sims = lattices()
        bt = Batch(sims)
        bt.set_resident(mode)
        bt.set_graph_steps(16)
        bt.set_probe(len(sims) - 1, 9, 6, 30, (20, 62), (30, 72))
        bt.run([70, 33, 51, 70, 64, 95])
        bt.run([9, 40, 0, 21, 17, 86])
        runs.append(([s.state_digest for s in sims], bt.probe_series(len(sims) - 1)))
    assert runs[0][0] == runs[1][0]
    assert runs[0][1].shape == (29, 3)          # odd steps 9, 15, ..., 177 of 181
    assert np.array_equal(runs[0][1], runs[1][1])


def test_resident_exchange_timeout_fails_the_run_and_the_device_stays_usable(gpu):
    """With a 10 ns wait limit a workgroup that finds a neighbour's exchange word
    not yet written gives up at once and raises the abort flag; every waiting
    workgroup returns and strh_batch_run reports the failure. The next batch,
    with the default limit, runs normally."""
    from strouhal_amd import Simulation
    from strouhal_amd._lib import StrouhalError

    def lattice():
        s = Simulation(gpu, "fp32", nx=640, ny=256, nz=1, omega=1.8)
        s.init_field(1, 0.04)
        return s

    a = lattice()
    bt = Batch([a])
    bt.set_resident(2)
    with pytest.raises(StrouhalError):
        bt.set_resident_timeout(0.0)
    with pytest.raises(StrouhalError):
        bt.set_resident_timeout(61.0)
    bt.set_resident_timeout(1e-8)
    with pytest.raises(StrouhalError):
        bt.run([256])
    bt.close()
    b, c = lattice(), lattice()
    bb, bc = Batch([b]), Batch([c])
    bb.set_resident(2)
    bc.set_resident(0)
    bb.run([130])
    bc.run([130])
    assert b.state_digest == c.state_digest
