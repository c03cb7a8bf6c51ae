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


def test_autotune_keeps_state(gpu):
    sims = _mixed(gpu, "fp16s")
    bt = Batch(sims)
    before = [s.state_digest for s in sims]
    bs = bt.autotune()
    assert bs in (128, 256, 512)
    assert [s.state_digest for s in sims] == before
    assert all(s.steps_done == 0 for s in sims)


def _planar(gpu, precision):
    """Lattices one cell deep of each kind the tiled launches cover: periodic
    wall-free, a lid-driven cavity (moving wall, ramp) and the vortex street
    (partial cells, ramped parabolic inflow, outflow, sponge, force probes).
    Sizes are not multiples of the tile patches."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import cavity, vortex_street
    a = Simulation(gpu, precision, nx=100, ny=40, nz=1, omega=1.7)
    a.init_field(1, 0.05)
    b = cavity(gpu, precision, n=66, re=1e3, ramp_steps=30)
    street = vortex_street(gpu, precision, D=8, uin_max=0.1)
    return [a, b, street.sim], (street.box_x, street.box_y)


def _run_planar(gpu, precision, mode, graph):
    sims, box = _planar(gpu, precision)
    bt = Batch(sims)
    bt.set_tiles(mode)
    bt.set_graph_steps(graph)
    bt.set_probe(2, 7, 6, 12, *box)      # odd steps 7, 13, ..., 73
    bt.run([36, 50, 44])      # all even at the start; segments of 36, 8 and 6 steps
    bt.run([3, 0, 17])        # the street alone from an odd step: one single step first
    bt.run([22, 9, 0])        # mixed parities (single steps), then one lattice even
    bt.run([40, 24, 16])      # all odd: one single step, then tiles
    return ([s.state_digest for s in sims], [s.steps_done for s in sims], bt.probe_series(2),
            bt.tiled_steps)


@pytest.mark.parametrize("precision", ["fp32", "fp16s", "bf16"])
@pytest.mark.parametrize("mode,graph", [(1, 16), (2, 16), (1, 0), (2, 12), (1, 64)])
def test_tiled_batch_matches_one_launch_per_step(gpu, precision, mode, graph):
    d0, n0, s0, t0 = _run_planar(gpu, precision, 0, graph)
    d1, n1, s1, t1 = _run_planar(gpu, precision, mode, graph)
    assert t0 == 0
    assert t1 >= 64, "the tiled launches did not run"
    assert n1 == n0
    assert d1 == d0
    assert s1.shape == s0.shape == (12, 3)
    assert s0[-1, 0] != 0.0          # the ramped flow has reached the cylinder
    assert np.array_equal(s1, s0)


def test_tiled_batch_matches_solo_runs(gpu):
    sims, _ = _planar(gpu, "fp32")
    solo, _ = _planar(gpu, "fp32")
    bt = Batch(sims)
    bt.set_tiles(1)
    bt.run([40, 56, 72])
    assert bt.tiled_steps >= 56
    for s, k in zip(solo, [40, 56, 72]):
        s.run(k)
    for x, y in zip(sims, solo):
        assert x.state_digest == y.state_digest


def test_tiles_skip_lattices_they_do_not_cover(gpu):
    # nz > 1 in the batch: every segment runs one launch per step
    sims = _mixed(gpu, "fp32")
    bt = Batch(sims)
    bt.set_tiles(1)
    bt.run([16, 16, 16])
    assert bt.tiled_steps == 0


def test_autotune_over_tile_modes_keeps_state(gpu):
    sims, _ = _planar(gpu, "fp16s")
    solo, _ = _planar(gpu, "fp16s")
    bt = Batch(sims)
    before = [s.state_digest for s in sims]
    bs = bt.autotune()
    assert bs in (128, 256, 512) and bt.tiles in (0, 1, 2)
    assert [s.state_digest for s in sims] == before
    bt.run([24, 24, 24])
    for s in solo:
        s.run(24)
    assert [s.state_digest for s in sims] == [s.state_digest for s in solo]


@pytest.mark.parametrize("mode", [1, 2])
def test_tiled_batch_with_several_samples_per_launch(gpu, mode):
    # probe period 2 < K: one tiled launch takes up to two samples, each in
    # its own buffer of the force ring
    def run(m):
        sims, box = _planar(gpu, "fp32")
        bt = Batch(sims)
        bt.set_tiles(m)
        bt.set_graph_steps(16)
        bt.set_probe(2, 5, 2, 40, *box)          # odd steps 5, 7, ..., 83
        bt.run([48, 48, 88])
        return [s.state_digest for s in sims], bt.probe_series(2), bt.tiled_steps
    d0, s0, _ = run(0)
    d1, s1, t1 = run(mode)
    assert t1 >= 80
    assert d1 == d0
    assert s1.shape == (40, 3) and np.array_equal(s1, s0)


@pytest.mark.parametrize("precision,nx,ny", [("fp32", 32, 32), ("fp32", 33, 45), ("fp16s", 64, 32),
                                              ("bf16", 70, 33)])
def test_tiled_batch_on_lattices_of_about_one_tile(gpu, precision, nx, ny):
    # tiles wrap onto themselves: wrapped duplicates evolve like the cells
    # they stand for and only the owned patch cells are written back
    from strouhal_amd import Simulation

    def make():
        a = Simulation(gpu, precision, nx=nx, ny=ny, nz=1, omega=1.6)
        a.init_field(1, 0.05)
        b = Simulation(gpu, precision, nx=nx, ny=ny, nz=1, omega=1.5, lid=(0.05, 0.0, 0.0),
                       flags=lambda x, y, z: np.where(y == 0, 1, np.where(y == ny - 1, 2, 0)))
        return [a, b]
    tiled, solo = make(), make()
    bt = Batch(tiled)
    bt.set_tiles(1)
    bt.run([40, 24])
    assert bt.tiled_steps >= 24
    for s, k in zip(solo, [40, 24]):
        s.run(k)
    assert [s.state_digest for s in tiled] == [s.state_digest for s in solo]


@pytest.mark.parametrize("mode", [0, 1])
def test_probe_series_do_not_depend_on_how_runs_are_split(gpu, mode):
    # the force ring is reduced once per 64 steps and at the end of each run:
    # splitting a run at odd places (spans that start mid-ring, runs that end
    # mid-span) must give the same samples and states as one long run
    def run(splits):
        sims, box = _planar(gpu, "fp32")
        bt = Batch(sims)
        bt.set_tiles(mode)
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
