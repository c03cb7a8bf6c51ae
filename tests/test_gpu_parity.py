"""GPU parity: the HIP step kernel against the FP32 oracle, bit for bit.

Every case starts both implementations from the same raw DDF image, runs the
same number of steps (odd and even counts, so both AA passes end a run) and
compares the raw storage bits. The FP32 oracle restates
Sources/StrouhalCore/Kernels.metal:105-355 in the reference's operation
order with no contraction, as the kernel is built.
"""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from tests.helpers import (CX, CY, CZ, STORAGE, assert_bit_exact, make_pair, random_state,
                           to_storage)

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp32", "fp16s", "bf16"]


def _flags(kind, nx, ny, nz, seed=0):
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    if kind == "periodic":
        return fl
    if kind == "cavity":
        fl[:, 0, :] = 1
        fl[:, :, 0] = 1
        fl[:, :, nx - 1] = 1
        fl[:, ny - 1, :] = 2
        return fl
    if kind == "channel":
        fl[:, 0, :] = 1
        fl[:, ny - 1, :] = 1
        fl[:, :, 0] = 3
        fl[:, :, nx - 1] = 3
        return fl
    if kind == "mixed":
        # every flag kind, scattered, including OUTFLOW (dormant ABB) walls
        rng = np.random.default_rng(seed)
        fl = rng.choice([0, 0, 0, 0, 0, 0, 1, 2, 3, 4], size=(nz, ny, nx)).astype(np.uint8)
        return fl
    raise ValueError(kind)


def _run_compare(g, r, steps_list, precision):
    for steps in steps_list:
        g.run(steps)
        r.run(steps)
        assert g.steps_done == r.steps_done
        assert_bit_exact(g.download_ddf(), r.download_ddf(), f"DDFs after {g.steps_done} steps")


def _seed_both(g, r, precision, seed):
    img = to_storage(random_state(g.cells, seed), precision)
    g.upload_ddf(img)
    r.upload_ddf(img if precision != "fp16s" else img.view(np.uint16))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "uploaded DDFs")


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("shape", [(16, 12, 10), (13, 7, 5), (70, 3, 1), (64, 64, 1)])
def test_periodic_srt_bulk(gpu, precision, shape):
    nx, ny, nz = shape
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.7)
    _seed_both(g, r, precision, seed=nx * 7 + ny)
    _run_compare(g, r, [1, 2, 5], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_periodic_trt_guo_force(gpu, precision):
    g, r = make_pair(gpu, precision, 18, 14, 9, 1.25, 1.6, force=(2e-5, -1e-5, 3e-6))
    _seed_both(g, r, precision, seed=3)
    _run_compare(g, r, [3, 4], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_les_srt_and_trt(gpu, precision):
    for om, omm in [(1.95, None), (1.9, 1.3)]:
        g, r = make_pair(gpu, precision, 17, 11, 6, om, omm, c_smago=0.04, force=(1e-6, 0, 0))
        _seed_both(g, r, precision, seed=11)
        _run_compare(g, r, [2, 3], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_cavity_moving_lid_with_ramp(gpu, precision):
    n = 20
    g, r = make_pair(gpu, precision, n + 2, n + 2, 3, 1.6, 1.1, lid=(0.1, 0.02, 0.0),
                     ramp_steps=7, flags=_flags("cavity", n + 2, n + 2, 3))
    _seed_both(g, r, precision, seed=5)
    _run_compare(g, r, [1, 4, 6], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("uniform", [False, True])
def test_channel_inflow_sponge(gpu, precision, uniform):
    nx, ny = 40, 14
    g, r = make_pair(gpu, precision, nx, ny, 2, 1.5, 1.7, uin=0.06, ramp_steps=5,
                     flags=_flags("channel", nx, ny, 2))
    g.inflow_uniform = uniform
    r.inflow_uniform = uniform
    g.sponge = (25.0, 10.0, 1.0)
    r.sponge = g.sponge
    _seed_both(g, r, precision, seed=7)
    _run_compare(g, r, [3, 3], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixed_flags_outflow_abb(gpu, precision):
    nx, ny, nz = 15, 9, 7
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.4, None, lid=(0.05, 0.0, 0.01), uin=0.04,
                     flags=_flags("mixed", nx, ny, nz, seed=2))
    _seed_both(g, r, precision, seed=9)
    _run_compare(g, r, [1, 1, 4], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_noble_torczynski_with_forces(gpu, precision):
    nx, ny, nz = 24, 18, 4
    rng = np.random.default_rng(4)
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    eps[:, 5:12, 6:14] = rng.uniform(0.0, 1.0, size=(nz, 7, 8)).astype(np.float32)
    eps[:, 8, 9] = 1.0
    eps[:, 6, 7] = 0.0
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.3, 1.5, uin=0.05, force=(1e-5, 0, 0),
                     flags=_flags("channel", nx, ny, nz), wants_forces=True)
    g.set_solid_fractions(eps)
    r.set_solid_fractions(eps)
    _seed_both(g, r, precision, seed=13)
    _run_compare(g, r, [2], precision)
    # force probe: one even+odd pair with momentum exchange on the odd pass
    fg = g.probe_force((3, 18), (2, 15))
    fr = r.probe_force((3, 18), (2, 15))
    assert g.steps_done == r.steps_done
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    field_g = g.download_force()
    field_r = r.force_field().astype(np.float32)
    assert_bit_exact(field_g, field_r, "per-cell force")
    # box sum: FP64, fixed tree order on the GPU vs sequential on the CPU
    scale = np.abs(field_r[:, :3].astype(np.float64)).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_partial_cells_les_sponge_walls(gpu, precision):
    """The LES sphere's kernels (walls, partial cells, Smagorinsky, sponge, no
    body force): plain steps build the pass without the per-cell force output,
    the probe's odd step with it; both stay bit-exact to the oracle."""
    nx, ny, nz = 26, 12, 6
    rng = np.random.default_rng(21)
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    eps[:, 3:9, 5:11] = rng.uniform(0.0, 1.0, size=(nz, 6, 6)).astype(np.float32)
    eps[:, 5, 7] = 1.0
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.85, None, uin=0.05, c_smago=0.1,
                     flags=_flags("channel", nx, ny, nz), wants_forces=True)
    for s in (g, r):
        s.set_solid_fractions(eps)
        s.sponge = (18.0, 7.0, 1.0)
    _seed_both(g, r, precision, seed=23)
    _run_compare(g, r, [2, 4], precision)
    fg = g.probe_force((3, 14), (1, 10))
    fr = r.probe_force((3, 14), (1, 10))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell force")
    scale = np.abs(r.force_field()[:, :3].astype(np.float64)).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)
    _run_compare(g, r, [2], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_partial_cells_in_few_blocks_of_a_large_lattice(gpu, precision):
    """A lattice of 2^22 cells whose partial cells lie in few of its 256-cell
    blocks runs those blocks with the partial-cell kernel and the others with
    the kernel without that path: still bit-exact to the oracle, probe steps
    included (the blocks on both sides of the split hold wall links)."""
    nx, ny, nz = 256, 128, 128
    rng = np.random.default_rng(31)
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    eps[58:70, 56:72, 40:60] = rng.uniform(0.0, 1.0, size=(12, 16, 20)).astype(np.float32)
    eps[64, 64, 50] = 1.0
    eps[3, 1, 10] = 0.5                     # a partial cell beside a wall
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.85, None, uin=0.05, c_smago=0.1,
                     flags=_flags("channel", nx, ny, nz), wants_forces=True)
    for s in (g, r):
        s.set_solid_fractions(eps)
        s.sponge = (200.0, 40.0, 1.0)
    _seed_both(g, r, precision, seed=37)
    _run_compare(g, r, [2], precision)
    fg = g.probe_force((30, 70), (1, 80))
    fr = r.probe_force((30, 70), (1, 80))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell force")
    scale = np.abs(r.force_field()[:, :3].astype(np.float64)).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind,shape,force", [("channel", (30, 14, 2), (1e-5, 0.0, 0.0)),
                                              ("periodic", (64, 8, 3), (0.0, 0.0, 0.0)),
                                              ("mixed", (15, 9, 7), (2e-5, -1e-5, 5e-6))])
def test_brinkman_porosity(gpu, precision, kind, shape, force):
    """Brinkman drag through the Guo terms (also with no body force, and on a
    wall-free lattice that would otherwise take the row-staged odd pass)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + nz)
    alpha = np.zeros((nz, ny, nx), dtype=np.float32)
    x0, x1 = nx // 4, 3 * nx // 4
    alpha[:, 2:ny - 2, x0:x1] = rng.uniform(0.0, 0.3, size=(nz, ny - 4, x1 - x0))
    alpha[:, 3, nx // 2] = 5.0
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.3, 1.6, uin=0.05, ramp_steps=4, force=force,
                     lid=(0.04, 0.0, 0.0), flags=_flags(kind, nx, ny, nz, seed=5), wants_forces=True)
    g.set_porosity(alpha)
    r.set_porosity(alpha)
    _seed_both(g, r, precision, seed=17)
    _run_compare(g, r, [1, 2, 3], precision)
    fg = g.probe_force((1, nx - 2), (1, ny - 2))
    fr = r.probe_force((1, nx - 2), (1, ny - 2))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    scale = np.abs(r.force_field()[:, :3]).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)
    g.set_porosity(None)           # back to the plain lattice
    r.set_porosity(np.zeros(g.cells, dtype=np.float32))
    _run_compare(g, r, [2], precision)


def test_brinkman_rejects_les_nt_and_batches(gpu):
    from strouhal_amd import StrouhalError
    from strouhal_amd.batch import Batch
    g, _ = make_pair(gpu, "fp32", 16, 8, 2, 1.3)
    with pytest.raises(StrouhalError):
        g.set_porosity(np.full(g.cells, -1.0, dtype=np.float32))
    g.set_porosity(np.full(g.cells, 0.1, dtype=np.float32))
    g.c_smago = 0.02
    with pytest.raises(StrouhalError):
        g.run(2)
    g.c_smago = 0.0
    g.set_solid_fractions(np.zeros(g.cells, dtype=np.float32))
    with pytest.raises(StrouhalError):
        g.run(2)
    with pytest.raises(StrouhalError):
        Batch([g])
    assert g.steps_done == 0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_row_staged_odd_pass(gpu, precision):
    """Wall-free lattices with nx a power of two >= 64 run the odd pass with LDS
    row staging (256- or 512-thread workgroups); same bits as the oracle."""
    cases = [((64, 8, 4), dict(omega=1.7)),
             ((128, 6, 3), dict(omega=1.2, omega_minus=1.6, force=(2e-5, 1e-5, -1e-5))),
             ((256, 2, 2), dict(omega=1.9, omega_minus=1.1, c_smago=0.04)),
             ((64, 4, 1), dict(omega=1.3)),
             ((512, 2, 2), dict(omega=1.6))]
    for (nx, ny, nz), kw in cases:
        om = kw.pop("omega")
        omm = kw.pop("omega_minus", None)
        g, r = make_pair(gpu, precision, nx, ny, nz, om, omm, **kw)
        _seed_both(g, r, precision, seed=nx + ny)
        _run_compare(g, r, [1, 2, 3], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_row_staged_odd_pass_partial_cells(gpu, precision):
    """The row-staged odd pass with Noble-Torczynski cells, LES and a body
    force on a periodic 128 x 5 x 3 lattice: DDFs and the per-cell
    momentum-exchange force of a force probe, bit for bit."""
    rng = np.random.default_rng(17)
    g, r = make_pair(gpu, precision, 128, 5, 3, 1.5, 1.3, force=(1e-5, 0.0, -2e-5), c_smago=0.02,
                     wants_forces=True)
    eps = np.where(rng.random(g.cells) < 0.3, rng.random(g.cells), 0.0).astype(np.float32)
    g.set_solid_fractions(eps)
    r.set_solid_fractions(eps)
    _seed_both(g, r, precision, seed=59)
    _run_compare(g, r, [1, 3], precision)          # the probe starts from an even step count
    g.probe_force((0, 127), (0, 4))
    r.probe_force((0, 127), (0, 4))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell force")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pow2_walls_blocked_layout(gpu, precision):
    """Wall lattices with nx a power of two (64, 128, 512) and the blocked DDF
    layout, across block boundaries: bounce-back, moving walls, dormant ABB,
    NT partial cells with LES and momentum exchange; same bits as the oracle,
    DDFs and per-cell forces."""
    rng = np.random.default_rng(21)
    # channel with parabolic inflow, sponge and TRT
    g, r = make_pair(gpu, precision, 64, 12, 3, 1.5, 1.7, uin=0.06, ramp_steps=5,
                     flags=_flags("channel", 64, 12, 3))
    g.sponge = (40.0, 16.0, 1.0)
    r.sponge = g.sponge
    _seed_both(g, r, precision, seed=41)
    _run_compare(g, r, [1, 2, 3], precision)
    # every flag kind scattered, moving lid, body force, forces probed
    fl = rng.choice([0] * 8 + [1, 2, 3, 4], size=(4, 6, 128)).astype(np.uint8)
    g, r = make_pair(gpu, precision, 128, 6, 4, 1.4, 1.2, lid=(0.05, 0.0, 0.01), uin=0.04,
                     force=(1e-5, -2e-5, 0.0), flags=fl, wants_forces=True)
    _seed_both(g, r, precision, seed=43)
    _run_compare(g, r, [1, 3], precision)
    fg = g.probe_force((0, 127), (0, 5))
    fr = r.probe_force((0, 127), (0, 5))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell force")
    scale = np.abs(r.force_field()[:, :3]).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)
    # NT partial cells with LES in a 512-wide channel (512-thread workgroups)
    nx, ny, nz = 512, 6, 2
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    eps[:, 2:4, 100:140] = rng.uniform(0.0, 1.0, size=(nz, 2, 40)).astype(np.float32)
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.8, None, uin=0.05, c_smago=0.04,
                     flags=_flags("channel", nx, ny, nz), wants_forces=True)
    g.set_solid_fractions(eps)
    r.set_solid_fractions(eps)
    _seed_both(g, r, precision, seed=47)
    _run_compare(g, r, [1, 3], precision)
    fg = g.probe_force((90, 150), (0, 5))
    fr = r.probe_force((90, 150), (0, 5))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after NT force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell NT force")


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind", ["channel", "periodic"])
def test_pow2_brinkman_blocked_layout(gpu, precision, kind):
    nx, ny, nz = 64, 10, 3
    rng = np.random.default_rng(5)
    alpha = rng.uniform(0.0, 0.5, size=(nz, ny, nx)).astype(np.float32)
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.3, 1.6, uin=0.05, force=(1e-5, 0.0, 2e-6),
                     flags=_flags(kind, nx, ny, nz), wants_forces=True)
    g.set_porosity(alpha)
    r.set_porosity(alpha)
    _seed_both(g, r, precision, seed=53)
    _run_compare(g, r, [1, 2, 3], precision)
    fg = g.probe_force((1, nx - 2), (1, ny - 2))
    fr = r.probe_force((1, nx - 2), (1, ny - 2))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_row_staged_noble_torczynski_force(gpu, precision):
    nx, ny, nz = 64, 12, 4
    rng = np.random.default_rng(8)
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    eps[:, 3:9, 20:30] = rng.uniform(0.0, 1.0, size=(nz, 6, 10)).astype(np.float32)
    g, r = make_pair(gpu, precision, nx, ny, nz, 1.4, 1.7, force=(1e-5, 0, 0), wants_forces=True)
    g.set_solid_fractions(eps)
    r.set_solid_fractions(eps)
    _seed_both(g, r, precision, seed=31)
    _run_compare(g, r, [2], precision)
    fg = g.probe_force((15, 35), (1, 10))
    fr = r.probe_force((15, 35), (1, 10))
    assert_bit_exact(g.download_ddf(), r.download_ddf(), "DDFs after force probe")
    assert_bit_exact(g.download_force(), r.force_field().astype(np.float32), "per-cell force")
    scale = np.abs(r.force_field()[:, :3]).sum() + 1e-300
    assert np.all(np.abs(fg - fr) <= 1e-13 * scale), (fg, fr)


def test_streaming_lone_ddf(gpu):
    """Cases.swift:85-125: omega = 0, each DDF lands at c0 + T c_i, 18 nonzero slots."""
    from strouhal_amd import Simulation
    for precision in PRECISIONS:
        n = 16
        sim = Simulation(gpu, precision, nx=n, ny=n, nz=n, omega=0.0)
        N = sim.cells
        f = sim.download_ddf()
        c0 = (8 * n + 8) * n + 8
        inj = to_storage(np.array([[i * 0.001] for i in range(19)]), precision)[:, 0]
        for i in range(1, 19):
            f[i, c0] = inj[i]
        sim.upload_ddf(f)
        T = 4
        sim.run(T)
        g = sim.download_ddf()
        for i in range(1, 19):
            dest = (((8 + T * CZ[i]) % n) * n + (8 + T * CY[i]) % n) * n + (8 + T * CX[i]) % n
            assert g[i, dest].tobytes() == inj[i].tobytes(), (precision, i)
        assert np.count_nonzero(g.view(np.uint16 if g.dtype.itemsize == 2 else np.uint32)) == 18
        assert N == n ** 3


def test_rest_fixed_points(gpu):
    """Cases.swift:58-83: rest states are bitwise fixed points (periodic and cavity)."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import cavity_flags
    for precision in PRECISIONS:
        sim = Simulation(gpu, precision, nx=32, ny=32, nz=32, omega=1.7)
        d0 = sim.state_digest
        sim.run(100)
        assert sim.state_digest == d0
        sim = Simulation(gpu, precision, nx=34, ny=34, nz=1, omega=1.7, flags=cavity_flags(32))
        d0 = sim.state_digest
        sim.run(100)
        assert sim.state_digest == d0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_digest_matches_hashlib_and_oracle(gpu, precision):
    g, r = make_pair(gpu, precision, 33, 9, 5, 1.8)
    _seed_both(g, r, precision, seed=21)
    g.run(3)
    r.run(3)
    img = g.download_ddf()
    assert g.state_digest == hashlib.sha256(img.tobytes()).hexdigest() == r.state_digest


@pytest.mark.parametrize("precision", PRECISIONS)
def test_moments_and_mass_and_energy(gpu, precision):
    g, r = make_pair(gpu, precision, 21, 13, 7, 1.6, force=(1e-5, 2e-5, 0.0),
                     flags=_flags("cavity", 21, 13, 7))
    _seed_both(g, r, precision, seed=17)
    g.run(4)
    r.run(4)
    mg = g.probe_moments()
    mr = r.probe_moments().astype(np.float32)
    assert_bit_exact(mg, mr, "moments")
    ddf = r.ddf_float()
    scale = np.abs(ddf).sum()
    assert abs(g.mass_sum() - r.mass_sum()) <= 1e-13 * scale
    ke_ref = float(((mg[:, 0] * mg[:, 0] + mg[:, 1] * mg[:, 1]) + mg[:, 2] * mg[:, 2])
                   .astype(np.float64).sum())
    assert abs(g.kinetic_energy() - ke_ref) <= 1e-13 * ke_ref
    # viscous dissipation: GPU FP64 fixed-order sum vs the numpy FP64 sum of the
    # same (bit-identical) state; walls excluded
    from tests.helpers import dissipation_sum
    d_ref = dissipation_sum(ddf, g.omega, fluid=(g.flags.reshape(-1) == 0))
    assert d_ref > 0 and abs(g.dissipation() - d_ref) <= 1e-12 * d_ref
    g.run(1)
    from strouhal_amd import StrouhalError
    with pytest.raises(StrouhalError):
        g.probe_moments()
    with pytest.raises(StrouhalError):
        g.dissipation()


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("mode", [1, 2])
def test_init_field_matches_oracle(gpu, precision, mode):
    """initField (Kernels.metal:393-453): GPU sinf/cosf vs glibc, within 4 ulp of storage."""
    from oracle.oracle import OracleSim
    from strouhal_amd import Simulation
    nx, ny, nz = 24, 20, 12
    g = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=1.3)
    r = OracleSim(nx, ny, nz, g.omega, arith="f32", storage=STORAGE[precision])
    g.init_field(mode, 0.05)
    r.init_field(mode, 0.05)
    a = r.ddf_float()
    gg = g.download_ddf()
    if precision == "bf16":
        b = (gg.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        ulp = 2.0 ** -7
    elif precision == "fp16s":
        b = gg.astype(np.float64)
        ulp = 2.0 ** -10
    else:
        b = gg.astype(np.float64)
        ulp = 2.0 ** -23
    # GPU sinf/cosf and glibc differ by an ulp or two of the trig values; the
    # stored sum feq + fneq carries that at the scale of each slot's largest
    # value: allow 16 storage ulps of the slot maximum.
    tol = 16 * ulp * np.abs(a).max(axis=1, keepdims=True)
    assert np.all(np.abs(a - b) <= tol)
    # and the field is the same one: relative L2 difference at the FP32 level
    assert np.linalg.norm(a - b) <= 16 * ulp * np.linalg.norm(a)


def test_determinism_run_twice(gpu):
    """Cases.swift:374-399: two identical runs give identical digests."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import cavity
    for precision in ["fp32", "fp16s"]:
        digests = []
        for _ in range(2):
            sim = cavity(gpu, precision, n=128, re=1000.0, lam=0.25)
            sim.run(10_000)
            digests.append(sim.state_digest)
        assert digests[0] == digests[1]
        d2 = []
        for _ in range(2):
            sim = Simulation(gpu, precision, nx=128, ny=128, nz=128, omega=1.9)
            sim.init_field(1, 0.05)
            sim.run(100)
            d2.append((sim.state_digest, sim.mass_sum(), sim.kinetic_energy()))
        assert d2[0] == d2[1]


def test_snapshot_restore(gpu):
    from strouhal_amd import Simulation
    sim = Simulation(gpu, "fp32", nx=20, ny=20, nz=20, omega=1.8)
    sim.init_field(2, 0.05)
    sim.run(4)
    sim.snapshot()
    d = sim.state_digest
    sim.run(6)
    after = sim.state_digest
    sim.restore()
    assert sim.steps_done == 4 and sim.state_digest == d
    sim.run(6)
    assert sim.state_digest == after


@pytest.mark.parametrize("precision,shape", [("fp32", (40, 24, 18)), ("fp16s", (33, 7, 5)),
                                              ("bf16", (128, 80, 2))])
def test_state_file_round_trip(gpu, precision, shape, tmp_path):
    """save_state / load_state: the file holds the canonical image, the step
    count and the digest; a loaded lattice continues bit for bit; a corrupted
    or mismatched file is refused before anything is uploaded."""
    import hashlib
    from strouhal_amd import Simulation, StrouhalError
    nx, ny, nz = shape
    fl = _flags("mixed", nx, ny, nz, seed=3)
    mk = lambda: Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=1.6, omega_minus=1.3,
                            lid=(0.04, 0.0, 0.0), uin=0.05, ramp_steps=7, flags=fl)
    a = mk()
    a.upload_ddf(to_storage(random_state(a.cells, 61), precision))
    a.run(9)
    path = str(tmp_path / "state.strh")
    a.save_state(path)
    raw = open(path, "rb").read()
    img = a.download_ddf()
    assert raw[:8] == b"STRHDDF1" and len(raw) == 128 + img.nbytes
    assert raw[128:] == img.tobytes()
    assert raw[48:80].hex() == a.state_digest == hashlib.sha256(img.tobytes()).hexdigest()
    b = mk()
    b.load_state(path)
    assert b.steps_done == 9 and b.state_digest == a.state_digest
    a.run(5)
    b.run(5)
    assert b.state_digest == a.state_digest
    bad = bytearray(raw)
    bad[200] ^= 0x40
    open(path, "wb").write(bytes(bad))
    c = mk()
    before = c.state_digest
    with pytest.raises(StrouhalError):
        c.load_state(path)
    assert c.state_digest == before and c.steps_done == 0
    open(path, "wb").write(raw)
    other = Simulation(gpu, precision, nx=nx + 1, ny=ny, nz=nz, omega=1.6)
    with pytest.raises(StrouhalError):
        other.load_state(path)
    with pytest.raises(StrouhalError):
        c.load_state(str(tmp_path / "missing.strh"))


def test_error_paths(gpu):
    from strouhal_amd import Simulation, StrouhalError
    sim = Simulation(gpu, "fp32", nx=8, ny=8, nz=8, omega=1.0)
    with pytest.raises(StrouhalError) as e:
        sim.probe_force((0, 3), (0, 3))
    assert e.value.code == 6
    with pytest.raises(ValueError):
        sim.upload_ddf(np.zeros((19, 10), dtype=np.float32))
    import ctypes
    from strouhal_amd._lib import lib
    small = np.zeros(10, dtype=np.float32)
    assert lib().strh_upload_ddf(sim._h, small.ctypes.data_as(ctypes.c_void_p), small.nbytes) == 1
    with pytest.raises(StrouhalError):
        sim.init_field(7, 0.1)
    with pytest.raises(StrouhalError):
        Simulation(gpu, "fp32", nx=0, ny=8, nz=8, omega=1.0)
    sim2 = Simulation(gpu, "fp32", nx=8, ny=8, nz=8, omega=1.0, wants_forces=True)
    with pytest.raises(StrouhalError):
        sim2.probe_force((0, 9), (0, 3))


def test_blocked_image_beyond_4gib(gpu):
    """A wall-free FP32 lattice whose blocked DDF image exceeds 4 GiB
    (512 x 512 x 216: 4.30 GB). The even pass addresses it from the wave's
    block base, the row-staged odd pass from row bases; two steps from the
    Taylor-Green field equal the FP32 oracle's bit for bit. Porosity and
    batches, which run the direct odd pass, refuse the lattice."""
    import ctypes
    from oracle.oracle import OracleSim
    from strouhal_amd import Simulation, StrouhalError
    from strouhal_amd._lib import lib
    from strouhal_amd.batch import Batch
    nx, ny, nz = 512, 512, 216
    g = Simulation(gpu, "fp32", nx=nx, ny=ny, nz=nz, omega=1.9, omega_minus=1.2)
    sz = ctypes.c_size_t()
    lib().strh_ddf_bytes(g._h, ctypes.byref(sz))
    assert sz.value >= 1 << 32
    g.init_field(2, 0.05)
    r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f32", storage="f32", threads=16)
    r.upload_ddf(g.download_ddf())
    g.run(2)
    r.run(2)
    assert g.state_digest == r.state_digest
    with pytest.raises(StrouhalError) as e:
        g.set_porosity(np.zeros(g.cells, dtype=np.float32))
    assert e.value.code == 8
    with pytest.raises(StrouhalError):
        Batch([g])
    g.close()
