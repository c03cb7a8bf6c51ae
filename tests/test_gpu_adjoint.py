"""GPU: Brinkman-porosity objective and its discrete adjoint (strh_vjp).

The FP64 oracle adjoint is pinned to central finite differences on the CPU;
here the HIP path (FP32 forward, FP32 adjoint state, FP64 accumulation) is
held to it:

* the recording forward leaves the lattice bit for bit where strh_run does;
* J within 1e-6 relative on short runs (1e-5 over 400 steps, the FP32
  forward's drift from FP64) and the gradients within 1e-4 of the largest
  component of the FP64 oracle's (the bar for gradients is 1e-4);
* the gradient does not depend on the checkpoint interval, bit for bit;
* lattices with features the adjoint does not cover are refused;
* Smagorinsky LES and the outlet sponge (rates that depend on the state and
  on x), with and without partial cells, SRT and TRT, match the FP64 oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import OracleSim
from tests.helpers import STORAGE

pytestmark = pytest.mark.gpu

INFLOW, SOLID, LID, OUTFLOW = 3, 1, 2, 4


def _channel(nx, ny, nz=1):
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, :, 0] = INFLOW
    fl[:, :, nx - 1] = INFLOW
    fl[:, 0, :] = SOLID
    fl[:, ny - 1, :] = SOLID
    return fl


def _mixed(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    fl = rng.choice([0] * 12 + [1, 2, 3, 4], size=(nz, ny, nx)).astype(np.uint8)
    fl[:, :, 0] = INFLOW
    return fl


def _alpha(shape, seed):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    a = np.zeros(shape, dtype=np.float32)
    a[:, 2:ny - 2, nx // 3:2 * nx // 3] = rng.uniform(0.0, 0.1, size=(nz, ny - 4, 2 * nx // 3 - nx // 3))
    return a


def _pair(gpu, precision, flags, alpha, uin=0.05, ramp=5, force=(1e-5, 0.0, 0.0), lid=(0.0, 0.0, 0.0)):
    from strouhal_amd import Simulation, trt_omegas
    nz, ny, nx = flags.shape
    wp, wm = trt_omegas(0.7, 3.0 / 16.0)
    g = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, uin=uin,
                   ramp_steps=ramp, force=force, lid=lid, flags=flags)
    g.init_field(0, 0.0)
    r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f64", storage="f64", lid=g.lid_vel,
                  uin=g.uin_target, ramp_steps=g.ramp_steps, force=g.force, flags=flags)
    r.init_field(0, 0.0)
    if alpha is not None:
        g.set_porosity(alpha)
        r.set_porosity(alpha)
    return g, r


OBJ = dict(inlet_x=2, outlet_x=15, spinup=6, avg=11, w_p=1.0, w_d=0.5)


def _close(a, b, rel):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    assert scale > 0
    err = np.max(np.abs(a - b)) / scale
    assert err <= rel, err


@pytest.mark.parametrize("porous", [True, False])
def test_vjp_matches_fp64_oracle_channel(gpu, porous):
    fl = _channel(18, 10)
    g, r = _pair(gpu, "fp32", fl, _alpha(fl.shape, 1) if porous else None)
    J, ga, gu, info = g.vjp(**OBJ, checkpoint_every=0)
    Jr, gar, gur = r.vjp(**OBJ)
    assert J == pytest.approx(Jr, rel=1e-6)
    _close(ga.reshape(-1), gar, 1e-4)
    assert gu == pytest.approx(gur, rel=1e-4)
    assert info["checkpoint_every"] == 5 and info["checkpoints"] == 3
    assert np.all(ga[:, [0, 9], :] == 0.0)


def test_vjp_matches_fp64_oracle_3d_mixed_walls(gpu):
    """Every wall kind (solid, moving lid, inflow, dormant outflow) in 3D."""
    fl = _mixed(14, 11, 5, seed=3)
    g, r = _pair(gpu, "fp32", fl, _alpha(fl.shape, 2), lid=(0.03, 0.0, 0.01),
                 force=(1e-5, -5e-6, 2e-6))
    obj = dict(inlet_x=1, outlet_x=12, spinup=3, avg=8, w_p=1.0, w_d=1.0)
    J, ga, gu, _ = g.vjp(**obj, checkpoint_every=4)
    Jr, gar, gur = r.vjp(**obj)
    assert J == pytest.approx(Jr, rel=1e-6)
    _close(ga.reshape(-1), gar, 1e-4)
    assert gu == pytest.approx(gur, rel=1e-4)


@pytest.mark.parametrize("precision", ["fp32", "fp16s", "bf16"])
def test_recording_forward_keeps_the_state(gpu, precision):
    """strh_vjp and strh_objective_eval advance the lattice exactly as strh_run."""
    fl = _channel(20, 9, 2)
    a = _alpha(fl.shape, 4)
    sims = [_pair(gpu, precision, fl, a)[0] for _ in range(3)]
    sims[0].run(17)
    sims[1].vjp(**OBJ)
    sims[2].objective(**OBJ)
    d = [s.state_digest for s in sims]
    assert d[0] == d[1] == d[2]
    assert all(s.steps_done == 17 for s in sims)
    plain = [_pair(gpu, precision, fl, None)[0] for _ in range(2)]
    plain[0].run(17)
    plain[1].vjp(**OBJ)
    assert plain[0].state_digest == plain[1].state_digest


def test_gradient_independent_of_checkpoint_interval(gpu):
    fl = _channel(18, 10)
    out = []
    for k in (1, 3, 17, 0, -1):
        g, _ = _pair(gpu, "fp32", fl, _alpha(fl.shape, 1))
        out.append(g.vjp(**OBJ, checkpoint_every=k))
    for J, ga, gu, _ in out[1:]:
        assert J == out[0][0] and gu == out[0][2]
        assert np.array_equal(ga, out[0][1])


def test_replays_alternate_or_overlap_with_the_same_bits(gpu):
    # One checkpoint level with replays: lattices of at most 2^20 cells replay
    # segment s - 1 on a second stream while the reverse steps of segment s
    # run; larger ones alternate the two on one stream. Either way J and the
    # gradients are the bits of the plan that replays nothing (K = T), and the
    # alternating pass takes at least its replays plus its reverse steps.
    for nx, ny in ((18, 10), (1040, 1024)):
        fl = _channel(nx, ny)
        out = []
        for k in (3, -1):
            g, _ = _pair(gpu, "fp32", fl, _alpha(fl.shape, 2))
            out.append(g.vjp(**OBJ, checkpoint_every=k))
            g.close()
        (J3, ga3, gu3, i3), (J0, ga0, gu0, i0) = out
        assert i0["checkpoints"] == 0 and i3["checkpoint_every"] == 3 and i3["checkpoints"] == 5
        assert J3 == J0 and gu3 == gu0 and np.array_equal(ga3, ga0), (nx, ny)
        if nx * ny > (1 << 20):
            assert i3["reverse_pass_seconds"] >= 0.999 * (i3["recompute_seconds"] + i3["backward_seconds"])


def test_vjp_porous_channel_design_field(gpu):
    """The design-field channel at 128x64 (alpha in {0, 1}, 30% porous) over
    400 steps: gradient within 1e-4 of the FP64 oracle's largest component,
    and the largest components in the same cells."""
    from strouhal_amd.channel_adjoint import channel_case
    g, alpha = channel_case(gpu, "fp32", nx=128, ny=64, ramp_steps=100, margin=8)
    r = OracleSim(128, 64, 1, g.omega, g.omega_minus, arith="f64", storage="f64", uin=g.uin_target,
                  ramp_steps=g.ramp_steps, flags=g.flags, threads=8)
    r.set_porosity(alpha)
    r.init_field(0, 0.0)
    obj = dict(inlet_x=4, outlet_x=123, spinup=150, avg=250, w_p=1.0, w_d=1.0)
    J, ga, gu, info = g.vjp(**obj)
    assert info["checkpoints"] == 0 and info["checkpoint_every"] == 400
    Jr, gar, gur = r.vjp(**obj)
    # FP32 forward against FP64 over 400 steps: J agrees to a few 1e-6
    assert J == pytest.approx(Jr, rel=1e-5)
    _close(ga.reshape(-1), gar, 1e-4)
    assert gu == pytest.approx(gur, rel=1e-4)
    top = np.argsort(-np.abs(gar))[:50]
    assert set(np.argsort(-np.abs(ga.reshape(-1)))[:50]) == set(top)


def test_objective_eval_matches_vjp(gpu):
    fl = _channel(18, 10)
    g1, _ = _pair(gpu, "fp32", fl, _alpha(fl.shape, 1))
    g2, _ = _pair(gpu, "fp32", fl, _alpha(fl.shape, 1))
    assert g1.objective(**OBJ) == g2.vjp(**OBJ)[0]


def test_vjp_refuses_unsupported_lattices(gpu):
    from strouhal_amd import StrouhalError
    fl = _channel(18, 10)
    g, _ = _pair(gpu, "fp32", fl, _alpha(fl.shape, 1))
    g.c_smago = 0.02                      # porosity with LES
    with pytest.raises(StrouhalError) as e:
        g.vjp(**OBJ)
    assert e.value.code == 8
    g.c_smago = 0.0
    with pytest.raises(StrouhalError) as e:
        g.vjp(**dict(OBJ, inlet_x=18))
    assert e.value.code == 1
    with pytest.raises(StrouhalError):
        g.vjp(**dict(OBJ, avg=0))
    assert g.steps_done == 0


def _nt_pair(gpu, precision, nx=20, ny=12, nz=1, sphere=False):
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.cases import disk_solid_fractions, sphere_solid_fractions
    fl = _channel(nx, ny, nz)
    if sphere:
        eps = sphere_solid_fractions(nx, ny, nz, 8.3, ny / 2 - 0.3, nz / 2 + 0.2, 2.7).astype(np.float32)
    else:
        eps = np.repeat(disk_solid_fractions(nx, ny, 8.3, 5.7, 2.6)[None], nz, axis=0).astype(np.float32)
    wp, wm = trt_omegas(0.7, 3.0 / 16.0)
    g = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, uin=0.05,
                   ramp_steps=4, force=(5e-6, 0.0, 0.0), flags=fl)
    g.set_solid_fractions(eps)
    g.init_field(0, 0.0)
    r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f64", storage="f64", uin=g.uin_target,
                  ramp_steps=g.ramp_steps, force=g.force, flags=fl, threads=8)
    r.set_solid_fractions(eps)
    r.init_field(0, 0.0)
    return g, r, eps


NT_OBJ = dict(inlet_x=2, outlet_x=17, spinup=5, avg=10, w_p=1.0, w_d=0.0, w_f=3.0)


@pytest.mark.parametrize("sphere", [False, True])
def test_vjp_solid_fractions_match_fp64_oracle(gpu, sphere):
    """dJ/deps (pressure drop + partial-cell drag) on a Noble-Torczynski disk
    in 2-D and a sphere in 3-D, against the FP64 oracle adjoint."""
    nz = 9 if sphere else 1
    g, r, eps = _nt_pair(gpu, "fp32", nz=nz, sphere=sphere)
    J, _, gu, info = g.vjp(**NT_OBJ, checkpoint_every=4, want_eps=True)
    Jr, _, ger, gur = r.vjp(**NT_OBJ, with_eps=True)
    ge = info["grad_eps"].reshape(-1)
    assert J == pytest.approx(Jr, rel=1e-6)
    _close(ge, ger, 1e-4)
    assert gu == pytest.approx(gur, rel=1e-4)
    assert np.all(ge[eps.reshape(-1) == 0] == 0.0)
    assert np.count_nonzero(ge) > 10


@pytest.mark.parametrize("precision", ["fp32", "fp16s"])
def test_nt_recording_forward_keeps_the_state_and_interval(gpu, precision):
    a, _, _ = _nt_pair(gpu, precision)
    b, _, _ = _nt_pair(gpu, precision)
    c, _, _ = _nt_pair(gpu, precision)
    a.run(15)
    J1, _, u1, i1 = b.vjp(**NT_OBJ, checkpoint_every=0, want_eps=True)
    J2, _, u2, i2 = c.vjp(**NT_OBJ, checkpoint_every=-1, want_eps=True)
    assert a.state_digest == b.state_digest == c.state_digest
    assert J1 == J2 and u1 == u2 and np.array_equal(i1["grad_eps"], i2["grad_eps"])


def _les_pair(gpu, precision, lam, nt, nz=1):
    """The oracle's LES + sponge case: tau0 = 0.56, Cs^2 = 0.3 (strong eddy
    viscosity, so the rate chain carries weight), sponge over x > 15."""
    from strouhal_amd import Simulation
    from strouhal_amd.cases import disk_solid_fractions, sphere_solid_fractions
    nx, ny = 22, 12
    fl = _channel(nx, ny, nz)
    tau = 0.56
    wp = 1.0 / tau
    wm = wp if lam is None else 1.0 / (0.5 + lam / (tau - 0.5))
    g = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, uin=0.06,
                   ramp_steps=3, c_smago=0.3, flags=fl)
    g.sponge = (15.0, 5.0, 0.9)
    eps = None
    if nt:
        if nz > 1:
            eps = sphere_solid_fractions(nx, ny, nz, 8.3, ny / 2 - 0.3, nz / 2 + 0.2, 2.7).astype(np.float32)
        else:
            eps = disk_solid_fractions(nx, ny, 8.3, 5.7, 2.6)[None].astype(np.float32)
        g.set_solid_fractions(eps)
    g.init_field(0, 0.0)
    r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f64", storage="f64", uin=g.uin_target,
                  ramp_steps=g.ramp_steps, c_smago=g.c_smago, flags=fl, threads=8)
    r.sponge = g.sponge
    if nt:
        r.set_solid_fractions(eps)
    r.init_field(0, 0.0)
    return g, r, eps


LES_OBJ = dict(inlet_x=2, outlet_x=19, spinup=4, avg=9, w_p=1.0, w_d=0.0, w_f=2.0)


@pytest.mark.parametrize("lam,nt,nz", [(None, True, 1), (3.0 / 16.0, True, 1), (None, False, 1),
                                       (3.0 / 16.0, True, 7)])
def test_vjp_les_and_sponge_match_fp64_oracle(gpu, lam, nt, nz):
    """dJ/deps and dJ/du_in through the Smagorinsky rate chain and the sponge
    ramp, SRT and TRT, 2-D disk, plain channel and 3-D sphere."""
    g, r, eps = _les_pair(gpu, "fp32", lam, nt, nz)
    J, _, gu, info = g.vjp(**LES_OBJ, checkpoint_every=5, want_eps=nt)
    if nt:
        Jr, _, ger, gur = r.vjp(**LES_OBJ, with_eps=True)
    else:
        Jr, _, gur = r.vjp(**LES_OBJ)
    assert J == pytest.approx(Jr, rel=1e-6)
    assert gu == pytest.approx(gur, rel=1e-4)
    if nt:
        ge = info["grad_eps"].reshape(-1)
        _close(ge, ger, 1e-4)
        assert np.all(ge[eps.reshape(-1) == 0] == 0.0)
        assert np.count_nonzero(ge) > 10


@pytest.mark.parametrize("precision", ["fp32", "fp16s"])
def test_les_recording_forward_keeps_the_state_and_interval(gpu, precision):
    a, _, _ = _les_pair(gpu, precision, 3.0 / 16.0, True)
    b, _, _ = _les_pair(gpu, precision, 3.0 / 16.0, True)
    c, _, _ = _les_pair(gpu, precision, 3.0 / 16.0, True)
    a.run(13)
    J1, _, u1, i1 = b.vjp(**LES_OBJ, checkpoint_every=0, want_eps=True)
    J2, _, u2, i2 = c.vjp(**LES_OBJ, checkpoint_every=-1, want_eps=True)
    assert a.state_digest == b.state_digest == c.state_digest
    assert J1 == J2 and u1 == u2 and np.array_equal(i1["grad_eps"], i2["grad_eps"])
    assert i1["checkpoint_every"] == 4 and i2["checkpoint_every"] == 13


@pytest.mark.parametrize("precision", ["fp16s", "bf16"])
def test_vjp_les_16bit_storage_near_fp64_oracle(gpu, precision):
    """16-bit lattices record the populations in their storage format and the
    reverse step re-applies the inflow corrections of the odd pass: J and the
    gradients stay within the storage precision of the FP64 oracle (FP16 about
    1e-3, BF16 about 1e-2 of the largest component after 13 steps)."""
    g, r, eps = _les_pair(gpu, precision, 3.0 / 16.0, True)
    J, _, gu, info = g.vjp(**LES_OBJ, checkpoint_every=5, want_eps=True)
    Jr, _, ger, gur = r.vjp(**LES_OBJ, with_eps=True)
    ge = info["grad_eps"].reshape(-1)
    tol = 5e-3 if precision == "fp16s" else 5e-2
    assert abs(J - Jr) <= tol * abs(Jr), (J, Jr)
    assert abs(gu - gur) <= tol * abs(gur), (gu, gur)
    err = np.max(np.abs(ge - ger)) / np.max(np.abs(ger))
    assert err <= tol, err
    assert np.all(ge[eps.reshape(-1) == 0] == 0.0)


@pytest.mark.parametrize("kind", ["porous", "partial_cells", "les"])
def test_two_level_checkpoints_match_one_level(gpu, kind):
    """strh_vjp_nested: outer segments replayed once to lay down inner
    checkpoints, inner segments replayed again while recording. Every replay
    starts from the forward pass's bits, so J, the gradients and the final
    state are the single-level plan's bits for any outer / inner lengths,
    ragged ones included."""
    if kind == "porous":
        fl = _channel(18, 10)
        make = lambda: _pair(gpu, "fp32", fl, _alpha(fl.shape, 1))[0]
        obj, kw, plans = OBJ, {}, [(8, 3), (17, 4), (6, 1), (5, 5), (4, 9)]
    elif kind == "partial_cells":
        make = lambda: _nt_pair(gpu, "fp32")[0]
        obj, kw, plans = NT_OBJ, {"want_eps": True}, [(6, 2), (15, 4), (7, 3)]
    else:
        make = lambda: _les_pair(gpu, "fp16s", 3.0 / 16.0, True)[0]
        obj, kw, plans = LES_OBJ, {"want_eps": True}, [(5, 2), (13, 3), (4, 1)]
    ref_sim = make()
    J0, ga0, gu0, i0 = ref_sim.vjp(**obj, checkpoint_every=3, **kw)
    for outer, inner in plans:
        s = make()
        J, ga, gu, info = s.vjp(**obj, checkpoint_every=outer, inner_every=inner, **kw)
        assert J == J0 and gu == gu0, (outer, inner)
        if ga0 is not None:
            assert np.array_equal(ga, ga0), (outer, inner)
        if "grad_eps" in i0:
            assert np.array_equal(info["grad_eps"], i0["grad_eps"]), (outer, inner)
        assert s.state_digest == ref_sim.state_digest
        T = obj["spinup"] + obj["avg"]
        o_eff = min(outer, T)
        i_eff = min(inner, o_eff)
        assert info["outer_every"] == o_eff and info["checkpoint_every"] == i_eff
        assert info["checkpoints"] == (-(-T // o_eff) - 1) + (-(-o_eff // i_eff) - 1)
