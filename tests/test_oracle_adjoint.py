"""CPU: Brinkman porosity and the discrete adjoint of the FP64 oracle.

Neither has a counterpart in the reference (it has only Noble-Torczynski
solid fractions, Kernels.metal:97-104, and no gradients), so nothing there
pins them. They are pinned here by mathematics:

* a uniform-porosity channel driven by a body force relaxes to the analytic
  Brinkman profile u(y) = F/(rho alpha) (1 - cosh(k y) / cosh(k h)),
  k = sqrt(alpha / nu);
* alpha = 0 everywhere is the plain forced update, bit for bit;
* the adjoint's dJ/dalpha and dJ/d(uin) equal central finite differences of
  the FP64 objective to 1e-6 relative (the bar for the product is 1e-4), with
  the perturbation measured after the float32 rounding the porosity field
  goes through.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.oracle import OracleSim

INFLOW, SOLID = 3, 1


def _channel(nx, ny):
    fl = np.zeros((1, ny, nx), dtype=np.uint8)
    fl[:, :, 0] = INFLOW
    fl[:, :, nx - 1] = INFLOW
    fl[:, 0, :] = SOLID
    fl[:, ny - 1, :] = SOLID
    return fl


def _walls(nx, ny):
    fl = np.zeros((1, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = SOLID
    fl[:, ny - 1, :] = SOLID
    return fl


@pytest.mark.parametrize("lam", [None, 3.0 / 16.0])
def test_brinkman_channel_matches_cosh_profile(lam):
    ny, tau, alpha, fb = 34, 0.8, 1e-3, 1e-6
    wp = 1.0 / tau
    wm = wp if lam is None else 1.0 / (0.5 + lam / (tau - 0.5))
    sim = OracleSim(2, ny, 1, wp, wm, arith="f64", storage="f64", force=(fb, 0.0, 0.0),
                    flags=_walls(2, ny))
    sim.set_porosity(np.full(2 * ny, alpha, dtype=np.float32))
    sim.init_field(0, 0.0)
    sim.run(20000)
    m = sim.probe_moments().reshape(ny, 2, 4)
    u = m[1:ny - 1, 0, 0]            # moments are (ux, uy, uz, rho)
    a = float(np.float32(alpha))
    nu = (tau - 0.5) / 3.0
    k = math.sqrt(a / nu)
    h = (ny - 2) / 2.0                        # halfway walls at y = 0.5 and ny - 1.5
    y = np.arange(1, ny - 1) - (ny - 1) / 2.0
    ua = fb / a * (1.0 - np.cosh(k * y) / math.cosh(k * h))
    err = np.max(np.abs(u - ua)) / np.max(ua)
    assert err < 5e-3, err
    # total porous drag = the integral of alpha u over the gap, F (2h - 2 tanh(kh) / k);
    # the rest of the body force goes into wall shear
    rho = m[1:ny - 1, 0, 3]
    drag = fb * (2.0 * h - 2.0 * math.tanh(k * h) / k)
    assert float(np.sum(a * rho * u)) == pytest.approx(drag, rel=2e-3)


@pytest.mark.parametrize("arith,storage", [("f64", "f64"), ("f32", "f32"), ("f32", "f16")])
def test_zero_porosity_is_the_plain_forced_update(arith, storage):
    rng = np.random.default_rng(3)
    fl = _channel(12, 9)

    def make():
        s = OracleSim(12, 9, 2, 1.3, 1.1, arith=arith, storage=storage, force=(2e-5, -1e-5, 0.0),
                      uin=0.04, ramp_steps=6, flags=np.repeat(fl, 2, axis=0))
        s.init_field(0, 0.0)
        return s
    a, b = make(), make()
    b.set_porosity(np.zeros(a.cells, dtype=np.float32))
    raw = a.download_ddf()
    if raw.dtype == np.float64:
        raw = raw + 1e-3 * rng.standard_normal(raw.shape)
    a.upload_ddf(raw)
    b.upload_ddf(raw)
    a.run(9)
    b.run(9)
    assert a.state_digest == b.state_digest


def _adjoint_case(uin=0.05, ramp=5, seed=0):
    nx, ny = 18, 10
    rng = np.random.default_rng(seed)
    fl = _channel(nx, ny)
    alpha = np.zeros((ny, nx), dtype=np.float32)
    alpha[2:8, 6:11] = rng.uniform(0.005, 0.08, size=(6, 5)).astype(np.float32)

    def make(al=alpha, u=uin):
        tau = 0.7
        wp = 1.0 / tau
        wm = 1.0 / (0.5 + 0.1875 / (tau - 0.5))
        s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=u, ramp_steps=ramp,
                      force=(1e-5, 0.0, 0.0), flags=fl)
        s.set_porosity(al)
        s.init_field(0, 0.0)
        return s
    return make, alpha


OBJ = dict(inlet_x=2, outlet_x=15, spinup=6, avg=11, w_p=1.0, w_d=0.5)


def test_adjoint_alpha_gradient_matches_central_differences():
    make, alpha = _adjoint_case()
    J, ga, gu = make().vjp(**OBJ)
    assert J == make().objective(**OBJ)
    ny, nx = alpha.shape
    cells = [(3, 7), (5, 9), (7, 10), (4, 4), (2, 13), (1, 8)]   # porous, plain fluid, wall-adjacent
    h = 1e-4
    for (y, x) in cells:
        ap, am = alpha.copy(), alpha.copy()
        ap[y, x] = np.float32(alpha[y, x] + h)
        am[y, x] = np.float32(alpha[y, x] - h)      # the map is smooth through alpha = 0
        d = float(ap[y, x]) - float(am[y, x])
        fd = (make(ap).objective(**OBJ) - make(am).objective(**OBJ)) / d
        g = ga[y * nx + x]
        assert abs(g - fd) <= 1e-6 * abs(fd), ((y, x), g, fd)
    assert np.all(ga.reshape(ny, nx)[[0, ny - 1], :] == 0.0)   # walls carry no gradient


@pytest.mark.parametrize("ramp", [0, 5])
def test_adjoint_uin_gradient_matches_central_differences(ramp):
    make, alpha = _adjoint_case(ramp=ramp)
    u0, h = 0.05, 1e-5
    J, ga, gu = make(u=u0).vjp(**OBJ)
    fd = (make(u=u0 + h).objective(**OBJ) - make(u=u0 - h).objective(**OBJ)) / (2 * h)
    assert gu == pytest.approx(fd, rel=1e-6)


def test_adjoint_objective_terms_separately():
    """Each objective term alone: pressure drop only, Brinkman drag only."""
    make, alpha = _adjoint_case(seed=1)
    nx = alpha.shape[1]
    for wts in [(1.0, 0.0), (0.0, 1.0)]:
        o = dict(OBJ, w_p=wts[0], w_d=wts[1])
        J, ga, gu = make().vjp(**o)
        y, x = 4, 8
        h = 1e-4
        ap, am = alpha.copy(), alpha.copy()
        ap[y, x] += np.float32(h)
        am[y, x] -= np.float32(h)
        d = float(ap[y, x]) - float(am[y, x])
        fd = (make(ap).objective(**o) - make(am).objective(**o)) / d
        assert ga[y * nx + x] == pytest.approx(fd, rel=1e-6)


def test_adjoint_independent_of_checkpoint_interval():
    """Replaying segments from checkpoints gives the same bits as keeping every
    state, also with several threads, and leaves the lattice where run() does."""
    make, _ = _adjoint_case(seed=2)
    ref = make()
    J0, g0, u0 = ref.vjp(**OBJ, checkpoint_every=17)
    for k, threads in [(1, 1), (4, 3), (0, 4), (6, 2)]:
        s = make()
        s.set_threads(threads)
        J, ga, gu = s.vjp(**OBJ, checkpoint_every=k)
        assert J == J0 and gu == u0 and np.array_equal(ga, g0)
        assert s.state_digest == ref.state_digest
    plain = make()
    plain.run(OBJ["spinup"] + OBJ["avg"])
    assert plain.state_digest == ref.state_digest


def test_adjoint_rejects_unsupported_features():
    make, _ = _adjoint_case()
    s = make()
    s.c_smago = 0.01
    J, ga, gu = s.vjp(**OBJ)
    assert math.isnan(J)


def _nt_case(seed=0):
    """Channel with a Noble-Torczynski disk: partial cells on its rim, solid core."""
    from strouhal_amd.cases import disk_solid_fractions
    nx, ny = 20, 12
    eps = disk_solid_fractions(nx, ny, 8.3, 5.7, 2.6).astype(np.float32)
    fl = _channel(nx, ny)

    def make(e=eps, u=0.05):
        tau = 0.7
        wp = 1.0 / tau
        wm = 1.0 / (0.5 + 0.1875 / (tau - 0.5))
        s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=u, ramp_steps=4,
                      force=(5e-6, 0.0, 0.0), flags=fl)
        s.set_solid_fractions(e)
        s.init_field(0, 0.0)
        return s
    return make, eps


NT_OBJ = dict(inlet_x=2, outlet_x=17, spinup=5, avg=10, w_p=1.0, w_d=0.0, w_f=3.0)


def test_adjoint_solid_fraction_gradient_matches_central_differences():
    """dJ/deps of the pressure drop plus the partial-cell drag, on rim cells
    (0 < eps < 1) and core cells (eps = 1), against central differences."""
    make, eps = _nt_case()
    J, ga, ge, gu = make().vjp(**NT_OBJ, with_eps=True)
    assert J == make().objective(**NT_OBJ)
    ny, nx = eps.shape
    flat = eps.reshape(-1)
    rim = [k for k in range(flat.size) if 0.0 < flat[k] < 1.0]
    core = [k for k in range(flat.size) if flat[k] == 1.0]
    assert len(rim) > 8 and len(core) > 4
    cells = rim[::max(1, len(rim) // 6)][:6] + core[:2]
    h = 1e-4
    for k in cells:
        ep, em = flat.copy(), flat.copy()
        ep[k] = np.float32(flat[k] + h)
        em[k] = np.float32(flat[k] - h)
        d = float(ep[k]) - float(em[k])
        fd = (make(ep.reshape(eps.shape)).objective(**NT_OBJ)
              - make(em.reshape(eps.shape)).objective(**NT_OBJ)) / d
        assert abs(ge[k] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(ge))), (k, flat[k], ge[k], fd)
    assert np.all(ge[flat == 0.0] == 0.0)       # only cells with a solid fraction are design cells
    u0, hu = 0.05, 1e-5
    fdu = (make(u=u0 + hu).objective(**NT_OBJ) - make(u=u0 - hu).objective(**NT_OBJ)) / (2 * hu)
    assert gu == pytest.approx(fdu, rel=1e-6)


def test_adjoint_drag_term_alone():
    """The partial-cell drag objective alone (w_p = 0): J is the time-averaged
    x momentum exchange of the NT cells, its gradient matches differences."""
    make, eps = _nt_case()
    obj = dict(NT_OBJ, w_p=0.0, w_f=1.0)
    J, ga, ge, gu = make().vjp(**obj, with_eps=True)
    assert J > 0.0                                # the body is dragged downstream
    flat = eps.reshape(-1)
    k = int(np.argmax((flat > 0.2) & (flat < 0.8)))
    h = 1e-4
    ep, em = flat.copy(), flat.copy()
    ep[k] += np.float32(h)
    em[k] -= np.float32(h)
    d = float(ep[k]) - float(em[k])
    fd = (make(ep.reshape(eps.shape)).objective(**obj) - make(em.reshape(eps.shape)).objective(**obj)) / d
    assert ge[k] == pytest.approx(fd, rel=1e-6)


@pytest.mark.parametrize("lam", [None, 3.0 / 16.0])
def test_adjoint_with_les_and_sponge_matches_central_differences(lam):
    """Smagorinsky LES (rates from |Pi_neq|) and the outlet sponge on the NT
    disk channel: dJ/deps and dJ/du_in against central differences, SRT and
    TRT (the LES rate enters w+ and w-, and B through tw = 1/w+ - 1/2)."""
    from strouhal_amd.cases import disk_solid_fractions
    nx, ny = 22, 12
    eps = disk_solid_fractions(nx, ny, 8.3, 5.7, 2.6).astype(np.float32)
    fl = _channel(nx, ny)

    def make(e=eps, u=0.06):
        tau = 0.56
        wp = 1.0 / tau
        wm = wp if lam is None else 1.0 / (0.5 + lam / (tau - 0.5))
        s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=u, ramp_steps=3,
                      c_smago=0.3, flags=fl)
        s.sponge = (15.0, 5.0, 0.9)
        s.set_solid_fractions(e)
        s.init_field(0, 0.0)
        return s
    obj = dict(inlet_x=2, outlet_x=19, spinup=4, avg=9, w_p=1.0, w_d=0.0, w_f=2.0)
    J, ga, ge, gu = make().vjp(**obj, with_eps=True)
    assert np.isfinite(J) and J == make().objective(**obj)
    flat = eps.reshape(-1)
    cells = [k for k in range(flat.size) if flat[k] > 0][::4][:7]
    # tau - 1/2 = 0.06 makes B(eps) strongly curved at eps = 1: a smaller step
    # keeps the central difference's truncation error under the bound
    h = 1e-5
    for k in cells:
        ep, em = flat.copy(), flat.copy()
        ep[k] = np.float32(flat[k] + h)
        em[k] = np.float32(flat[k] - h)
        d = float(ep[k]) - float(em[k])
        fd = (make(ep.reshape(eps.shape)).objective(**obj) - make(em.reshape(eps.shape)).objective(**obj)) / d
        assert abs(ge[k] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(ge))), (k, flat[k], ge[k], fd)
    u0, hu = 0.06, 1e-5
    fdu = (make(u=u0 + hu).objective(**obj) - make(u=u0 - hu).objective(**obj)) / (2 * hu)
    assert gu == pytest.approx(fdu, rel=1e-6)
