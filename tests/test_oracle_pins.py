"""Pins of the CPU oracle (oracle/strouhal_oracle.cpp) — CPU only.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand) fails it:

* storage rounding vs numpy float16 / torch bfloat16 (IEEE RNE);
* AA streaming and rest fixed points, bit-exact (Cases.swift:58-125);
* collide vs textbook forms: exact mass conservation, momentum change = F
  (Guo forcing), SRT/TRT relaxation of the symmetric / antisymmetric parts,
  Noble-Torczynski blending and its eps=1 full bounce, Smagorinsky tau_eff
  (Hou et al. 1996), the viscous sponge's tau ramp with Lambda preserved;
* initField moments equal the analytic Taylor-Green fields and the
  non-equilibrium stress equals -(2/3) tau S;
* walls: Poiseuille (TRT 3/16, wall-exact, Cases.swift:286-317), Couette with
  a moving lid (Ladd correction), wall momentum exchange = body force input,
  inflow walls inject the expected mass flux;
* Taylor-Green observed order >= 1.9 (Cases.swift:329-370);
* the lid-driven cavity at Re 400 against Ghia et al. (Cases.swift:11-45).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleSim
from tests.helpers import CX, CY, CZ, OPP, W, random_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C = np.stack([CX, CY, CZ]).astype(np.float64)   # (3, 19)


def trt(tau, lam):
    wp = 1.0 / tau
    if lam is None:
        return float(np.float32(wp)), float(np.float32(wp))
    return float(np.float32(wp)), float(np.float32(1.0 / (0.5 + lam / (tau - 0.5))))


def feq_textbook(rho, u):
    """Shifted D3Q19 equilibrium W_i (rho - 1 + rho (3 c.u + 9/2 (c.u)^2 - 3/2 u^2))."""
    cu = C.T @ u
    return W * (rho - 1.0 + rho * (3.0 * cu + 4.5 * cu * cu - 1.5 * (u @ u)))


def moments(f, F=np.zeros(3)):
    rho = 1.0 + f.sum()
    u = (C @ f + 0.5 * F) / rho
    return rho, u


# ---------------------------------------------------------------- rounding --
def test_f16_rounding_matches_numpy():
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, size=400_000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e-8, 5.96e-8,
                        2.98e-8, 2.9802322e-8, 6.1035156e-05, 6.097555e-05, -3e-5, 1.0 + 2 ** -11,
                        1.0 + 3 * 2 ** -11, np.inf, -np.inf], dtype=np.float32)
    scaled = (rng.standard_normal(200_000) * 1e-3).astype(np.float32)
    for v in (vals, special, scaled):
        ours = orc.f32_to_f16(v)
        ref = v.astype(np.float16).view(np.uint16)
        assert np.array_equal(ours, ref)
        back = orc.f16_to_f32(ref)
        assert np.array_equal(back.view(np.uint32), ref.view(np.float16).astype(np.float32).view(np.uint32))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2 ** 32, size=400_000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    ours = orc.f32_to_bf16(vals)
    ref = torch.from_numpy(vals.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref)
    # the tests' own bf16 rounding helper agrees too
    from tests.helpers import to_storage
    assert np.array_equal(to_storage(vals.astype(np.float64), "bf16"), ref)


# --------------------------------------------------------- streaming / rest --
@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
def test_lone_ddf_streaming(storage):
    n, T = 16, 4
    arith = "f64" if storage == "f64" else "f32"
    s = OracleSim(n, n, n, 0.0, arith=arith, storage=storage)
    f = s.download_ddf()
    c0 = (8 * n + 8) * n + 8
    inj = np.array([i * 0.001 for i in range(19)], dtype=np.float32)
    conv = {"f64": lambda a: a.astype(np.float64), "f32": lambda a: a,
            "f16": orc.f32_to_f16, "bf16": orc.f32_to_bf16}[storage]
    injs = conv(inj)
    f[1:, c0] = injs[1:]
    s.upload_ddf(f)
    s.run(T)
    g = s.download_ddf()
    for i in range(1, 19):
        d = (((8 + T * CZ[i]) % n) * n + (8 + T * CY[i]) % n) * n + (8 + T * CX[i]) % n
        assert g[i, d] == injs[i]
    assert np.count_nonzero(g) == 18


@pytest.mark.parametrize("wall", [1, 4])
def test_lone_ddf_bounces_off_a_wall_cell(wall):
    """Halfway bounce-back at a SOLID cell and the dormant anti-bounce-back of an
    OUTFLOW cell (Kernels.metal:299-300, 328-329: sign -1, no moving-wall term).
    With omega = 0 the collision is the identity, so a lone DDF v in direction i
    streams to A = c0 + (k - 1) c_i, meets the wall cell c0 + k c_i on step k,
    comes back at A in direction opp(i), and after T = 8 steps sits at
    A - (T - k) c_i: +v (SOLID) or -v (OUTFLOW), in slot opp(i), the only
    non-zero slot. k = 3 bounces on an odd pass's pull, k = 4 on its push. Lid
    and inlet speeds are non-zero, so a wrongly applied moving-wall term would
    show."""
    n, T = 16, 8
    for i, k in ((i, k) for i in range(1, 19) for k in (3, 4)):
        c = np.array([CX[i], CY[i], CZ[i]])
        p0 = np.array([8, 8, 8])
        w = (p0 + k * c) % n
        fl = np.zeros((n, n, n), np.uint8)
        fl[w[2], w[1], w[0]] = wall
        s = OracleSim(n, n, n, 0.0, arith="f64", storage="f64", flags=fl, lid=(0.03, 0.02, 0.01),
                      uin=0.05)
        f = s.download_ddf()
        v = 0.001 * i
        c0 = (p0[2] * n + p0[1]) * n + p0[0]
        f[i, c0] = v
        s.upload_ddf(f)
        s.run(T)
        g = s.download_ddf()
        e = (p0 + (k - 1) * c - (T - k) * c) % n      # A - (T - k) c_i
        d = (e[2] * n + e[1]) * n + e[0]
        assert g[OPP[i], d] == (v if wall == 1 else -v), (i, k, g[OPP[i], d])
        assert np.count_nonzero(g) == 1, (i, k)


def test_rest_fixed_points():
    s = OracleSim(32, 32, 32, 1.7, storage="f32")
    d0 = s.state_digest
    s.run(100)
    assert s.state_digest == d0
    n = 34
    fl = np.zeros((1, n, n), np.uint8)
    fl[0, :, 0] = fl[0, :, n - 1] = fl[0, 0, :] = 1
    fl[0, n - 1, :] = 2
    s = OracleSim(n, n, 1, 1.7, storage="f16", flags=fl)
    d0 = s.state_digest
    s.run(100)
    assert s.state_digest == d0


# ------------------------------------------------------------------ collide --
def _cell(seed, umax=0.06):
    return random_state(1, seed, umax=umax, noise=3e-4)[:, 0]


@pytest.mark.parametrize("lam", [None, 3.0 / 16.0, 0.25])
def test_collide_conserves_mass_and_adds_force(lam):
    F = np.array([3e-4, -2e-4, 1e-4])
    wp, wm = trt(0.7, lam)
    s = OracleSim(4, 4, 4, wp, wm, arith="f64", storage="f64", force=tuple(F))
    for seed in range(20):
        f = _cell(seed)
        out, rho, _ = s.collide_cell(f)
        assert abs(out.sum() - f.sum()) < 2e-16
        assert np.allclose(C @ out - C @ f, F, rtol=0, atol=2e-16)
        assert abs(rho - (1.0 + f.sum())) < 1e-16


def test_collide_srt_is_bgk_with_guo_source():
    F = np.array([2e-4, 1e-4, -3e-4])
    tau = 0.66
    wp, _ = trt(tau, None)
    s = OracleSim(4, 4, 4, wp, arith="f64", storage="f64", force=tuple(F))
    for seed in range(20):
        f = _cell(seed)
        rho, u = moments(f, F)
        cu = C.T @ u
        cF = C.T @ F
        guo = (1.0 - 0.5 * wp) * W * (3.0 * (cF - u @ F) + 9.0 * cu * cF)
        expect = f + wp * (feq_textbook(rho, u) - f) + guo
        out, _, _ = s.collide_cell(f)
        assert np.allclose(out, expect, rtol=0, atol=1e-16)


def test_collide_trt_relaxes_symmetric_and_antisymmetric_parts():
    wp, wm = trt(0.9, 3.0 / 16.0)
    s = OracleSim(4, 4, 4, wp, wm, arith="f64", storage="f64")
    for seed in range(20):
        f = _cell(seed)
        rho, u = moments(f)
        fe = feq_textbook(rho, u)
        sym = lambda a: 0.5 * (a + a[OPP])
        asym = lambda a: 0.5 * (a - a[OPP])
        out, _, _ = s.collide_cell(f)
        assert np.allclose(sym(out), sym(f) + wp * (sym(fe) - sym(f)), rtol=0, atol=1e-16)
        assert np.allclose(asym(out), asym(f) + wm * (asym(fe) - asym(f)), rtol=0, atol=1e-16)


def test_collide_noble_torczynski():
    tau = 0.8
    wp, wm = trt(tau, 3.0 / 16.0)
    s = OracleSim(4, 4, 4, wp, wm, arith="f64", storage="f64")
    for seed in range(10):
        f = _cell(seed)
        plain, _, nt0 = s.collide_cell(f, eps=0.0)
        # eps = 1: B = 1, the cell bounces every population into its opposite
        full, _, ntf = s.collide_cell(f, eps=1.0)
        assert np.allclose(full, f[OPP], rtol=0, atol=1e-17)
        assert np.allclose(ntf, 2.0 * (C @ f), rtol=0, atol=5e-16)
        assert np.all(nt0 == 0.0)
        for eps in (0.1, 0.5, 0.93):
            tw = 1.0 / wp - 0.5
            B = eps * tw / ((1.0 - eps) + tw)
            out, _, ntv = s.collide_cell(f, eps=eps)
            expect = f + (1.0 - B) * (plain - f) + B * (f[OPP] - f)
            assert np.allclose(out, expect, rtol=0, atol=1e-16)
            # F = -sum_i c_i B (f_opp - f_i)
            assert np.allclose(ntv, -(C @ (B * (f[OPP] - f))), rtol=0, atol=5e-16)


def test_collide_smagorinsky_closed_form():
    tau0 = 0.52
    cs2 = 0.01
    wp, _ = trt(tau0, None)
    s = OracleSim(4, 4, 4, wp, arith="f64", storage="f64", c_smago=cs2)
    s_off = OracleSim(4, 4, 4, wp, arith="f64", storage="f64", c_smago=0.0)
    for seed in range(10):
        f = _cell(seed)
        rho, u = moments(f)
        fneq = f - feq_textbook(rho, u)
        Pi = (C * fneq) @ C.T
        Q = np.sqrt((Pi * Pi).sum())
        t0 = 1.0 / wp
        teff = 0.5 * (t0 + np.sqrt(t0 * t0 + 25.455844 * cs2 * Q / rho))
        expect = f + (1.0 / teff) * (feq_textbook(rho, u) - f)
        out, _, _ = s.collide_cell(f)
        assert np.allclose(out, expect, rtol=0, atol=1e-16)
        assert not np.allclose(out, s_off.collide_cell(f)[0], rtol=0, atol=1e-12)


def test_collide_sponge_ramps_tau_and_keeps_lambda():
    tau0, lam, tauS = 0.6, 3.0 / 16.0, 1.0
    wp, wm = trt(tau0, lam)
    s = OracleSim(40, 4, 1, wp, wm, arith="f64", storage="f64")
    s.sponge = (10.0, 20.0, tauS)
    for x, frac in ((5, 0.0), (20, 0.5), (35, 1.0)):
        f = _cell(x)
        rho, u = moments(f)
        t0 = 1.0 / wp
        tS = t0 + frac * (tauS - t0)
        L = (1.0 / wp - 0.5) * (1.0 / wm - 0.5)
        wpe, wme = 1.0 / tS, 1.0 / (0.5 + L / (tS - 0.5))
        fe = feq_textbook(rho, u)
        sym = lambda a: 0.5 * (a + a[OPP])
        asym = lambda a: 0.5 * (a - a[OPP])
        out, _, _ = s.collide_cell(f, x=x)
        assert np.allclose(sym(out), sym(f) + wpe * (sym(fe) - sym(f)), rtol=0, atol=1e-16)
        assert np.allclose(asym(out), asym(f) + wme * (asym(fe) - asym(f)), rtol=0, atol=1e-16)


# ---------------------------------------------------------------- initField --
@pytest.mark.parametrize("mode", [1, 2])
def test_init_field_moments_and_stress(mode):
    nx, ny, nz = 16, 12, 10
    A, tau = 0.05, 0.8
    s = OracleSim(nx, ny, nz, 1.0 / tau, arith="f64", storage="f64")
    s.init_field(mode, A)
    m = s.probe_moments()
    z, y, x = np.indices((nz, ny, nx)).reshape(3, -1)
    kx, ky, kz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    if mode == 1:
        xa, ya = kx * (x + 0.5), ky * (y + 0.5)
        ux = A * np.sin(xa) * np.cos(ya)
        uy = -A * np.cos(xa) * np.sin(ya)
        sxx = A * kx * np.cos(xa) * np.cos(ya)
        rhom1 = 0.75 * A * A * (np.cos(2 * xa) + np.cos(2 * ya))
        sxz = syz = np.zeros_like(ux)
    else:
        xa, ya, za = kx * (x + 0.5), ky * (y + 0.5), kz * (z + 0.5)
        ux = A * np.sin(xa) * np.cos(ya) * np.cos(za)
        uy = -A * np.cos(xa) * np.sin(ya) * np.cos(za)
        sxx = A * kx * np.cos(xa) * np.cos(ya) * np.cos(za)
        sxz = -0.5 * A * kz * np.sin(xa) * np.cos(ya) * np.sin(za)
        syz = 0.5 * A * kz * np.cos(xa) * np.sin(ya) * np.sin(za)
        rhom1 = 0.1875 * A * A * (np.cos(2 * xa) + np.cos(2 * ya)) * (np.cos(2 * za) + 2)
    assert np.allclose(m[:, 0], ux, rtol=0, atol=1e-15)
    assert np.allclose(m[:, 1], uy, rtol=0, atol=1e-15)
    assert np.allclose(m[:, 2], 0.0, rtol=0, atol=1e-15)
    assert np.allclose(m[:, 3], 1.0 + rhom1, rtol=0, atol=1e-15)
    f = s.ddf_float()
    rho = m[:, 3]
    u = m[:, :3].T
    cu = C.T @ u
    feq = W[:, None] * (rho - 1.0 + rho * (3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(0)))
    fneq = f - feq
    Pi = np.einsum("ai,bi,in->abn", C, C, fneq)
    assert np.allclose(Pi[0, 0], -(2.0 / 3.0) * tau * sxx, rtol=0, atol=1e-15)
    assert np.allclose(Pi[1, 1], (2.0 / 3.0) * tau * sxx, rtol=0, atol=1e-15)
    assert np.allclose(Pi[0, 2], -(2.0 / 3.0) * tau * sxz, rtol=0, atol=1e-15)
    assert np.allclose(Pi[1, 2], -(2.0 / 3.0) * tau * syz, rtol=0, atol=1e-15)
    assert np.allclose(Pi[0, 1], 0.0, rtol=0, atol=1e-15)


def test_mass_sum_and_moments_on_constructed_state():
    s = OracleSim(5, 4, 3, 1.0, arith="f64", storage="f64")
    f = np.zeros((19, 60))
    f[1, :] = 0.01     # +x population only
    f[0, 7] = 0.5
    s.upload_ddf(f)
    assert s.mass_sum() == pytest.approx(0.01 * 60 + 0.5, abs=1e-15)
    m = s.probe_moments()
    assert m[0, 0] == pytest.approx(0.01 / 1.01, abs=1e-16)
    assert m[7, 3] == pytest.approx(1.51, abs=1e-15)


# ------------------------------------------------------------------- walls --
def test_poiseuille_wall_exact_and_force_balance():
    """Cases.swift:286-317 (H reduced for the CPU): TRT 3/16 halfway bounce-back is
    viscosity-exact; at steady state the wall momentum exchange equals the body force."""
    H, tau, umax = 16, 0.8, 0.05
    nu = (tau - 0.5) / 3.0
    F = 8.0 * nu * umax / H ** 2
    wp, wm = trt(tau, 3.0 / 16.0)
    ny = H + 2
    fl = np.zeros((1, ny, 8), np.uint8)
    fl[0, 0, :] = fl[0, ny - 1, :] = 1
    s = OracleSim(8, ny, 1, wp, wm, arith="f64", storage="f64", force=(F, 0, 0), flags=fl,
                  wants_forces=True)
    s.run((int(8 * H * H / nu) + 1) & ~1)
    m = s.probe_moments()
    err = 0.0
    for j in range(1, H + 1):
        yd = j - 0.5
        exact = F / (2 * nu) * yd * (H - yd)
        err = max(err, abs(m[j * 8 + 4, 0] - exact) / umax)
    assert err < 1e-6
    fx, fy, fz = s.probe_force((0, 7), (0, ny - 1))
    assert fx == pytest.approx(F * 8 * H, rel=1e-6)
    assert abs(fy) < 1e-12 and abs(fz) < 1e-12


def test_couette_moving_lid():
    """Ladd moving-wall correction: a lid at u_w over a resting wall gives the exact
    linear profile u(y) = u_w (y - 1/2) / H (halfway walls at y = 1/2, H + 1/2)."""
    H, tau, uw = 12, 0.8, 0.04
    wp, wm = trt(tau, 3.0 / 16.0)
    ny = H + 2
    fl = np.zeros((1, ny, 4), np.uint8)
    fl[0, 0, :] = 1
    fl[0, ny - 1, :] = 2
    s = OracleSim(4, ny, 1, wp, wm, arith="f64", storage="f64", lid=(uw, 0, 0), flags=fl)
    s.run((int(10 * H * H / ((tau - 0.5) / 3.0)) + 1) & ~1)
    m = s.probe_moments()
    for j in range(1, H + 1):
        assert m[j * 4 + 1, 0] == pytest.approx(uw * (j - 0.5) / H, abs=1e-9)
        assert abs(m[j * 4 + 1, 1]) < 1e-12


def test_inflow_walls_inject_the_flux():
    """Velocity walls at both channel ends (Cases.swift:406-437): at steady state every
    column carries the flux the inlet links inject, 6 w_i c_i.u_w per link."""
    nx, ny, uin = 24, 10, 0.04
    tau = 0.8
    wp, wm = trt(tau, 3.0 / 16.0)
    fl = np.zeros((1, ny, nx), np.uint8)
    fl[0, :, 0] = fl[0, :, nx - 1] = 3
    fl[0, 0, :] = fl[0, ny - 1, :] = 1
    s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=uin, flags=fl, threads=4)
    s.inflow_uniform = True
    s.run(30000)
    m = s.probe_moments().reshape(ny, nx, 4)
    # per fluid row: links 1 (w=1/18), 7, 9, 11, 13 (w=1/36) pull from the inlet
    # column; the two rows next to the solid corners lose one diagonal link each.
    injected = uin * (ny - 2) - 2 * 6.0 / 36.0 * uin
    for x in (1, nx // 2, nx - 2):
        flux = float((m[1:ny - 1, x, 0] * m[1:ny - 1, x, 3]).sum())
        assert flux == pytest.approx(injected, rel=2e-3)


def test_parabolic_inflow_walls_inject_the_flux():
    """The parabolic velocity wall (Kernels.metal:81-87): u_w(y) = max(4 uin
    yd (H - yd) / H^2, 0), yd = y - 1/2, H = ny - 2, taken at the wall cell's
    row. At steady state every column carries what the inlet links inject:
    per fluid row y, links 1, 11, 13 pull from the inlet cell of row y (6 w c_x
    u_w = u_w/3, u_w/6, u_w/6), link 7 from row y - 1 and link 9 from row y + 1
    (u_w/6 each, where that row is an inlet cell, not a solid corner)."""
    nx, ny, uin = 24, 10, 0.04
    H = ny - 2
    tau = 0.8
    wp, wm = trt(tau, 3.0 / 16.0)
    fl = np.zeros((1, ny, nx), np.uint8)
    fl[0, :, 0] = fl[0, :, nx - 1] = 3
    fl[0, 0, :] = fl[0, ny - 1, :] = 1
    s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=uin, flags=fl, threads=4)
    s.run(30000)
    m = s.probe_moments().reshape(ny, nx, 4)

    def uw(y):
        if y < 1 or y > ny - 2:
            return 0.0                       # solid corner: plain bounce-back
        yd = y - 0.5
        return max(4.0 * uin * yd * (H - yd) / (H * H), 0.0)
    injected = sum(uw(y) / 3.0 + uw(y) / 6.0 + uw(y) / 6.0 + uw(y - 1) / 6.0 + uw(y + 1) / 6.0
                   for y in range(1, ny - 1))
    for x in (1, nx // 2, nx - 2):
        flux = float((m[1:ny - 1, x, 0] * m[1:ny - 1, x, 3]).sum())
        assert flux == pytest.approx(injected, rel=2e-3)
    # the profile is the parabola, not the uniform speed: 5.375 uin against 8 uin
    assert injected < 0.75 * uin * H


# ----------------------------------------------------------- accuracy gates --
def test_taylor_green_observed_order():
    """Cases.swift:329-370: diffusive scaling, min observed order >= 1.9 (FP64)."""
    nu = 0.02
    tau = 3.0 * nu + 0.5
    wp, wm = trt(tau, 0.25)
    errs = []
    for N in (32, 64, 128):
        u0 = 0.20 * 32.0 / N
        k = 2 * np.pi / N
        steps = int(np.log(2.0) / (2.0 * nu * k * k)) & ~1
        s = OracleSim(N, N, 1, wp, wm, arith="f64", storage="f64", threads=8)
        s.init_field(1, u0)
        s.run(steps)
        m = s.probe_moments()
        decay = np.exp(-2.0 * nu * k * k * steps)
        y, x = np.indices((N, N)).reshape(2, -1)
        xa, ya = x + 0.5, y + 0.5
        ue = u0 * decay * np.sin(k * xa) * np.cos(k * ya)
        ve = -u0 * decay * np.cos(k * xa) * np.sin(k * ya)
        l2 = np.sqrt(((m[:, 0] - ue) ** 2 + (m[:, 1] - ve) ** 2).sum() / (2 * N * N)) / (u0 * decay)
        errs.append(l2)
    orders = [np.log2(errs[i - 1] / errs[i]) for i in range(1, 3)]
    assert min(orders) >= 1.9, (errs, orders)


def ghia_rms(m, n, nx, ulid, column, u_only=False):
    with open(os.path.join(GOLDEN, "ghia1982.json")) as fh:
        g = json.load(fh)

    def u_at(iy):
        return (m[iy * nx + n // 2, 0] + m[iy * nx + n // 2 + 1, 0]) / 2.0 / ulid

    def v_at(ix):
        return (m[(n // 2) * nx + ix, 1] + m[(n // 2 + 1) * nx + ix, 1]) / 2.0 / ulid

    def interp(coord, fn):
        sp = coord * n + 0.5
        k0 = min(max(int(np.floor(sp)), 1), n - 1)
        fr = sp - k0
        return fn(k0) * (1 - fr) + fn(k0 + 1) * fr

    d = [interp(r["y"], u_at) - r[column] for r in g["u"]]
    if not u_only:
        d += [interp(r["x"], v_at) - r[column] for r in g["v"]]
    return float(np.sqrt(np.mean(np.square(d))))


def test_cavity_re400_against_ghia():
    """The Re 400 cavity gate of Cases.swift:645-686 on a 2-D lattice (128^2 + frame,
    TRT Lambda=1/4, u_lid 0.1, cosine ramp 5000), FP64: u centre-line RMS vs
    Ghia, Ghia & Shin (1982) <= 0.02. (The v column carries the table's suspected
    Re=400 entry at x=0.9063, Cases.swift:7-9, so the gate is on u.)"""
    n, re, ulid = 128, 400.0, 0.1
    nu = float(np.float32(ulid)) * n / re
    wp, wm = trt(3 * nu + 0.5, 0.25)
    nx = n + 2
    fl = np.zeros((1, nx, nx), np.uint8)
    fl[0, :, 0] = fl[0, :, nx - 1] = fl[0, 0, :] = 1
    fl[0, nx - 1, :] = 2
    s = OracleSim(nx, nx, 1, wp, wm, arith="f64", storage="f64", lid=(ulid, 0, 0),
                  ramp_steps=5000, flags=fl, threads=8)
    s.run(30000)
    m = s.probe_moments()
    assert ghia_rms(m, n, nx, float(np.float32(ulid)), "re400", u_only=True) <= 0.02


def test_dissipation_of_the_taylor_green_initial_field():
    """2 nu <S:S> of the 3-D TGV initial field is 3 nu U0^2 / (4 L^2), i.e.
    eps* = 3 / (4 Re) = 4.6875e-4 at Re 1600 (the first entry of the DNS
    table, Cases.swift:1184). initField writes f_neq from the analytic strain
    rate, so the Chapman-Enskog estimate recovers it to O(A^2) (density)."""
    from tests.helpers import dissipation_sum
    n, u0, re = 32, 0.05, 1600.0
    L = n / (2 * np.pi)
    tau = 3 * u0 * L / re + 0.5
    sim = OracleSim(n, n, n, 1.0 / tau, arith="f64", storage="f64", threads=4)
    sim.init_field(2, u0)
    f = sim.download_ddf().reshape(19, -1)
    eps_star = dissipation_sum(f, 1.0 / tau) / n ** 3 * L / u0 ** 3
    assert eps_star == pytest.approx(3.0 / (4.0 * re), rel=2e-3)
