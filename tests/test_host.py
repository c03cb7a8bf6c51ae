"""Host-side checks (CPU only): the C-ABI library loads and exports every
declared symbol, SHA-256 known answers, host geometry builders."""
from __future__ import annotations

import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "strouhal_hip.h")).read()
    return sorted(set(re.findall(r"STRH_API\s+[\w\s\*]+?\b(strh_\w+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    from strouhal_amd import _lib
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for name in syms:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == syms


def test_no_device_reports_status_not_abort():
    from strouhal_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.strh_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        L.strh_ctx_destroy(h)
    else:
        assert rc == 2
        assert b"device" in L.strh_last_error()
    # null arguments are reported, never dereferenced
    assert L.strh_run(None, 1) == 1
    assert L.strh_mass_sum(None, None) == 1


KATS = {
    b"": "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855",
    b"abc": "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad",
    b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq":
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1",
    b"a" * 1_000_000: "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0",
}


@pytest.mark.parametrize("portable", [False, True])
def test_sha256_known_answers(portable):
    from strouhal_amd import sha256, sha256_force_portable
    try:
        sha256_force_portable(portable)
        for msg, digest in KATS.items():
            assert sha256(msg) == digest
        rng = np.random.default_rng(0)
        for n in [1, 55, 56, 63, 64, 65, 119, 120, 127, 128, 1000, 4097]:
            data = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
            assert sha256(data) == hashlib.sha256(data).hexdigest()
    finally:
        sha256_force_portable(False)


def _disk_literal(nx, ny, cx, cy, r):
    """CaseSetups.swift:6-25 transcribed as scalar loops."""
    eps = np.zeros(nx * ny, dtype=np.float32)
    r2, sub = r * r, 16
    lo, hi = int(cx - r) - 2, int(cx + r) + 2
    loY, hiY = int(cy - r) - 2, int(cy + r) + 2
    for y in range(max(0, loY), min(ny - 1, hiY) + 1):
        for x in range(max(0, lo), min(nx - 1, hi) + 1):
            inside = 0
            for sy in range(sub):
                for sx in range(sub):
                    px = float(x) - 0.5 + (float(sx) + 0.5) / float(sub)
                    py = float(y) - 0.5 + (float(sy) + 0.5) / float(sub)
                    dx, dy = px - cx, py - cy
                    if dx * dx + dy * dy <= r2:
                        inside += 1
            eps[y * nx + x] = np.float32(inside) / np.float32(sub * sub)
    return eps.reshape(ny, nx)


def _sphere_literal(n, c, r):
    eps = np.zeros((n, n, n), dtype=np.float32)
    r2, sub = r * r, 8
    cx, cy, cz = c
    for z in range(max(0, int(cz - r) - 2), min(n - 1, int(cz + r) + 2) + 1):
        for y in range(max(0, int(cy - r) - 2), min(n - 1, int(cy + r) + 2) + 1):
            for x in range(max(0, int(cx - r) - 2), min(n - 1, int(cx + r) + 2) + 1):
                inside = 0
                for sz in range(sub):
                    for sy in range(sub):
                        for sx in range(sub):
                            px = float(x) - 0.5 + (float(sx) + 0.5) / float(sub)
                            py = float(y) - 0.5 + (float(sy) + 0.5) / float(sub)
                            pz = float(z) - 0.5 + (float(sz) + 0.5) / float(sub)
                            dx, dy, dz = px - cx, py - cy, pz - cz
                            if dx * dx + dy * dy + dz * dz <= r2:
                                inside += 1
                eps[z, y, x] = np.float32(inside) / np.float32(sub ** 3)
    return eps


def test_disk_and_sphere_fractions_match_the_loops():
    from strouhal_amd.cases import disk_solid_fractions, sphere_solid_fractions
    d = disk_solid_fractions(20, 18, 9.3, 8.7, 4.5)
    assert np.array_equal(d, _disk_literal(20, 18, 9.3, 8.7, 4.5))
    assert abs(float(d.sum()) - np.pi * 4.5 ** 2) < 0.05
    s = sphere_solid_fractions(10, 10, 10, 5.2, 4.9, 5.05, 2.6)
    assert np.array_equal(s, _sphere_literal(10, (5.2, 4.9, 5.05), 2.6))
    assert abs(float(s.sum()) - 4.0 / 3.0 * np.pi * 2.6 ** 3) < 0.1


def test_trt_omegas_and_flags():
    from strouhal_amd import Cell, flags_array, trt_omegas
    wp, wm = trt_omegas(0.8, 3.0 / 16.0)
    assert wp == float(np.float32(1.25))
    assert wm == float(np.float32(1.0 / (0.5 + 0.1875 / 0.3)))
    assert trt_omegas(0.6, None) == (float(np.float32(1 / 0.6)),) * 2
    fl = flags_array(lambda x, y, z: np.where(y == 0, Cell.solid, Cell.fluid), 5, 4, 3)
    assert fl.shape == (3, 4, 5) and fl[:, 0, :].min() == 1 and fl[:, 1:, :].max() == 0


def test_cycle_stats_on_a_sine():
    from strouhal_amd.ladder import cycle_stats
    """Strouhal from upward zero crossings of a pure sine equals its frequency."""
    t = np.arange(0, 4000, 7.0)
    D, um, f = 20, 0.05, 1.0 / 400.0
    cl = 0.8 * np.sin(2 * np.pi * f * t + 0.3)
    cd = 1.5 + 0.1 * np.sin(4 * np.pi * f * t)
    s = cycle_stats(cl, cd, t, D, um)
    assert s["cycles"] >= 8
    assert s["strouhal"] == pytest.approx(f * D / um, rel=1e-3)
    assert s["max_cl"] == pytest.approx(0.8, abs=2e-3)
    assert s["mean_cd"] == pytest.approx(1.5, abs=1e-3)


def test_channel_design_field_fixture():
    """The porous-channel design field is reproducible: its float32 bytes hash
    to the committed SHA-256, 30% of the interior is porous, and the smoothing
    kernel has unit mass."""
    import json
    from strouhal_amd.channel_adjoint import _gauss_kernel, design_field
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "channel_design_field.json")))
    a = design_field(fx["nx"], fx["ny"], sigma=fx["sigma"], solid=fx["solid"], seed=fx["seed"],
                     margin=fx["margin"])
    assert hashlib.sha256(a.tobytes()).hexdigest() == fx["sha256_float32"]
    assert int(a.sum()) == fx["porous_cells"]
    inner = a[0, 1:fx["ny"] - 1, fx["margin"]:fx["nx"] - fx["margin"]]
    assert abs(inner.mean() - fx["solid"]) < 1e-3
    assert a[0, :, :fx["margin"]].max() == 0.0 and a[0, [0, -1], :].max() == 0.0
    assert abs(_gauss_kernel(8.0).sum() - 1.0) < 1e-15


def test_tgv_post_processing():
    """Central differences of a quadratic are exact; the parabolic peak
    refinement recovers the vertex of a sampled parabola; the DNS table's RMS
    distance to itself is zero."""
    import json
    from strouhal_amd.tgv import central_eps, peak, rms_vs, tgv_setup
    t = np.arange(0, 3.0001, 0.1)
    e = 1.0 - 0.3 * t + 0.05 * t * t
    tc, ec = central_eps(t, e)
    assert np.allclose(ec, 0.3 - 0.1 * tc, atol=1e-12)
    ts = np.arange(0, 12, 0.1)
    pt, pe = peak(ts, 0.0128 - 0.002 * (ts - 8.97) ** 2, 0.1)
    assert pt == pytest.approx(8.97, abs=1e-9) and pe == pytest.approx(0.0128, abs=1e-12)
    dns = json.load(open(os.path.join(ROOT, "tests", "golden", "tgv_dns_re1600.json")))
    assert rms_vs(dns["t_star"], dns["eps_star"], dns["t_star"], dns["eps_star"], 14.0) == 0.0
    assert dns["eps_star"][0] == pytest.approx(3.0 / (4.0 * 1600.0), rel=1e-3)
    L, tau, tc_, stride = tgv_setup(256)
    assert stride == 80 and tau == pytest.approx(0.503820, abs=1e-6)


def test_qoi_csv_format(tmp_path):
    """The ladder's QoI export writes the app's `step,C_D,C_L` CSV."""
    from strouhal_amd.ladder import write_qoi_csv
    p = tmp_path / "q.csv"
    write_qoi_csv(str(p), [100.0, 202.0], [3.2, 3.25], [-0.1234567, 0.5])
    assert p.read_text() == "step,C_D,C_L\n100,3.200000,-0.123457\n202,3.250000,0.500000\n"


def _top_level_args(text: str, start: int) -> int:
    """Number of comma-separated arguments of the call whose '(' is at start."""
    depth, n, i, empty = 0, 1, start, True
    while True:
        ch = text[i]
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                return 0 if empty else n
        elif ch == "," and depth == 1:
            n += 1
        elif not ch.isspace() and depth >= 1 and i > start:
            empty = False
        i += 1


def test_swift_facade_calls_only_declared_entry_points():
    # No Swift toolchain here: check the facade against the header instead
    # (names and argument counts of every C call it makes).
    hdr = open(os.path.join(ROOT, "include", "strouhal_hip.h")).read()
    arity = {}
    for m in re.finditer(r"STRH_API\s+[\w\s\*]+?\b(strh_\w+)\s*\(", hdr):
        void = re.match(r"\(\s*void\s*\)", hdr[m.end() - 1:])
        arity[m.group(1)] = 0 if void else _top_level_args(hdr, m.end() - 1)
    src = open(os.path.join(ROOT, "swift", "Sources", "StrouhalCore", "Sim.swift")).read()
    types = set(re.findall(r"\}\s*(strh_\w+)\s*;", hdr)) | set(re.findall(r"typedef\s+struct\s+\w+\s+(strh_\w+)\s*;", hdr))
    calls = [m for m in re.finditer(r"\b(strh_\w+)\s*\(", src) if m.group(1) not in types]
    assert len(calls) >= 20
    for m in calls:
        name = m.group(1)
        assert name in arity, name
        assert _top_level_args(src, m.end() - 1) == arity[name], name
