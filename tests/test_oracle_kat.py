"""Pins the CPU oracle with the reference's own known-answer tests and with
closed forms (CPU only).  Each test cites the reference test it restates."""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

from spectral_amd import _abi as A
from spectral_amd import scenes
from spectral_amd.core import Camera, Dielectric, Lambertian, Scene, axis_box, opts, wedge

GOLD = json.loads((Path(__file__).parent / "golden" / "rng_vectors.json").read_text())
F3 = C.c_float * 3


def v(x):
    return F3(*x)


# ---------------------------------------------------------------- rng.rs
def test_rng_frozen_first_draws(oracle):
    # rng.rs:75-82
    out = (C.c_uint32 * 16)()
    oracle.lib().or_rng_draws_u32(0xABCD, 0x1234, 16, out)
    assert list(out)[:4] == [0xA1DD5847, 0x13248A2D, 0x37278DEF, 0xDE1D5B62]
    assert list(out) == GOLD["draws_abcd_1234_16"]


def test_rng_vectors_and_mix2(oracle):
    L = oracle.lib()
    for e in GOLD["draws_more"]:
        out = (C.c_uint32 * 8)()
        L.or_rng_draws_u32(e["key"], e["seed"], 8, out)
        assert list(out) == e["draws"]
    for e in GOLD["mix2"]:
        assert L.or_mix2(e["a"], e["b"]) == e["out"]
    f = (C.c_float * 8)()
    L.or_rng_draws_f32(7, 12345, 8, f)
    assert [float(x) for x in f] == GOLD["f32_bits"][0]["f32"]


def test_rng_properties(oracle):
    # rng.rs:46-71: same key -> same; different key -> different; [0,1)
    L = oracle.lib()
    a, b = (C.c_float * 1000)(), (C.c_float * 1000)()
    L.or_rng_draws_f32(7, 12345, 16, a)
    L.or_rng_draws_f32(7, 12345, 16, b)
    assert list(a)[:16] == list(b)[:16]
    L.or_rng_draws_f32(8, 12345, 4, b)
    assert list(a)[:4] != list(b)[:4]
    L.or_rng_draws_f32(1, 2, 1000, a)
    x = np.array(a)
    assert (x >= 0).all() and (x < 1).all()


# ---------------------------------------------------------------- spectrum.rs
def test_hero_comb_stratified(oracle):
    # spectrum.rs:39-55
    out = (C.c_float * 4)()
    oracle.lib().or_hero_lambdas(0.1, out)
    lam = list(out)
    assert all(380.0 <= l < 730.0 for l in lam)
    step = 350.0 / 4
    for j in range(4):
        d = (lam[(j + 1) % 4] - lam[j]) % 350.0
        assert abs(d - step) < 1e-3
    assert abs(lam[0] - (380.0 + 0.1 * 350.0)) < 1e-4


# ---------------------------------------------------------------- sellmeier.rs
def test_sellmeier_published_lines(oracle):
    g = oracle.lib().or_glass_n
    # sellmeier.rs:53-58
    assert abs(g(A.GLASS_BK7, 486.1) - 1.52238) < 5e-4
    assert abs(g(A.GLASS_BK7, 587.6) - 1.51680) < 5e-4
    assert abs(g(A.GLASS_BK7, 656.3) - 1.51432) < 5e-4
    # sellmeier.rs:60-78
    assert g(A.GLASS_BK7, 450) > g(A.GLASS_BK7, 650)
    assert g(A.GLASS_SF11, 450) > g(A.GLASS_SF11, 650)
    assert g(A.GLASS_SF11, 450) - g(A.GLASS_SF11, 650) > g(A.GLASS_BK7, 450) - g(A.GLASS_BK7, 650)
    assert abs(g(A.GLASS_WATER, 589.0) - 1.333) < 2e-3
    assert g(A.GLASS_WATER, 450) > g(A.GLASS_WATER, 650)


def test_sellmeier_matches_f64_closed_form(oracle):
    # The f64 formula evaluated independently in numpy, rounded to f32, must
    # equal the oracle bit for bit; f32 evaluation stays within 1e-4
    # (f32_budget.rs:27-40).
    g = oracle.lib().or_glass_n
    B = {0: (1.03961212, 0.231792344, 1.01046945), 1: (1.73759695, 0.313747346, 1.89878101)}
    Cc = {0: (0.00600069867, 0.0200179144, 103.560653), 1: (0.013188707, 0.0623068142, 155.23629)}
    for glass in (0, 1):
        for lam in np.linspace(380, 730, 141, dtype=np.float32):
            l2 = (float(lam) / 1000.0) ** 2
            s = 0.0
            for b, c in zip(B[glass], Cc[glass]):
                s = s + b * l2 / (l2 - c)
            n64 = np.float32(math.sqrt(1.0 + s))
            assert g(glass, float(lam)) == n64
            l2f = np.float32(np.float32(lam) / np.float32(1000.0)) ** 2
            n32 = np.sqrt(np.float32(1) + sum(np.float32(b) * l2f / (l2f - np.float32(c))
                                              for b, c in zip(B[glass], Cc[glass])))
            assert abs(float(n32) - float(n64)) < 1e-4


# ---------------------------------------------------------------- cie.rs
def test_d65_white_point(oracle):
    # cie.rs:151-158
    out = (C.c_float * 3)()
    oracle.lib().or_integrate_illuminant(A.ILLUM_D65, out)
    s = sum(out)
    assert abs(out[0] / s - 0.3127) < 2e-3 and abs(out[1] / s - 0.3290) < 2e-3


def test_cmf_in_band_and_ybar_peak(oracle):
    # cie.rs:160-166, data.rs:71-79
    c = (C.c_float * 3)()
    oracle.lib().or_cmf(555.0, c)
    assert c[0] >= 0 and c[1] > 0 and c[2] >= 0
    assert abs(c[1] - 1.0) < 0.05
    assert oracle.lib().or_illuminant(A.ILLUM_D65, 555.0) > 0


def test_table_interpolation_and_clamp(oracle):
    # Table::sample, cie.rs:47-61: linear between rows, clamped at both ends
    L = oracle.lib()
    rows = np.loadtxt(Path(oracle.DATA_DIR) / "cie_1931_2deg.csv", delimiter=",", dtype=np.float32)
    c = (C.c_float * 3)()
    L.or_cmf(380.0, c)
    assert list(c) == [float(x) for x in rows[0, 1:]]
    L.or_cmf(100.0, c)
    assert list(c) == [float(x) for x in rows[0, 1:]]
    L.or_cmf(900.0, c)
    assert list(c) == [float(x) for x in rows[-1, 1:]]
    L.or_cmf(552.5, c)
    i = (552 - 380) // 5
    mid = rows[i, 1:] + (rows[i + 1, 1:] - rows[i, 1:]) * np.float32(0.5)
    assert np.allclose(list(c), mid, rtol=0, atol=1e-7)


def test_xyz_to_srgb_white(oracle):
    # cie.rs:168-174
    out = (C.c_float * 3)()
    oracle.lib().or_xyz_to_linear_srgb(F3(0.9505, 1.0, 1.0890), out)
    assert all(abs(x - 1.0) < 1e-2 for x in out)


def test_tonemap_closed_form(oracle):
    # cie.rs:131-144 at known points: 0 -> 0, v=1 -> mapped 0.5 -> 188
    out = (C.c_uint8 * 3)()
    oracle.lib().or_tonemap_to_u8(F3(0.0, 1.0, 1000.0), 1.0, out)
    g = 1.055 * 0.5 ** (1 / 2.4) - 0.055
    assert list(out)[:2] == [0, int(g * 255 + 0.5)]


# ---------------------------------------------------------------- optics.rs
def _refract(L, d, n, n1, n2):
    out = F3()
    ok = L.or_refract(v(d), v(n), n1, n2, out)
    return np.array(out) if ok else None


def test_refract_matches_snell(oracle):
    # optics.rs:58-71
    L = oracle.lib()
    for deg in (5.0, 15.0, 30.0, 45.0, 60.0):
        ti = math.radians(deg)
        d = np.array([math.sin(ti), -math.cos(ti), 0.0])
        t = _refract(L, d, (0, 1, 0), 1.0, 1.5)
        tt = math.acos(max(-1, min(1, -t[1])))
        assert abs(tt - math.asin(math.sin(ti) / 1.5)) < 1e-5
        assert abs(t[2]) < 1e-6


def test_refract_forward_tangential_reversible(oracle):
    # optics.rs:77-118
    L = oracle.lib()
    ti = math.radians(40.0)
    d = np.array([math.sin(ti), -math.cos(ti), 0.0])
    t = _refract(L, d, (0, 1, 0), 1.0, 1.5)
    assert t @ d > 0 and t[1] < 0
    ti = math.radians(50.0)
    d = np.array([math.sin(ti), -math.cos(ti), 0.0])
    t = _refract(L, d, (0, 1, 0), 1.0, 1.5)
    assert abs(1.0 * abs(d[0]) - 1.5 * abs(t[0])) < 1e-5 and d[0] * t[0] > 0
    ti = math.radians(35.0)
    d = np.array([math.sin(ti), -math.cos(ti), 0.0], np.float32)
    t = _refract(L, d, (0, 1, 0), 1.0, 1.5)
    back = _refract(L, t, (0, 1, 0), 1.5, 1.0)
    assert np.allclose(back, d, atol=1e-5)


def test_tir_critical_angle_and_predicates(oracle):
    # optics.rs:120-131, 164-177
    L = oracle.lib()
    tc = math.asin(1 / 1.5)
    for ang, tir in ((tc + 0.05, True), (tc - 0.05, False)):
        d = (math.sin(ang), -math.cos(ang), 0.0)
        assert (_refract(L, d, (0, 1, 0), 1.5, 1.0) is None) == tir
    for deg in (10, 30, 41, 42, 50, 70):
        th = math.radians(deg)
        d = (math.sin(th), -math.cos(th), 0.0)
        r = L.or_fresnel_reflectance(abs(math.cos(th)), 1.5, 1.0)
        assert (_refract(L, d, (0, 1, 0), 1.5, 1.0) is not None) == (r < 1.0)


def test_fresnel_closed_forms(oracle):
    # optics.rs:133-158
    L = oracle.lib()
    assert abs(L.or_fresnel_reflectance(1.0, 1.0, 1.5) - ((1 - 1.5) / (1 + 1.5)) ** 2) < 1e-6
    assert L.or_fresnel_reflectance(1e-5, 1.0, 1.5) > 0.999
    prev = L.or_fresnel_reflectance(1.0, 1.0, 1.5)
    for s in range(1, 21):
        r = L.or_fresnel_reflectance(1.0 - s / 20 * 0.999, 1.0, 1.5)
        assert 0.0 <= r <= 1.0 and r >= prev - 1e-6
        prev = r


def test_reflect_is_mirror(oracle):
    # optics.rs:179-190
    L = oracle.lib()
    out = F3()
    d = np.array([0.4, -0.7, 0.3])
    d = d / np.linalg.norm(d)
    L.or_reflect(v(d), v((0, 1, 0)), out)
    assert np.allclose(list(out), [d[0], -d[1], d[2]], atol=1e-6)


# ---------------------------------------------------------------- geom.rs
def _sphere(L, c, r, o, d):
    h = (C.c_float * 8)()
    return np.array(h) if L.or_sphere_intersect(v(c), r, v(o), v(d), 1e-4, math.inf, h) else None


def _solid(L, planes, o, d):
    P = (A.Plane * len(planes))(*[A.Plane(F3(*n), dd) for n, dd in planes])
    h = (C.c_float * 8)()
    return np.array(h) if L.or_solid_intersect(P, len(planes), v(o), v(d), 1e-4, math.inf, h) else None


def test_geom_known_answers(oracle):
    # geom.rs:183-286
    L = oracle.lib()
    h = _sphere(L, (0, 0, -5), 1.0, (0, 0, 0), (0, 0, -1))
    assert abs(h[0] - 4.0) < 1e-4 and np.allclose(h[4:7], (0, 0, 1), atol=1e-4)
    assert _sphere(L, (5, 0, -5), 1.0, (0, 0, 0), (0, 0, -1)) is None
    cube = axis_box((-1, -1, -1), (1, 1, 1))
    h = _solid(L, cube, (0, 0, 5), (0, 0, -1))
    assert abs(h[0] - 4) < 1e-4 and np.allclose(h[4:7], (0, 0, 1)) and h[7] == 1.0
    assert _solid(L, cube, (2, 0, -5), (0, 0, 1)) is None
    assert _solid(L, cube, (-5, 2, 0), (1, 0, 0)) is None
    h = _solid(L, cube, (5, 0, 0), (-1, 0, 0))
    assert abs(h[0] - 4) < 1e-4 and np.allclose(h[4:7], (1, 0, 0))
    h = _solid(L, cube, (0, 0, 0), (0, 0, 1))
    assert abs(h[0] - 1) < 1e-4 and np.allclose(h[4:7], (0, 0, -1)) and h[7] == 0.0
    tri = [((math.cos(a), math.sin(a), 0.0), -1.0)
           for a in (math.pi / 2 + k * 2 * math.pi / 3 for k in range(3))]
    tri += [((0, 0, 1), -1.0), ((0, 0, -1), -1.0)]
    h = _solid(L, tri, (0, 5, 0), (0, -1, 0))
    assert abs(h[0] - 4) < 1e-3 and np.allclose(h[4:7], (0, 1, 0), atol=1e-3)
    assert _solid(L, tri, (5, 5, 0), (0, -1, 0)) is None
    assert len(wedge(30.0, 1.0, 2.0)) == 5


def test_scene_intersect_and_background(oracle):
    # scene.rs:85-120
    from oracle.oracle import scene_view
    L = oracle.lib()
    s = Scene()
    s.add_sphere((0, 0, -3), 1.0, Lambertian(0.5))
    s.add_sphere((0, 0, -6), 1.0, Lambertian(0.5))
    sv = scene_view(s)
    h = (C.c_float * 8)()
    m = C.c_uint32()
    assert L.or_scene_intersect(C.byref(sv), v((0, 0, 0)), v((0, 0, -1)), h, C.byref(m))
    assert abs(h[0] - 2.0) < 1e-3
    assert not L.or_scene_intersect(C.byref(sv), v((0, 0, 0)), v((0, 0, 1)), h, C.byref(m))
    env = Scene(background=0.5).env()
    assert L.or_background_radiance(C.byref(env), v((0, 1, 0))) == 0.5
    env = Scene(background=0.5, horizon=2.0).env()
    assert L.or_background_radiance(C.byref(env), v((0, 1, 0))) == 2.5
    assert L.or_background_radiance(C.byref(env), v((0, -1, 0))) == 0.5


# ---------------------------------------------------------------- camera.rs
def test_camera_known_answers(oracle):
    # camera.rs:76-110
    L = oracle.lib()
    cam = A.CameraBasis()
    L.or_camera_look_at(v((0, 0, 0)), v((0, 0, -1)), v((0, 1, 0)), 60.0, 1.0, C.byref(cam))
    d = F3()
    L.or_camera_primary_ray(C.byref(cam), 0.5, 0.5, d)
    assert -d[2] > 0.99
    for s, t in ((0, 0), (1, 1), (0.3, 0.7)):
        L.or_camera_primary_ray(C.byref(cam), s, t, d)
        assert abs(np.linalg.norm(list(d)) - 1) < 1e-5
    L.or_camera_look_at(v((0, 1, 5)), v((0, 0, 0)), v((0, 1, 0)), 50.0, 1.5, C.byref(cam))
    out = F3()
    for s, t in ((0.5, 0.5), (0.2, 0.7), (0.8, 0.35)):
        L.or_camera_primary_ray(C.byref(cam), s, t, d)
        p = np.array(cam.origin[:]) + np.array(list(d)) * 4.0
        assert L.or_camera_project(C.byref(cam), v(p), out)
        assert abs(out[0] - s) < 1e-3 and abs(out[1] - t) < 1e-3 and out[2] > 0


def test_host_look_at_matches_oracle_bitwise(oracle):
    # The library's host Camera::look_at (no device needed) equals the oracle's.
    L = oracle.lib()
    for args in (((0.5, -0.5, 12.0), (0.5, -0.5, 0.0), 42.0, 1.0),
                 ((0, 1, 5), (0, 0, 0), 50.0, 1.5), ((0.4, -0.2, 4.2), (0.4, -0.2, 0), 40.0, 16 / 9)):
        o, t, fov, asp = args
        ref = A.CameraBasis()
        L.or_camera_look_at(v(o), v(t), v((0, 1, 0)), fov, asp, C.byref(ref))
        got = Camera.look_at(o, t, (0, 1, 0), fov, asp).basis
        assert bytes(ref) == bytes(got)


# ---------------------------------------------------------------- material.rs
def test_material_known_answers(oracle):
    # material.rs:94-159
    L = oracle.lib()
    hit = (C.c_float * 8)(1.0, 0, 0, 0, 0, 1, 0, 1.0)
    lam = A.Material(0.8, 0, A.MAT_LAMBERTIAN, 0)
    d, w = F3(), C.c_float()
    for s in range(200):
        st = C.c_uint32(oracle.lib().or_rng_new(s, 7))
        L.or_material_scatter(C.byref(lam), v((0, -1, 0)), hit, 1.5, C.byref(st), d, C.byref(w))
        assert d[1] > 0 and abs(np.linalg.norm(list(d)) - 1) < 1e-5 and 0 <= w.value <= 1
    bk7 = A.Material(0.0, A.GLASS_BK7, A.MAT_DIELECTRIC, 0)
    th = math.radians(60.0)
    wo = (math.sin(th), -math.cos(th), 0.0)
    exit_hit = (C.c_float * 8)(1.0, 0, 0, 0, 0, 1, 0, 0.0)
    r = F3()
    L.or_reflect(v(wo), v((0, 1, 0)), r)
    for s in range(8):
        st = C.c_uint32(L.or_rng_new(s, 5))
        L.or_material_scatter(C.byref(bk7), v(wo), exit_hit, L.or_glass_n(0, 550.0), C.byref(st), d,
                              C.byref(w))
        assert np.allclose(list(d), list(r), atol=1e-5)
    th = math.radians(30.0)
    wo = (math.sin(th), -math.cos(th), 0.0)
    nh = L.or_glass_n(0, 550.0)
    ent, ext = [], []
    for s in range(32):
        for hh, acc in ((hit, ent), (exit_hit, ext)):
            st = C.c_uint32(L.or_rng_new(s, 99))
            L.or_material_scatter(C.byref(bk7), v(wo), hh, nh, C.byref(st), d, C.byref(w))
            acc.append(tuple(d))
    assert ent != ext


def test_hg_phase(oracle):
    # volume.rs:334-340, vol_photons.rs:877-904
    L = oracle.lib()
    assert L.or_phase_hg(0.6, 1.0) > L.or_phase_hg(0.6, -1.0)
    assert abs(L.or_phase_hg(0.0, 0.5) - L.or_phase_hg(0.0, -0.5)) < 1e-6
    for g, c in ((0.0, 0.5), (0.5, 1.0), (0.5, -1.0), (0.9, 0.0), (-0.3, 0.7)):
        ref = (1 - g * g) / (4 * math.pi * max(1 + g * g - 2 * g * c, 1e-6) ** 1.5)
        assert abs(L.or_phase_hg(g, c) - ref) < 1e-5 * max(1.0, ref)
