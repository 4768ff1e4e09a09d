"""Independent pins of the oracle functions that no reference golden value
covers: the point estimator's camera connection (volume.rs:130-148), the
lighttrace Screen's bilinear splat (lighttrace.rs:65-91) and the surface rim
colour (prism_dsotm.rs:59-74).  Each is checked against numbers computed here
without the oracle -- a float32 numpy restatement on a handful of photons, a
closed-form integral, an analytic sphere -- chosen so that a plausible slip
(a dropped 1/d^2, the wrong sign of cos theta, a flipped screen row, the rim
exponent) fails one of them.  CPU only."""
import ctypes as C
import math
from pathlib import Path

import numpy as np

from spectral_amd import _abi as A
from spectral_amd import scenes
from spectral_amd.core import Camera, Lambertian, Scene, beam, opts, screen

F = np.float32
DATA = Path(__file__).resolve().parents[1] / "spectral_amd" / "data"
M32 = 0xFFFFFFFF


# ---- independent restatements (Python ints / numpy float32) ---------------
def _pcg32(x):
    state = (x * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


class _Rng:
    """PathRng (rng.rs:19-34); its draws are pinned by tests/golden/rng_vectors.json."""

    def __init__(self, key, seed):
        self.st = _pcg32(key ^ _pcg32(seed))

    def f32(self):
        self.st = (self.st + 1) & M32
        return F(_pcg32(self.st) >> 8) * F(1.0 / 16777216.0)


def _table(name, cols):
    rows = [ln.split(",") for ln in (DATA / name).read_text().splitlines() if ln.strip()]
    return np.array([[float(v) for v in r[1:1 + cols]] for r in rows], np.float32)


CMF = _table("cie_1931_2deg.csv", 3)
D65 = _table("illuminant_d65.csv", 1)[:, 0]


def _sample(tab, nm):
    """Linear interpolation at 5 nm from 380 nm, clamped (the CIE tables)."""
    f = min(max((F(nm) - F(380.0)) / F(5.0), F(0.0)), F(len(tab) - 1))
    i = int(np.floor(f))
    if i + 1 >= len(tab):
        return tab[-1]
    fr = f - F(i)
    return tab[i] + (tab[i + 1] - tab[i]) * fr


def _v(x):
    return np.array(x, np.float32)


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _length(a):
    return F(np.sqrt(_dot(a, a)))


def _basis(cam):
    b = cam.basis
    return _v(b.origin[:]), _v(b.lower_left[:]), _v(b.horizontal[:]), _v(b.vertical[:])


def _project(cam, p):
    """Camera::project (camera.rs:50-69): (s, t) or None."""
    o, ll, hz, vt = _basis(cam)
    u, v = hz / _length(hz), vt / _length(vt)
    vw, vh = _length(hz), _length(vt)
    w = o - hz * F(0.5) - vt * F(0.5) - ll
    d = p - o
    c = _dot(d, w)
    if c >= F(-1e-6):
        return None
    s = F(0.5) + _dot(d, u) / (-c * vw)
    t = F(0.5) + _dot(d, v) / (-c * vh)
    if F(0) <= s < F(1) and F(0) <= t < F(1):
        return s, t
    return None


def _hg(g, c):
    g, c = F(g), F(c)
    g2 = g * g
    den = F(np.power(max(F(1) + g2 - F(2) * g * c, F(1e-6)), F(1.5)))
    return (F(1) - g2) / (F(4) * F(math.pi) * den)


def _comb(u):
    """The stratified hero comb of spectrum.rs:20-32: 380 + rem_euclid(u 350 + k 87.5, 350)."""
    return [F(380.0) + F(np.fmod(F(u) * F(350.0) + F(k) * F(87.5), F(350.0))) for k in range(4)]


def _point_film(scene_empty_beam, cam, w, h, n, seed, vp, power_of, hero=False):
    """render_volumetric_scene_point (volume.rs:118-150) for an empty scene:
    one escaped segment of length max_dist per photon, four samples.  hero:
    the photon carries the 4-wavelength comb, spectrum sum_k cmf(l_k) s_k / 4
    and power 1 (no dielectric, so the comb never collapses)."""
    bm = scene_empty_beam
    corner, bu, bv, bd = _v(bm.corner[:]), _v(bm.u[:]), _v(bm.v[:]), _v(bm.dir[:])
    cam_o = _basis(cam)[0]
    img = np.zeros((h, w, 3), np.float32)
    hits = []
    for i in range(n):
        r = _Rng(i, seed)
        a, b = r.f32(), r.f32()
        o = corner + bu * a + bv * b
        u = r.f32()
        lam = F(380.0) + F(350.0) * u
        if hero:
            cmf = np.zeros(3, np.float32)
            for lk in _comb(u):
                cmf = cmf + _sample(CMF, lk) * power_of(lk) * F(0.25)
            power = F(1.0)
        else:
            cmf = _sample(CMF, lam)
            power = power_of(lam)
        seg = F(vp.max_dist)
        for _ in range(4):
            p = o + bd * (r.f32() * seg)
            st = _project(cam, p)
            if st is None:
                continue
            px = min(int(st[0] * F(w)), w - 1)
            py = min(int((F(1) - st[1]) * F(h)), h - 1)
            to = cam_o - p
            d = max(_length(to), F(1e-3))
            cos_t = _dot(bd, to / d)
            contrib = (power * F(vp.sigma_s) * _hg(vp.g, cos_t) * F(np.exp(-F(vp.sigma_t) * d)) / (d * d)
                       * seg / F(4))
            img[py, px] += cmf * contrib
            hits.append((py, px))
    return img, hits


# ---- the point estimator ---------------------------------------------------
def test_point_estimator_deposits_brute_force(oracle):
    # 8 photons of a 0.1-wide beam through an empty scene, two illuminants:
    # every deposit's pixel and value, recomputed here
    sc = Scene(background=0.0)
    bm = beam((-1.2, -0.3, -0.05), (0.0, 0.0, 0.1), (0.0, 0.1, 0.0), (1.0, 0.0, 0.0))
    cam = Camera.look_at((0.3, 0.5, 4.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 16.0 / 9.0)
    vp = A.VolParams(0.3, 0.2, 0.4, 2.5)
    w, h, n, seed = 96, 54, 8, 77
    scale = F(1.0 / (np.float64(D65).mean()))
    for ill, power_of in ((A.ILLUM_E, lambda lam: F(1.0)),
                          (A.ILLUM_D65, lambda lam: _sample(D65[:, None], lam)[0] * scale)):
        want, hits = _point_film(bm, cam, w, h, n, seed, vp, power_of)
        got, st = oracle.volume(sc, cam, bm, w, h, n, 0, vp, None, seed, A.EST_POINT, opts(ill), threads=1)
        assert len(hits) >= 16 and st.splats == len(hits)
        assert np.array_equal(got[..., 1] != 0, want[..., 1] != 0)
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)


def _segment_in_frame(oracle, cam, p0, d, L):
    L_ = oracle.lib()
    for x in np.linspace(0.0, L, 257):
        p = (C.c_float * 3)(*(np.asarray(p0) + x * np.asarray(d)))
        out = (C.c_float * 3)()
        if not L_.or_camera_project(C.byref(cam.basis), p, out):
            return False
    return True


def _expected_y(cam_o, p0, d, L, sigma_s, sigma_t, g, ybar_mean):
    """n-photon expectation per photon: sigma_s * <ybar> * integral over the
    segment of HG(g, cos) * exp(-sigma_t d) / d^2 (the four samples of
    length L/4 each sum to one integral), by Simpson's rule in float64."""
    x = np.linspace(0.0, L, 40001)
    p = np.asarray(p0)[None, :] + x[:, None] * np.asarray(d)[None, :]
    to = np.asarray(cam_o)[None, :] - p
    dist = np.linalg.norm(to, axis=1)
    cos = (to @ np.asarray(d)) / dist
    hg = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos) ** 1.5)
    f = hg * np.exp(-sigma_t * dist) / dist ** 2
    hstep = x[1] - x[0]
    simpson = hstep / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    return sigma_s * ybar_mean * simpson


def _ybar_mean():
    """(1/350) * integral of the piecewise-linear ybar over 380..730 nm."""
    y = CMF[:, 1].astype(np.float64)
    return 5.0 * (y.sum() - 0.5 * (y[0] + y[-1])) / 350.0


def test_point_estimator_energy_closed_form(oracle):
    # isotropic haze (g = 0), no extinction: total Y = n sigma_s <ybar> / (4 pi)
    # * integral dl / d(l)^2 = (1 / h) (atan((L - l0) / h) + atan(l0 / h))
    sc = Scene(background=0.0)
    bm = beam((-1.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    cam = Camera.look_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 16.0 / 9.0)
    vp = A.VolParams(0.25, 0.0, 0.0, 2.0)
    assert _segment_in_frame(oracle, cam, (-1, 0, 0), (1, 0, 0), 2.0)
    n = 1 << 20
    img, _ = oracle.volume(sc, cam, bm, 160, 90, n, 0, vp, None, 3, A.EST_POINT, opts(A.ILLUM_E), threads=8)
    integral = (1 / 5.0) * (math.atan(1 / 5.0) + math.atan(1 / 5.0))
    want = n * 0.25 * _ybar_mean() / (4 * math.pi) * integral
    got = float(img[..., 1].astype(np.float64).sum())
    print(f"closed form {want:.6g}, oracle {got:.6g}")
    assert abs(got / want - 1) < 6e-3


def test_point_estimator_phase_sign_front_and_behind(oracle):
    # forward scattering (g = 0.8) seen by a camera the beam heads toward
    # (cos theta near +1) and by its mirror image behind the beam's start:
    # each total matches its quadrature, and they differ ~40x, so a flipped
    # cos theta fails both
    sc = Scene(background=0.0)
    bm = beam((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    vp = A.VolParams(0.5, 0.1, 0.8, 2.0)
    n = 1 << 19
    totals = []
    for cam_o in ((4.0, 0.8, 0.0), (-2.0, 0.8, 0.0)):
        cam = Camera.look_at(cam_o, (1.0, 0.0, 0.0), (0, 1, 0), 60.0, 1.0)
        assert _segment_in_frame(oracle, cam, (0, 0, 0), (1, 0, 0), 2.0)
        img, _ = oracle.volume(sc, cam, bm, 128, 128, n, 0, vp, None, 9, A.EST_POINT, opts(A.ILLUM_E),
                               threads=8)
        got = float(img[..., 1].astype(np.float64).sum())
        want = n * _expected_y(cam.origin(), (0, 0, 0), (1, 0, 0), 2.0, 0.5, 0.1, 0.8, _ybar_mean())
        print(f"camera at {cam_o}: quadrature {want:.6g}, oracle {got:.6g}")
        assert abs(got / want - 1) < 1e-2
        totals.append(got)
    assert totals[0] > 20 * totals[1]


# ---- the Screen splat --------------------------------------------------------
def _screen_case(oracle, s, r, n=8, seed=5):
    """A zero-width beam through an empty scene straight onto a 4 x 2 screen
    at known (s, r): the film holds the four bilinear texels."""
    W, H = 16, 8
    scr = screen((0.0, 0.0, 0.0), (4.0, 0.0, 0.0), (0.0, 2.0, 0.0), W, H)
    bm = beam((4.0 * s, 2.0 * r, 1.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    got, st = oracle.lighttrace(Scene(background=0.0), scr, bm, n, 0, seed, opts(A.ILLUM_E), threads=1)
    lams = []
    for i in range(n):
        rg = _Rng(i, seed)
        rg.f32(), rg.f32()
        lams.append(F(380.0) + F(350.0) * rg.f32())
    return got, lams, W, H


def test_screen_splat_bilinear_texels_row_zero_at_top(oracle):
    # s = 27/64, r = 55/64: fx = s W - 1/2 = 6.25, fy = (1 - r) H - 1/2 = 0.625
    # -> texels (6, 0), (7, 0), (6, 1), (7, 1): the top rows, since r runs up
    got, lams, W, H = _screen_case(oracle, 27 / 64, 55 / 64)
    wts = {(0, 6): 0.75 * 0.375, (0, 7): 0.25 * 0.375, (1, 6): 0.75 * 0.625, (1, 7): 0.25 * 0.625}
    want = np.zeros((H, W, 3), np.float32)
    for lam in lams:
        c = _sample(CMF, lam)
        for (y, x), wk in wts.items():
            want[y, x] += c * F(1.0) * F(wk)
    assert np.array_equal(got != 0, want != 0)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    ysum = float(np.sum([_sample(CMF, l)[1] for l in lams], dtype=np.float64))
    assert abs(got[..., 1].astype(np.float64).sum() / ysum - 1) < 1e-6


def test_screen_splat_clamps_at_the_corner_keeping_energy(oracle):
    # s = r = 1/64: fx clamps to 0 and y1 = y0 = H - 1, so all the weight
    # lands on the bottom-left texel
    got, lams, W, H = _screen_case(oracle, 1 / 64, 1 / 64)
    nz = np.argwhere(got[..., 1] != 0)
    assert nz.tolist() == [[H - 1, 0]]
    ysum = float(np.sum([_sample(CMF, l)[1] for l in lams], dtype=np.float64))
    assert abs(float(got[H - 1, 0, 1]) / ysum - 1) < 1e-6


# ---- the surface rim -----------------------------------------------------------
def test_surface_rim_against_an_analytic_sphere(oracle):
    # (1 - |n . d|)^2 into (0.015 + 0.45 rim, 0.02 + 0.5 rim, 0.04 + 0.6 rim)
    # and depth = t, with n and t from the ray-sphere closed form
    sc = Scene(background=0.0)
    c0, rad = np.array([0.2, -0.1, 0.0]), 1.0
    sc.add_sphere(tuple(c0), rad, Lambertian(0.5))
    w = h = 64
    cam = Camera.look_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 1.0)
    rim, depth = oracle.surface_rim(sc, cam, w, h)
    o, ll, hz, vt = (x.astype(np.float64) for x in _basis(cam))
    checked = 0
    for py in range(h):
        for px in range(w):
            s, t = (px + 0.5) / w, 1.0 - (py + 0.5) / h
            d = ll + s * hz + t * vt - o
            d /= np.linalg.norm(d)
            oc = o - c0
            b = oc @ d
            disc = b * b - (oc @ oc - rad * rad)
            if abs(disc) < 1e-3:
                continue  # silhouette: float32 rounding may flip the hit
            if disc < 0:
                assert depth[py, px] == np.inf and not rim[py, px].any()
                continue
            tt = -b - math.sqrt(disc)
            nrm = (o + tt * d - c0) / rad
            q = (1.0 - min(abs(nrm @ d), 1.0)) ** 2
            # float32 rounding of the hit distance grows as 1 / sqrt(disc)
            # toward the silhouette; a wrong exponent or constant is off by > 1e-2
            np.testing.assert_allclose(rim[py, px], [0.015 + 0.45 * q, 0.02 + 0.5 * q, 0.04 + 0.6 * q],
                                       rtol=0, atol=1e-5 + 1e-5 / math.sqrt(disc))
            assert abs(depth[py, px] - tt) < 1e-4 * tt
            checked += 1
    assert checked > 0.2 * w * h


# ---- the 4-wavelength hero comb for photons (an extension) ---------------------
def test_hero_comb_point_deposits_brute_force(oracle):
    # the comb's spectrum on every deposit of 8 photons through an empty scene
    sc = Scene(background=0.0)
    bm = beam((-1.2, -0.3, -0.05), (0.0, 0.0, 0.1), (0.0, 0.1, 0.0), (1.0, 0.0, 0.0))
    cam = Camera.look_at((0.3, 0.5, 4.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 16.0 / 9.0)
    vp = A.VolParams(0.3, 0.2, 0.4, 2.5)
    w, h, n, seed = 96, 54, 8, 77
    scale = F(1.0 / (np.float64(D65).mean()))
    for ill, power_of in ((A.ILLUM_E, lambda lam: F(1.0)),
                          (A.ILLUM_D65, lambda lam: _sample(D65[:, None], lam)[0] * scale)):
        want, hits = _point_film(bm, cam, w, h, n, seed, vp, power_of, hero=True)
        got, st = oracle.volume(sc, cam, bm, w, h, n, 0, vp, None, seed, A.EST_POINT, opts(ill, hero_lanes=4),
                                threads=1)
        assert st.splats == len(hits) >= 16
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)


def test_hero_comb_energy_closed_form_per_channel(oracle):
    # E[sum_k cmf(l_k) / 4] over the stratified comb = the band mean of each
    # matching function, so every channel of an isotropic, unextinguished
    # segment totals n sigma_s <cmf> / (4 pi) * integral dl / d^2
    sc = Scene(background=0.0)
    bm = beam((-1.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    cam = Camera.look_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 16.0 / 9.0)
    vp = A.VolParams(0.25, 0.0, 0.0, 2.0)
    n = 1 << 18
    img, _ = oracle.volume(sc, cam, bm, 160, 90, n, 0, vp, None, 3, A.EST_POINT, opts(A.ILLUM_E, hero_lanes=4),
                           threads=8)
    integral = (1 / 5.0) * (math.atan(1 / 5.0) + math.atan(1 / 5.0))
    c = CMF.astype(np.float64)
    mean = 5.0 * (c.sum(0) - 0.5 * (c[0] + c[-1])) / 350.0
    got = img.reshape(-1, 3).astype(np.float64).sum(0)
    want = n * 0.25 * mean / (4 * math.pi) * integral
    print("hero comb closed form", want, "oracle", got)
    assert (np.abs(got / want - 1) < 3e-3).all()


def test_hero_comb_collapses_at_the_first_dielectric(oracle):
    # seen past the prism only (the incoming beam out of frame: its first
    # segment deposits nothing), the comb has collapsed onto the hero lane,
    # which is the one-wavelength photon: the films agree to rounding
    sc, bm = scenes.prism_scene(), scenes.prism_beam()
    cam = Camera.look_at((1.9, 0.9, 3.0), (1.9, 0.9, 0.0), (0, 1, 0), 40.0, 1.0)
    vp = A.VolParams(**scenes.HAZE)
    first, _ = oracle.volume(sc, cam, bm, 96, 96, 1 << 14, 0, vp, None, 4, A.EST_POINT,
                             opts(A.ILLUM_D65, max_depth=1), threads=8)
    assert not first.any()
    one, _ = oracle.volume(sc, cam, bm, 96, 96, 1 << 14, 0, vp, None, 4, A.EST_POINT, opts(A.ILLUM_D65), threads=8)
    comb, _ = oracle.volume(sc, cam, bm, 96, 96, 1 << 14, 0, vp, None, 4, A.EST_POINT,
                            opts(A.ILLUM_D65, hero_lanes=4), threads=8)
    assert one[..., 1].sum() > 0
    err = np.abs(comb.astype(np.float64) - one).sum() / np.abs(one.astype(np.float64)).sum()
    print("hero comb past the prism vs one wavelength: rel L1", err)
    assert err < 1e-5
    # and in front of the prism the comb differs (four wavelengths per photon)
    cam2 = scenes.prism_camera(96, 54)
    a, _ = oracle.volume(sc, cam2, bm, 96, 54, 1 << 14, 0, vp, None, 4, A.EST_POINT, opts(A.ILLUM_D65), threads=8)
    b, _ = oracle.volume(sc, cam2, bm, 96, 54, 1 << 14, 0, vp, None, 4, A.EST_POINT,
                         opts(A.ILLUM_D65, hero_lanes=4), threads=8)
    assert np.abs(a - b).sum() > 1e-3 * np.abs(a).sum()
