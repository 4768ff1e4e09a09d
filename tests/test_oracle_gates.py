"""The reference's physics acceptance gates restated over the oracle (CPU):
scene A's prism dispersion (spectral-core/tests/scene_a_prism.rs), the
Rayleigh sky and sunset colours (tests/rayleigh.rs) and the three spectral
versus RGB transport witnesses (tests/spectral_vs_rgb.rs).  Each gate uses
the oracle's CIE tables, illuminants, Sellmeier indices, refraction and (for
scene A) its camera tracer with the horizon sky."""
import ctypes as C
import math

import numpy as np

from spectral_amd import _abi as A
from spectral_amd import scenes
from spectral_amd.core import opts

F = np.float32


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _cmf(L, nm):
    out = (C.c_float * 3)()
    L.or_cmf(float(nm), out)
    return F(out[0]), F(out[1]), F(out[2])


def _nms():
    """380, 385, ..., 730 nm (the f32 loop of spectral_vs_rgb.rs:17-25)."""
    out, nm = [], F(380.0)
    while nm <= F(730.0):
        out.append(nm)
        nm = F(nm + F(5.0))
    return out


def _chromaticity(xyz):
    s = xyz[0] + xyz[1] + xyz[2]
    return xyz[0] / s, xyz[1] / s


def _srgb(L, xyz):
    out = (C.c_float * 3)()
    L.or_xyz_to_linear_srgb(_f3(xyz), out)
    return [F(out[0]), F(out[1]), F(out[2])]


# ---- scene A (scene_a_prism.rs) ---------------------------------------------
def _deviation(L, d, n_vec, lam):
    """Angle between d and its refraction entering BK7 (scene_a_prism.rs:18-23)."""
    n = L.or_glass_n(A.GLASS_BK7, float(lam))
    t = (C.c_float * 3)()
    assert L.or_refract(_f3(d), _f3(n_vec), 1.0, n, t)
    d64, t64 = np.array(d, np.float64), np.array(t[:], np.float64)
    return math.acos(min(1.0, d64 @ t64 / (np.linalg.norm(d64) * np.linalg.norm(t64)))), n


def test_scene_a_blue_deviates_more_than_red(oracle):
    # scene_a_prism.rs:25-32: 30 deg incidence on a BK7 interface
    L = oracle.lib()
    d = np.array([0.5, -0.8660254, 0.0], np.float32)
    d = d / F(np.sqrt(d @ d))
    blue, _ = _deviation(L, d, (0, 1, 0), 450.0)
    red, _ = _deviation(L, d, (0, 1, 0), 650.0)
    assert blue > red


def test_scene_a_angular_spread_matches_snell(oracle):
    # scene_a_prism.rs:34-48: deviation theta_i - asin(sin theta_i / n) within 3 %
    L = oracle.lib()
    ti = math.radians(30.0)
    d = np.array([math.sin(ti), -math.cos(ti), 0.0], np.float32)
    for lam in (440.0, 520.0, 600.0, 680.0):
        meas, n = _deviation(L, d, (0, 1, 0), lam)
        analytic = ti - math.asin(math.sin(ti) / n)
        assert abs(meas - analytic) / analytic < 0.03, lam


def _chroma_spread(mean):
    """scene_a_prism.rs:52-64: largest chromaticity distance from D65's white point."""
    px = mean.reshape(-1, 3).astype(np.float32)
    s = px.sum(1)
    ok = s > F(1e-5)
    x, y = px[ok, 0] / s[ok], px[ok, 1] / s[ok]
    return float(np.sqrt((x - F(0.3127)) ** 2 + (y - F(0.3290)) ** 2).max())


def test_scene_a_dispersive_prism_colors_horizon_more_than_diffuse(oracle):
    # scene_a_prism.rs:66-90: 128x128, 256 spp, seed 11, horizon sky 4.0 over
    # a 0.3 background; SF11 prism against a 0.8 Lambertian control
    cam = scenes.scene_a_camera()
    spread = []
    for disp in (True, False):
        s, st = oracle.camera_accumulate(scenes.scene_a_scene(disp), cam, 128, 128, 256, 0, 11,
                                         opts(A.ILLUM_D65), threads=8)
        assert st.paths == 128 * 128 * 256
        spread.append(_chroma_spread(s / F(256)))
    cg, cd = spread
    print(f"scene A chroma spread: glass {cg:.4f}, diffuse {cd:.4f}")
    assert cg > 0.05 and cg > 2.0 * cd


def test_scene_a_horizon_lights_the_upper_half(oracle):
    # the sky above the horizon (scene.rs:44-50) makes rows looking up brighter
    # than rows looking down wherever the prism does not cover the pixel
    cam = scenes.scene_a_camera()
    s, _ = oracle.camera_accumulate(scenes.scene_a_scene(False), cam, 64, 64, 16, 0, 3,
                                    opts(A.ILLUM_D65), threads=8)
    y = s[..., 1] / 16
    assert y[0, 0] > 10 * y[-1, 0] > 0


# ---- Rayleigh sky (rayleigh.rs) ---------------------------------------------------
def _integrate(L, weight):
    """rayleigh.rs:9-23: sum of D65 * weight * cmf * 5 nm in f32."""
    x = y = z = F(0.0)
    for nm in _nms():
        s = F(L.or_illuminant(A.ILLUM_D65, float(nm))) * F(weight(nm))
        xb, yb, zb = _cmf(L, nm)
        x = F(x + s * xb * F(5.0))
        y = F(y + s * yb * F(5.0))
        z = F(z + s * zb * F(5.0))
    return [x, y, z]


def _rayleigh(nm):
    return F(F(550.0) / F(nm)) ** 4


def test_rayleigh_scattered_sky_is_blue(oracle):
    # rayleigh.rs:29-40
    L = oracle.lib()
    xyz = _integrate(L, _rayleigh)
    x, y = _chromaticity(xyz)
    rgb = _srgb(L, xyz)
    assert x < 0.28 and y < 0.33
    assert rgb[2] > rgb[0]


def test_rayleigh_sunset_is_red(oracle):
    # rayleigh.rs:42-50: transmission exp(-1.5 * (550/nm)^4)
    L = oracle.lib()
    xyz = _integrate(L, lambda nm: F(np.exp(F(-1.5) * _rayleigh(nm))))
    x, _ = _chromaticity(xyz)
    rgb = _srgb(L, xyz)
    assert x > 0.40
    assert rgb[0] > rgb[2]


# ---- spectral versus RGB transport (spectral_vs_rgb.rs) ------------------------------
def _spectral_shade(L, nms, light, rho):
    xyz = [F(0.0)] * 3
    for i, nm in enumerate(nms):
        s = F(light[i] * rho[i])
        c = _cmf(L, nm)
        xyz = [F(xyz[k] + s * c[k] * F(5.0)) for k in range(3)]
    return xyz


def _albedo_xyz(L, nms, rho):
    num, den = [F(0.0)] * 3, [F(0.0)] * 3
    for i, nm in enumerate(nms):
        c = _cmf(L, nm)
        num = [F(num[k] + F(rho[i]) * c[k]) for k in range(3)]
        den = [F(den[k] + c[k]) for k in range(3)]
    return [num[k] / den[k] for k in range(3)]


def _rgb_shade(L, nms, light, rho):
    lx = [F(0.0)] * 3
    for i, nm in enumerate(nms):
        c = _cmf(L, nm)
        lx = [F(lx[k] + F(light[i]) * c[k] * F(5.0)) for k in range(3)]
    a = _albedo_xyz(L, nms, rho)
    return [lx[k] * a[k] for k in range(3)]


def _light(L, ill, nms):
    return [F(L.or_illuminant(ill, float(nm))) for nm in nms]


def test_spectral_vs_rgb_achromatic_surface_agrees(oracle):
    # spectral_vs_rgb.rs:84-101: flat 0.6 grey under D65, per channel within 1e-3
    L, nms = oracle.lib(), _nms()
    light, rho = _light(L, A.ILLUM_D65, nms), [F(0.6)] * len(nms)
    g, r = _spectral_shade(L, nms, light, rho), _rgb_shade(L, nms, light, rho)
    for k in range(3):
        assert abs(g[k] - r[k]) / max(g[k], F(1e-6)) < 1e-3


def _metameric_black(L, nms):
    """spectral_vs_rgb.rs:109-150: a 3-cycle wiggle minus its CMF-span part."""
    dot = lambda a, b: F(np.sum(np.asarray(a, np.float32) * np.asarray(b, np.float32), dtype=np.float32))
    cols = [[_cmf(L, nm)[k] for nm in nms] for k in range(3)]
    basis = []
    for v in cols:
        u = np.array(v, np.float32)
        for e in basis:
            u = u - dot(u, e) * e
        norm = F(np.sqrt(dot(u, u)))
        if norm > 1e-6:
            basis.append(u / norm)
    tau = F(2.0 * math.pi)
    w = np.array([np.sin((F(nm) - F(380.0)) / F(350.0) * F(3.0) * tau) for nm in nms], np.float32)
    for e in basis:
        w = w - dot(w, e) * e
    return w


def test_spectral_vs_rgb_metamers_separated_under_a(oracle):
    # spectral_vs_rgb.rs:159-197: equal sensor albedo, chromaticity gap > 5e-3 under A
    L, nms = oracle.lib(), _nms()
    delta = _metameric_black(L, nms)
    alpha = F(0.35) / max(F(np.abs(delta).max()), F(1e-6))
    rho1 = [F(0.5)] * len(nms)
    rho2 = [F(F(0.5) + alpha * delta[i]) for i in range(len(nms))]
    assert all(0.0 <= r <= 1.0 for r in rho2)
    a1, a2 = _albedo_xyz(L, nms, rho1), _albedo_xyz(L, nms, rho2)
    assert all(abs(a1[k] - a2[k]) < 1e-4 for k in range(3))
    la = _light(L, A.ILLUM_A, nms)
    (x1, y1), (x2, y2) = _chromaticity(_spectral_shade(L, nms, la, rho1)), \
        _chromaticity(_spectral_shade(L, nms, la, rho2))
    assert math.hypot(x1 - x2, y1 - y2) > 5e-3


def test_spectral_vs_rgb_narrowband_light_on_notch(oracle):
    # spectral_vs_rgb.rs:221-254: 589 nm light (8 nm) on a 40 nm notch
    L, nms = oracle.lib(), _nms()
    gauss = lambda c, wd: [F(np.exp(-(((F(nm) - F(c)) / F(wd)) ** 2))) for nm in nms]
    light, notch = gauss(589.0, 8.0), gauss(589.0, 40.0)
    rho = [F(1.0) - n for n in notch]
    g, r = _spectral_shade(L, nms, light, rho), _rgb_shade(L, nms, light, rho)
    gw = _spectral_shade(L, nms, light, [F(1.0)] * len(nms))
    assert g[1] < 0.05 * gw[1]
    assert r[1] > 0.5 * gw[1]
