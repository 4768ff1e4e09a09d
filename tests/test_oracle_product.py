"""Pins the oracle's restatement of the floor-caustic product passes
(prism_on_plane.rs:16-185) on the CPU: the box blur against an independent
edge-padded convolution and its impulse closed form, the surface pass
against ray-plane geometry computed here from the camera basis, and the
composite tonemap against a float64 sRGB curve."""
import ctypes as C

import numpy as np
import pytest

from spectral_amd import scenes
from spectral_amd.core import Scene


def _blur_ref(img, r):
    """Edge-padded separable box filter in float64."""
    k = 2 * r + 1
    p = np.pad(img.astype(np.float64), ((0, 0), (r, r), (0, 0)), mode="edge")
    t = sum(p[:, d:d + img.shape[1]] for d in range(k)) / k
    p = np.pad(t, ((r, r), (0, 0), (0, 0)), mode="edge")
    return sum(p[d:d + img.shape[0]] for d in range(k)) / k


@pytest.mark.parametrize("r", [0, 1, 2, 5])
def test_box_blur_is_the_edge_padded_box(oracle, r):
    rng = np.random.default_rng(r)
    img = rng.exponential(1.0, (23, 31, 3)).astype(np.float32)
    got = oracle.box_blur(img, r)
    np.testing.assert_allclose(got, _blur_ref(img, r), rtol=2e-6, atol=1e-7)


def test_box_blur_impulse_closed_form(oracle):
    img = np.zeros((15, 15, 3), np.float32)
    img[7, 7] = (1.0, 2.0, 3.0)
    got = oracle.box_blur(img, 2)
    norm = np.float32(1.0) / np.float32(5)
    for c, v in enumerate((1.0, 2.0, 3.0)):
        want = np.zeros((15, 15), np.float32)
        want[5:10, 5:10] = np.float32(v) * norm * norm  # one non-zero tap per sum: exact
        assert np.array_equal(got[..., c], want)


def _rays(cam, w, h):
    b = cam.basis
    o, ll, hz, vt = (np.array(list(x), np.float64) for x in (b.origin, b.lower_left, b.horizontal, b.vertical))
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    s = (px + 0.5) / w
    t = 1.0 - (py + 0.5) / h
    d = ll + s[..., None] * hz + t[..., None] * vt - o
    return o, d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_floor_surface_without_a_solid_is_the_shaded_floor(oracle):
    w, h = 80, 60
    cam, scr = scenes.plane_camera(w, h), scenes.plane_floor_screen(32, 32)
    const = np.float32(0.5)
    floor = np.full((32, 32, 3), const, np.float32)
    sh = scenes.PLANE_SHADE
    surf, depth = oracle.floor_surface(Scene(), cam, scr, floor, 2.0, sh["ambient"], sh["background"], sh["tint"],
                                       w, h)
    o, d = _rays(cam, w, h)
    t = (-1.0 - o[1]) / d[..., 1]                      # the floor plane y = -1
    p = o + d * t[..., None]
    s, r = (p[..., 2] + 5.0) / 10.0, (p[..., 0] + 5.0) / 14.0
    inside = (t > 1e-3) & (s > 0.01) & (s < 0.99) & (r > 0.01) & (r < 0.99)
    outside = (t <= 0) | (s < -0.01) | (s > 1.01) | (r < -0.01) | (r > 1.01)
    assert inside.sum() > 100 and outside.sum() > 100
    want = np.array(sh["ambient"]) + 0.5 * 2.0
    np.testing.assert_allclose(surf[inside], np.broadcast_to(want, surf[inside].shape), rtol=1e-6)
    np.testing.assert_allclose(depth[inside], t[inside], rtol=1e-5)
    assert (surf[outside] == np.array(sh["background"], np.float32)).all()
    assert np.isinf(depth[outside]).all()


def test_floor_surface_glass_over_floor(oracle):
    w, h = 96, 72
    cam, scr = scenes.plane_camera(w, h), scenes.plane_floor_screen(16, 16)
    floor = np.zeros((16, 16, 3), np.float32)
    sh = scenes.PLANE_SHADE
    sc = scenes.wedge_scene()
    surf, depth = oracle.floor_surface(sc, cam, scr, floor, 1.0, sh["ambient"], sh["background"], sh["tint"], w, h)
    sv = oracle.scene_view(sc)
    o, d = _rays(cam, w, h)
    glass = 0
    for py in range(0, h, 3):
        for px in range(0, w, 3):
            dd = (C.c_float * 3)(*d[py, px])
            hit, mi = (C.c_float * 8)(), C.c_uint32(0)
            if not oracle.lib().or_scene_intersect(C.byref(sv), cam.basis.origin, dd, hit, C.byref(mi)):
                continue
            if not abs(hit[0] - depth[py, px]) < 1e-4 * hit[0]:
                continue  # the floor is nearer than the glass here
            glass += 1
            cosv = min(abs(float(np.dot(np.array(hit[4:7]), d[py, px]))), 1.0)
            rim = (1.0 - cosv) ** 3
            # the floor behind the glass shades as ambient (zero sums) or, off the floor, the background
            ok = any(np.allclose(surf[py, px], np.array(base) * 0.7 + np.array(sh["tint"]) +
                                 rim * np.array([0.6, 0.6, 0.7]), rtol=1e-4, atol=1e-5)
                     for base in (sh["ambient"], sh["background"]))
            assert ok, (px, py, surf[py, px], rim)
    assert glass > 20


def test_composite_is_the_srgb_tonemap_of_the_sum(oracle):
    rng = np.random.default_rng(7)
    surf = rng.uniform(0.0, 0.3, (9, 13, 3)).astype(np.float32)
    vol = rng.exponential(0.5, (9, 13, 3)).astype(np.float32)
    got = oracle.composite_srgb8(surf, vol, 1.7)
    M = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])
    lin = (surf.astype(np.float64) + vol) @ M.T
    v = np.maximum(lin * 1.7, 0.0)
    m = v / (1.0 + v)
    g = np.where(m <= 0.0031308, 12.92 * m, 1.055 * m ** (1.0 / 2.4) - 0.055)
    want = np.floor(np.clip(g, 0.0, 1.0) * 255.0 + 0.5)
    assert np.abs(got.astype(np.int32) - want).max() <= 1
