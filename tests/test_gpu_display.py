"""The display side of the path on the device against the oracle: the sRGB8
resolve (xyz_to_linear_srgb + tonemap_to_u8, cie.rs:121-144), the viewer's
progressive display pass (spectral-viewer main.rs:131-174: Gaussian gather,
composite with the surface rim, exposure, Reinhard, gamma 1/2.2, unorm8)
and its exposure quantile (main.rs:614-644), on a real volume render plus
its rim layer.  Bytes: within one step (expf / powf last bits differ
between ROCm and glibc), and all but a few equal; the quantile: bit-exact
(same fp32 luminance, exact selection)."""
import ctypes as C

import numpy as np
import pytest

import spectral_amd as S
from spectral_amd import _abi as A
from spectral_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def render(ctx):
    """A 160x90 prism volume render (2^18 photons) and its surface rim."""
    w, h, n = 160, 90, 1 << 18
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera(w, h)
    vp = A.VolParams(**scenes.HAZE)
    vol = S.render_volume(ctx, sc, cam, bm, w, h, n, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist, None, 0)
    rim, depth = S.surface_rim(ctx, sc, cam, w, h)
    depth.free()
    yield vol, rim, n
    vol.free()
    rim.free()


def _close_bytes(a, b, what):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    frac = (d > 0).mean()
    print(f"{what}: max byte diff {d.max()}, differing {frac:.2e}")
    assert d.max() <= 1 and frac < 1e-3, what


def test_resolve_srgb8_matches_the_oracle(ctx, oracle, render):
    vol, _, n = render
    film = vol.image()
    scale, exposure = 1.0 / n * 4.0e6, 0.8
    got = S.resolve_srgb8(ctx, vol, scale, exposure)
    h, w = film.shape[:2]
    want = np.empty((h, w, 3), np.uint8)
    lin, out = (C.c_float * 3)(), (C.c_uint8 * 3)()
    for y in range(h):
        for x in range(w):
            xyz = (C.c_float * 3)(*(film[y, x].astype(np.float64) * np.float32(scale)).astype(np.float32))
            oracle.lib().or_xyz_to_linear_srgb(xyz, lin)
            oracle.lib().or_tonemap_to_u8(lin, exposure, out)
            want[y, x] = list(out)
    _close_bytes(got, want, "resolve_srgb8")
    assert got.max() > 0


@pytest.mark.parametrize("kradius", [0.0, 2.0, 3.5])
@pytest.mark.parametrize("with_surface", [False, True])
def test_display_resolve_matches_the_oracle(ctx, oracle, render, kradius, with_surface):
    vol, rim, n = render
    surf = rim if with_surface else None
    total, target = n, 4.0e6
    q = S.display_quantile(ctx, vol, surf, total, target, 0.997)
    exposure = 1.0 / q
    got = S.display_resolve(ctx, vol, surf, total, target, exposure, kradius)
    want = oracle.display_resolve(vol.image(), rim.image() if with_surface else None, total, target, exposure,
                                  kradius)
    _close_bytes(got, want, f"display kradius={kradius} surface={with_surface}")
    assert (got[..., 3] == 255).all()
    assert got[..., :3].max() > 0


@pytest.mark.parametrize("q", [0.0, 0.5, 0.995, 0.997, 1.0])
def test_display_quantile_is_exact(ctx, oracle, render, q):
    vol, rim, n = render
    for surf in (None, rim):
        got = S.display_quantile(ctx, vol, surf, n, 4.0e6, q)
        want = oracle.display_quantile(vol.image(), None if surf is None else surf.image(), n, 4.0e6, q)
        assert got == want, (q, surf is not None, got, want)


def test_display_quantile_full_frame_order_statistic(ctx):
    # 1920x1080 with ties, zeros and a heavy tail: the device's radix select
    # against numpy's sort at the same fp32 luminance
    rng = np.random.default_rng(5)
    w, h = 1920, 1080
    img = rng.pareto(1.5, (h, w, 3)).astype(np.float32)
    img[rng.random((h, w)) < 0.3] = 0.0
    img[..., 1] = np.round(img[..., 1] * 8.0) / 8.0  # many ties
    film = S.Film(ctx, w, h)
    film.upload(img.ravel())
    lum = np.sort((img[..., 1] * (np.float32(1.0) / np.float32(7) * np.float32(3.0))).ravel())
    for q in (0.25, 0.9, 0.997):
        idx = min(int(np.float32(w * h) * np.float32(q)), w * h - 1)
        assert S.display_quantile(ctx, film, None, 7, 3.0, q) == max(lum[idx], np.float32(1e-9))
    film.free()


def test_display_arguments_fail_loudly(ctx):
    film = S.Film(ctx, 8, 8)
    out = S.DeviceBuffer(ctx, 64)
    lib = A.lib()
    rc = lib.spx_display_resolve(ctx.handle, film.ptr, None, 8, 8, 1, 1.0, 1.0, 40.0, out.ptr, None)
    assert rc == A.SPX_EINVAL and b"kradius" in lib.spx_last_error(ctx.handle)
    q = C.c_float(0.0)
    rc = lib.spx_display_quantile(ctx.handle, film.ptr, None, 8, 8, 1, 1.0, 1.5, C.byref(q), None)
    assert rc == A.SPX_EINVAL
    rc = lib.spx_display_quantile(ctx.handle, None, None, 8, 8, 1, 1.0, 0.5, C.byref(q), None)
    assert rc == A.SPX_EINVAL
    film.free()
    out.free()
