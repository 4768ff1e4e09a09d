"""The floor-caustic product image (prism_on_plane.rs:76-185) on the device
against the oracle, pass by pass and end to end at a reduced size: the box
blur bit-exact, the surface pass and z-buffer to fp32 rounding (pixels on a
silhouette may flip), the composite tonemap within one byte, and the whole
image (caustic, blur, gain, surface, z-buffered beam-estimator volume,
99.5th-percentile exposure, sRGB8) with films at the 0.5 % bar and bytes
within a step on all but a few pixels."""
import numpy as np
import pytest

import spectral_amd as S
from spectral_amd import _abi as A
from spectral_amd import product, scenes
from spectral_amd.core import box_blur, composite_srgb8, floor_surface, opts

from _metrics import rel_l1_per_channel

pytestmark = pytest.mark.gpu
THREADS = 16


@pytest.mark.parametrize("r", [0, 2, 7])
def test_box_blur_bit_exact(ctx, oracle, r):
    rng = np.random.default_rng(r)
    img = rng.exponential(1.0, (70, 90, 3)).astype(np.float32)
    film = S.Film(ctx, 90, 70)
    film.upload(img.ravel())
    out = box_blur(ctx, film, r)
    assert np.array_equal(out.image(), oracle.box_blur(img, r))
    # in place, as prism_on_plane's floor.sum = box_blur(floor.sum)
    box_blur(ctx, film, r, out=film)
    assert np.array_equal(film.image(), oracle.box_blur(img, r))
    out.free()
    film.free()


def test_floor_surface_matches_the_oracle(ctx, oracle):
    w, h = 200, 150
    sc, cam, scr = scenes.wedge_scene(), scenes.plane_camera(w, h), scenes.plane_floor_screen(64, 64)
    rng = np.random.default_rng(3)
    floor = rng.exponential(1.0, (64, 64, 3)).astype(np.float32)
    fdev = S.Film(ctx, 64, 64)
    fdev.upload(floor.ravel())
    sh = scenes.PLANE_SHADE
    surf, depth = floor_surface(ctx, sc, cam, scr, fdev, 0.25, sh["ambient"], sh["background"], sh["tint"], w, h)
    gs, gd = surf.image(), depth.download().reshape(h, w)
    os_, od = oracle.floor_surface(sc, cam, scr, floor, 0.25, sh["ambient"], sh["background"], sh["tint"], w, h)
    same_kind = np.isinf(gd) == np.isinf(od)
    close = np.isclose(gd, od, rtol=1e-5) | (np.isinf(gd) & np.isinf(od))
    sclose = np.isclose(gs, os_, rtol=1e-5, atol=1e-7).all(axis=-1)
    bad = ~(same_kind & close & sclose)
    print(f"surface pass: {bad.sum()} of {w * h} pixels differ")
    assert bad.mean() < 2e-3
    for f in (fdev, surf, depth):
        f.free()


def test_composite_srgb8_matches_the_oracle(ctx, oracle):
    rng = np.random.default_rng(4)
    surf = rng.uniform(0.0, 0.3, (60, 80, 3)).astype(np.float32)
    vol = rng.exponential(0.5, (60, 80, 3)).astype(np.float32)
    a, b = S.Film(ctx, 80, 60), S.Film(ctx, 80, 60)
    a.upload(surf.ravel())
    b.upload(vol.ravel())
    got = composite_srgb8(ctx, a, b, 1.3)
    want = oracle.composite_srgb8(surf, vol, 1.3)
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    a.free()
    b.free()


def test_prism_on_plane_image_matches_the_oracle(ctx, oracle):
    w, h, side, n_c, n_v = 200, 150, 128, 1 << 18, 1 << 15
    img = product.prism_on_plane(ctx, w, h, n_c, n_v, side)
    # the oracle runs the same pipeline
    sc, scr, bm = scenes.wedge_scene(), scenes.plane_floor_screen(side, side), scenes.plane_beam()
    o = opts(A.ILLUM_E)
    floor, _ = oracle.lighttrace(sc, scr, bm, n_c, 0, 0x0F1, o, threads=THREADS)
    floor = oracle.box_blur(floor, 2)
    gain = np.float32(6.0) / max(floor[..., 1].max(), np.float32(1e-9))
    cam, sh, hz = scenes.plane_camera(w, h), scenes.PLANE_SHADE, scenes.PLANE_HAZE
    surf, depth = oracle.floor_surface(sc, cam, scr, floor, gain, sh["ambient"], sh["background"], sh["tint"], w, h)
    vp = A.VolParams(hz["sigma_s"], hz["sigma_t"], hz["g"], hz["max_dist"])
    vol, _ = oracle.volume(sc, cam, bm, w, h, n_v, 0, vp, depth.ravel(), 0xBEE5, A.EST_BEAM, o, threads=THREADS)
    p = oracle.display_quantile(vol, surf, 1, 1.0, 0.995)
    exposure = np.float32(1.1) / np.float32(p)
    rgb = oracle.composite_srgb8(surf, vol, exposure)
    ef, ev = rel_l1_per_channel(img.floor, floor), rel_l1_per_channel(img.volume, vol)
    print(f"floor rel L1 {ef}, volume rel L1 {ev}, gain {img.floor_gain} / {gain}, exposure {img.exposure} / "
          f"{exposure}")
    assert (ef < 5e-3).all() and (ev < 5e-3).all()
    assert abs(img.floor_gain / gain - 1) < 1e-3 and abs(img.exposure / exposure - 1) < 1e-3
    d = np.abs(img.rgb.astype(np.int32) - rgb.astype(np.int32))
    print(f"image: max byte diff {d.max()}, pixels differing by more than one {(d > 1).any(-1).mean():.2e}")
    assert (d > 1).any(-1).mean() < 5e-3
    assert img.rgb.max() > 0 and vol[..., 1].sum() > 0


def test_prism_on_plane_on_a_stream_equals_the_null_stream_run(ctx):
    # Every pass of the product image runs on the caller's stream: a
    # non-blocking stream gives the null-stream image and exposure (the
    # quantile, composite and downloads wait for that stream's renders).
    w, h, side, n_c, n_v = 200, 150, 128, 1 << 18, 1 << 16
    ref = product.prism_on_plane(ctx, w, h, n_c, n_v, side)
    st = ctx.stream()
    try:
        got = product.prism_on_plane(ctx, w, h, n_c, n_v, side, stream=st)
    finally:
        ctx._lib.spx_stream_destroy(ctx.handle, st)
    ev = rel_l1_per_channel(got.volume, ref.volume)
    print(f"stream vs null stream: exposure {got.exposure} / {ref.exposure}, volume rel L1 {ev}")
    assert abs(got.exposure / ref.exposure - 1) < 1e-5 and abs(got.floor_gain / ref.floor_gain - 1) < 1e-5
    assert (ev < 1e-5).all()
    # the depth pass is deterministic; the surface shades the floor film,
    # whose fp32 atomic sums vary in their last bits from run to run
    assert bool(np.array_equal(got.depth, ref.depth)), "depth differs"
    sd = float(np.abs(got.surface - ref.surface).max())
    assert sd < 1e-5 * max(float(np.abs(ref.surface).max()), 1.0), f"surface differs by {sd}"
    d = int(np.abs(got.rgb.astype(np.int32) - ref.rgb.astype(np.int32)).max())
    assert d <= 1, f"image differs by {d}"
