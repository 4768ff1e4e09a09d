"""Film parity of the benchmark configurations, engine by engine, against the
oracle at the configuration's film size, seed and sample count:

  C0  DSOTM prism, SF11, 0.02-wide D65 beam, haze sigma_s = 0.05, 320x180,
      2^20 photons, seed 0 (full size)
  C1  rainbow: unit water sphere, D65 sun + dim sky, 512x512, 64 spp, depth 8,
      Russian roulette from depth 3, seed 1 (full size)
  C2  prism on a Lambertian floor (albedo 0.8), 1920x1080 with the surface
      z-buffer, under D65, A and dispersion off, plus the floor caustic on a
      1024x1024 screen: full size, 2^28 photons (the oracle on 16 threads)

Bar: relative L1 per XYZ channel below 0.5 %.  Each test prints its errors."""
import ctypes as C

import numpy as np
import pytest

import spectral_amd as S
from spectral_amd import _abi as A
from spectral_amd import scenes
from spectral_amd.core import _check, opts

from _metrics import rel_l1_per_channel

pytestmark = pytest.mark.gpu

FILM_TOL = 5e-3
THREADS = 16


def _batched(ctx, job):
    """One job through the batched engine (spx_batch_run)."""
    tbl = (A.Job * 1)(job)
    _check(ctx, A.lib().spx_batch_run(ctx.handle, tbl, 1, None, None))
    ctx.sync()


def _volume_job(ds, cam, bm, w, h, n, seed, ill, disp, film, zbuf=None):
    j = A.Job()
    j.kind, j.width, j.height, j.n_photons, j.seed = A.JOB_VOLUME, w, h, n, seed
    j.estimator = A.EST_POINT
    j.opts = opts(ill, disp)
    j.camera = cam.basis
    j.beam = bm
    j.vol = A.VolParams(**scenes.HAZE)
    j.scene = ds.handle.value
    j.film = film.ptr
    j.zbuf = zbuf.ptr if zbuf is not None else None
    return j


def _report(name, errs):
    for eng, e in errs.items():
        print(f"{name} {eng}: rel L1 X {e[0]:.2e} Y {e[1]:.2e} Z {e[2]:.2e}")
        assert (e < FILM_TOL).all(), f"{name} {eng}: {e}"


def test_config0_prism(ctx, oracle):
    w, h, n = 320, 180, 1 << 20
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera(w, h)
    vp = A.VolParams(**scenes.HAZE)
    c, _ = oracle.volume(sc, cam, bm, w, h, n, 0, vp, None, 0, A.EST_POINT, opts(A.ILLUM_D65),
                         threads=THREADS)
    errs = {}
    for eng, name in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront")):
        f = S.render_volume(ctx, sc, cam, bm, w, h, n, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist, None, 0,
                            illuminant=A.ILLUM_D65, engine=eng)
        errs[name] = rel_l1_per_channel(f.image(), c)
        f.free()
    ds = sc.upload(ctx)
    film = S.Film(ctx, w, h)
    _batched(ctx, _volume_job(ds, cam, bm, w, h, n, 0, A.ILLUM_D65, True, film))
    errs["batched"] = rel_l1_per_channel(film.image(), c)
    film.free()
    ds.close()
    assert c[..., 1].sum() > 0
    _report("C0 prism 320x180 2^20", errs)


def test_config1_rainbow(ctx, oracle):
    w = h = 512
    spp = 64
    sc, cam = scenes.rainbow_scene(), scenes.rainbow_camera(w, h)
    o = opts(A.ILLUM_D65, True, 8, 3)
    c, _ = oracle.camera_accumulate(sc, cam, w, h, spp, 0, 1, o, threads=THREADS)
    errs = {}
    for eng, name in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront")):
        tr = S.HipTracer(ctx, sc, cam, w, h, A.ILLUM_D65, 1, engine=eng, rr_depth=3)
        tr.accumulate(spp)
        errs[name] = rel_l1_per_channel(tr.sum(), c)
        exact = np.array_equal(tr.sum(), c)
        print(f"C1 {name}: film bit-identical to the oracle: {exact}")
        tr.film.free()
    ds = sc.upload(ctx)
    film = S.Film(ctx, w, h)
    j = A.Job()
    j.kind, j.width, j.height, j.spp, j.seed = A.JOB_CAMERA, w, h, spp, 1
    j.opts = o
    j.camera = cam.basis
    j.scene = ds.handle.value
    j.film = film.ptr
    _batched(ctx, j)
    errs["batched"] = rel_l1_per_channel(film.image(), c)
    film.free()
    ds.close()
    _report("C1 rainbow 512x512x64", errs)


@pytest.mark.parametrize("variant", ["D65", "A", "disp_off"])
def test_config2_floor(ctx, oracle, variant):
    w, h, n = 1920, 1080, 1 << 28
    ill = A.ILLUM_A if variant == "A" else A.ILLUM_D65
    disp = variant != "disp_off"
    fs, fc, bm = scenes.prism_scene(floor=True), scenes.floor_camera(w, h), scenes.prism_beam()
    vp = A.VolParams(**scenes.HAZE)
    _, depth = oracle.surface_rim(fs, fc, w, h)
    zdev = S.DeviceBuffer(ctx, w * h, None)
    zdev.upload(depth)
    c, _ = oracle.volume(fs, fc, bm, w, h, n, 0, vp, depth.ravel(), 0, A.EST_POINT, opts(ill, disp),
                         threads=THREADS)
    errs = {}
    f = S.render_volume(ctx, fs, fc, bm, w, h, n, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist, zdev, 0,
                        illuminant=ill, dispersion=disp, engine=A.ENGINE_PLAIN)
    errs["plain"] = rel_l1_per_channel(f.image(), c)
    f.free()
    ds = fs.upload(ctx)
    film = S.Film(ctx, w, h)
    _batched(ctx, _volume_job(ds, fc, bm, w, h, n, 0, ill, disp, film, zdev))
    errs["batched"] = rel_l1_per_channel(film.image(), c)
    film.free()
    ds.close()
    # the floor caustic: lighttrace onto the floor as a screen
    ps, scr = scenes.prism_scene(), scenes.floor_screen(1024, 1024)
    cc, _ = oracle.lighttrace(ps, scr, bm, n, 0, 0x0F1, opts(ill, disp), threads=THREADS)
    g = S.lighttrace_trace(ctx, ps, scr, bm, n, 0x0F1, illuminant=ill, dispersion=disp)
    errs["caustic plain"] = rel_l1_per_channel(g.image(), cc)
    g.free()
    zdev.free()
    assert c[..., 1].sum() > 0 and cc[..., 1].sum() > 0
    _report(f"C2 floor 1920x1080 2^28 {variant}", errs)
