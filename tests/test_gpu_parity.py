"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on
the same inputs.  Bars: RNG bit-exact; per-photon states within 1e-6
relative with identical bounce counts; films within 0.5 % relative L1 per
XYZ channel."""
import ctypes as C
import math

import numpy as np
import pytest

import spectral_amd as S
from spectral_amd import _abi as A
from spectral_amd import scenes
from spectral_amd.core import opts

from _metrics import diff_gate_stats, rel_l1_per_channel

pytestmark = pytest.mark.gpu

FILM_TOL = 5e-3   # relative L1 per XYZ channel
PATH_RTOL = 1e-6  # per-photon states, relative
BEAM_TILE_PIX = 1280  # the wavefront beam kernel's LDS tile (kBeamTilePix, k_beam.hip)


def test_rng_parity(ctx, oracle):
    # rng_parity.rs:19-76: 16 draws of PathRng(0xABCD, 0x1234), bit-equal
    gpu = S.rng_draws(ctx, 0xABCD, 0x1234, 16)
    cpu = (C.c_uint32 * 16)()
    oracle.lib().or_rng_draws_u32(0xABCD, 0x1234, 16, cpu)
    assert list(gpu) == list(cpu)
    assert list(gpu[:4]) == [0xA1DD5847, 0x13248A2D, 0x37278DEF, 0xDE1D5B62]
    for key, seed in ((0, 0), (0xFFFFFFFF, 42), (123456789, 0xDED)):
        g = S.rng_draws(ctx, key, seed, 64)
        c = (C.c_uint32 * 64)()
        oracle.lib().or_rng_draws_u32(key, seed, 64, c)
        assert list(g) == list(c)


def test_sellmeier_fp64_bit_exact(ctx, oracle):
    # Glass::n in fp64 on the device equals the CPU bit for bit (sellmeier.rs:28-44)
    lam = np.linspace(380.0, 730.0, 3501, dtype=np.float32)
    for glass in (A.GLASS_BK7, A.GLASS_SF11, A.GLASS_WATER):
        g = S.glass_n(ctx, glass, lam)
        c = np.array([oracle.lib().or_glass_n(glass, float(l)) for l in lam], np.float32)
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32))


def _compare_records(gpu_recs, cpu_rec_fn, n):
    stats = {"max_rel": 0.0, "bounces": 0, "tir": 0, "violet": None, "tir_probe": None}
    for i in range(n):
        g = gpu_recs[i]
        c = cpu_rec_fn(i)
        assert g.lambda_ == c.lambda_, f"photon {i}: lambda {g.lambda_} vs {c.lambda_}"
        assert g.num_states == c.num_states, f"photon {i} (lambda {c.lambda_:.1f}): bounce count"
        assert g.tir_count == c.tir_count
        k = min(c.num_states, 9)
        gs = np.array(g.states, np.float64)[:k, :, :3]
        cs = np.array(c.states, np.float64)[:k, :, :3]
        err = np.abs(gs - cs) / np.maximum(np.abs(cs), 1.0)
        stats["max_rel"] = max(stats["max_rel"], float(err.max()))
        assert err.max() <= PATH_RTOL, f"photon {i}: rel err {err.max():.3e}"
        stats["bounces"] += c.num_states
        stats["tir"] += c.tir_count
        if c.lambda_ < 430 and c.num_states > 1 and stats["violet"] is None:
            stats["violet"] = i
        if c.tir_count and stats["tir_probe"] is None:
            stats["tir_probe"] = i
    return stats


@pytest.mark.parametrize("which", ["dsotm", "tir", "prism"])
def test_per_photon_parity(ctx, oracle, which):
    # forward_parity_gpu_matches_cpu, vol_photons.rs:1361-1414 (512 photons, seed 42)
    sc, dsotm, _ = scenes.dsotm_scene()
    bm = {"dsotm": dsotm, "tir": scenes.tir_beam(), "prism": scenes.prism_beam()}[which]
    if which == "prism":
        sc = scenes.prism_scene()
    o = opts(A.ILLUM_E)
    n = 512
    recs = S.photon_records(ctx, sc, bm, 0, n, 42, 14.0, 0)
    st = _compare_records(recs, lambda i: oracle.photon_record(sc, bm, i, 42, 14.0, 0, o), n)
    print(f"[{which}] {n} photons, {st['bounces']} states, {st['tir']} TIR, max rel {st['max_rel']:.2e}")
    if which == "dsotm":
        assert st["violet"] is not None
    if which == "tir":
        assert st["tir"] > 0 and st["tir_probe"] is not None


def test_per_photon_parity_lighttrace_and_dispersion_off(ctx, oracle):
    sc, bm, scr = scenes.prism_scene(), scenes.prism_beam(), scenes.floor_screen(64, 64)
    for disp in (True, False):
        o = opts(A.ILLUM_E, dispersion=disp)
        recs = S.photon_records(ctx, sc, bm, 1000, 256, 5, 14.0, 2, scr=scr, dispersion=disp)
        _compare_records(recs, lambda i: oracle.photon_record(sc, bm, 1000 + i, 5, 14.0, 2, o, scr),
                         256)


def _volume_pair(ctx, oracle, sc, cam, bm, w, h, n, vp, seed, est=A.EST_POINT, ill=A.ILLUM_E,
                 disp=True, engine=A.ENGINE_PLAIN, zbuf=None, base=0):
    zdev = None
    if zbuf is not None:
        zdev = S.DeviceBuffer(ctx, w * h, None)
        zdev.upload(zbuf)
    film = S.render_volume(ctx, sc, cam, bm, w, h, n, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist,
                           zdev, seed, est, photon_base=base, illuminant=ill, dispersion=disp,
                           engine=engine)
    g = film.image()
    film.free()
    c, _ = oracle.volume(sc, cam, bm, w, h, n, base, vp, zbuf, seed, est, opts(ill, disp), threads=8)
    return g, c


def test_vol_diff_gate_reference_config(ctx, oracle):
    # vol_diff_gate, vol_photons.rs:1425-1513: 200x200, 200k photons, seed 7,
    # VolWeights::default(), all-INF z-buffer -- at the 0.5 % per-channel bar.
    sc, bm, cam = scenes.dsotm_scene()
    vp = A.VolParams(0.5, 0.06, 0.5, 14.0)
    g, c = _volume_pair(ctx, oracle, sc, cam, bm, 200, 200, 200_000, vp, 7,
                        zbuf=np.full(200 * 200, np.inf, np.float32))
    err = rel_l1_per_channel(g, c)
    print("vol_diff_gate rel L1 per channel", err)
    assert c[..., 1].sum() > 0 and (err < FILM_TOL).all()


@pytest.mark.parametrize("engine", [A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT])
@pytest.mark.parametrize("ill,disp", [(A.ILLUM_D65, True), (A.ILLUM_A, True), (A.ILLUM_D65, False)])
def test_prism_volume_film(ctx, oracle, engine, ill, disp):
    # The DSOTM prism workload at 2^18 photons (320x180, sigma_s = 0.05, g = 0)
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera()
    vp = A.VolParams(**scenes.HAZE)
    g, c = _volume_pair(ctx, oracle, sc, cam, bm, 320, 180, 1 << 18, vp, 0, ill=ill, disp=disp,
                        engine=engine)
    err = rel_l1_per_channel(g, c)
    print(f"prism volume engine={engine} ill={ill} disp={disp} rel L1", err)
    assert (err < FILM_TOL).all()


def test_hero_comb_photon_films(ctx, oracle):
    # photons carrying the 4-wavelength hero comb (hero_lanes = 4), the
    # prism workload under D65 and A, plain / wavefront / batched, and the
    # floor caustic's lighttrace: films against the oracle's comb
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera()
    vp = A.VolParams(**scenes.HAZE)
    n = 1 << 18
    for ill in (A.ILLUM_D65, A.ILLUM_A):
        c, _ = oracle.volume(sc, cam, bm, 320, 180, n, 0, vp, None, 0, A.EST_POINT, opts(ill, hero_lanes=4),
                             threads=8)
        one, _ = oracle.volume(sc, cam, bm, 320, 180, n, 0, vp, None, 0, A.EST_POINT, opts(ill), threads=8)
        assert rel_l1_per_channel(c, one).max() > 1e-3  # the comb is a different estimate
        for eng in (A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT):
            f = S.render_volume(ctx, sc, cam, bm, 320, 180, n, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist, None, 0,
                                illuminant=ill, engine=eng, hero_lanes=4)
            err = rel_l1_per_channel(f.image(), c)
            f.free()
            print(f"hero comb ill={ill} engine={eng}: rel L1 {err}")
            assert (err < FILM_TOL).all() and (err < 1e-5).all()
        ds = sc.upload(ctx)
        film = S.Film(ctx, 320, 180)
        j = A.Job()
        j.kind, j.width, j.height, j.n_photons, j.seed = A.JOB_VOLUME, 320, 180, n, 0
        j.estimator, j.opts = A.EST_POINT, opts(ill, hero_lanes=4)
        j.camera, j.beam, j.vol = cam.basis, bm, vp
        j.scene, j.film = ds.handle.value, film.ptr
        assert A.lib().spx_batch_run(ctx.handle, (A.Job * 1)(j), 1, None, None) == 0
        ctx.sync()
        err = rel_l1_per_channel(film.image(), c)
        print(f"hero comb ill={ill} batched: rel L1 {err}")
        assert (err < 1e-5).all()
        film.free()
        ds.close()
    scr = scenes.floor_screen(256, 256)
    cc, _ = oracle.lighttrace(sc, scr, bm, 1 << 17, 0, 0x0F1, opts(A.ILLUM_D65, hero_lanes=4), threads=8)
    for eng in (A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT):
        g = S.lighttrace_trace(ctx, sc, scr, bm, 1 << 17, 0x0F1, illuminant=A.ILLUM_D65, engine=eng, hero_lanes=4)
        err = rel_l1_per_channel(g.image(), cc)
        g.free()
        print(f"hero comb lighttrace engine={eng}: rel L1 {err}")
        assert (err < 1e-5).all()
    with pytest.raises(A.SpxError):
        S.render_volume(ctx, sc, cam, bm, 320, 180, 16, 0.05, 0.05, 0.0, 14.0, None, 0, estimator=A.EST_BEAM,
                        hero_lanes=4)


@pytest.mark.parametrize("engine", [A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT])
def test_floor_volume_film_with_zbuffer(ctx, oracle, engine):
    # prism + Lambertian floor, z-buffer from the surface pass (prism_on_plane.rs:125-164)
    w, h = 192, 108
    sc, bm, cam = scenes.prism_scene(floor=True), scenes.prism_beam(), scenes.floor_camera(w, h)
    _, depth = oracle.surface_rim(sc, cam, w, h)
    vp = A.VolParams(**scenes.HAZE)
    g, c = _volume_pair(ctx, oracle, sc, cam, bm, w, h, 1 << 17, vp, 3, ill=A.ILLUM_D65,
                        engine=engine, zbuf=depth.ravel())
    err = rel_l1_per_channel(g, c)
    print("floor volume rel L1", err)
    assert (err < FILM_TOL).all()


def test_legacy_and_beam_estimators(ctx, oracle):
    sc, bm, cam = scenes.dsotm_scene()
    vp = A.VolParams(0.5, 0.06, 0.5, 14.0)
    for est, n in ((A.EST_LEGACY, 50_000), (A.EST_BEAM, 2_000)):
        g, c = _volume_pair(ctx, oracle, sc, cam, bm, 96, 96, n, vp, 11, est=est)
        err = rel_l1_per_channel(g, c)
        print("estimator", est, "rel L1", err)
        assert (err < FILM_TOL).all()
        if est == A.EST_BEAM:
            # every texel of every tent, in the same order: the films differ
            # only in the order of fp32 atomic sums (the oracle with each
            # tent's last column left out of its deposits reads 5.5e-3)
            assert (err < 2e-5).all()


@pytest.mark.parametrize("chunk_log2", [12, 22])
def test_beam_estimator_wavefront_engine(ctx, oracle, chunk_log2):
    # the beam estimator with its segments queued and their steps spread over
    # waves: the oracle at the bar, the per-photon form to summation order;
    # chunk_log2 = 12 splits 2000 photons into 4 chunks of 512
    sc, bm, cam = scenes.dsotm_scene()
    vp = A.VolParams(0.5, 0.06, 0.5, 14.0)
    old = ctx.tuning()
    ctx.set_tuning(chunk_log2=chunk_log2)
    try:
        g, c = _volume_pair(ctx, oracle, sc, cam, bm, 96, 96, 2_000, vp, 11, est=A.EST_BEAM,
                            engine=A.ENGINE_WAVEFRONT)
    finally:
        ctx.set_tuning(chunk_log2=old.chunk_log2)
    p, _ = _volume_pair(ctx, oracle, sc, cam, bm, 96, 96, 2_000, vp, 11, est=A.EST_BEAM)
    err, err_p = rel_l1_per_channel(g, c), rel_l1_per_channel(g, p)
    print("beam wavefront: rel L1 vs oracle", err, "vs per-photon", err_p)
    assert (err < 2e-5).all() and (err_p < 1e-4).all() and c[..., 1].sum() > 0


def test_chunked_equals_single_dispatch(ctx, oracle):
    # photon_base chunking (vol_photons.rs:590-619): two halves == one dispatch
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera(160, 90)
    vp = A.VolParams(**scenes.HAZE)
    one = S.render_volume(ctx, sc, cam, bm, 160, 90, 1 << 16, vp.sigma_s, vp.sigma_t, vp.g,
                          vp.max_dist, None, 9)
    two = S.render_volume(ctx, sc, cam, bm, 160, 90, 1 << 15, vp.sigma_s, vp.sigma_t, vp.g,
                          vp.max_dist, None, 9)
    S.render_volume(ctx, sc, cam, bm, 160, 90, 1 << 15, vp.sigma_s, vp.sigma_t, vp.g, vp.max_dist,
                    None, 9, film=two, photon_base=1 << 15)
    assert (rel_l1_per_channel(two.image(), one.image()) < 1e-5).all()


@pytest.mark.parametrize("scene_name", ["diff_gate", "lambertian", "rainbow"])
@pytest.mark.parametrize("engine", [A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT])
def test_camera_film(ctx, oracle, scene_name, engine):
    # diff_gate.rs:95-144 / lambertian_gate.rs:28-92 (64x64, 64 spp, seed 0xD1FF)
    # and the rainbow workload (RR after depth 3), per-channel rel L1 < 0.5 %.
    if scene_name == "rainbow":
        sc, cam, o = scenes.rainbow_scene(), scenes.rainbow_camera(64, 64), dict(rr_depth=3)
    else:
        sc = scenes.diff_gate_scene() if scene_name == "diff_gate" else scenes.lambertian_scene()
        cam, o = scenes.gate_camera(), {}
    w = h = 64
    spp, seed = 64, 0xD1FF
    tr = S.HipTracer(ctx, sc, cam, w, h, A.ILLUM_D65, seed, engine=engine, **o)
    tr.accumulate(spp)
    g = tr.buffer()
    c, _ = oracle.camera_accumulate(sc, cam, w, h, spp, 0, seed, opts(A.ILLUM_D65, **o), threads=8)
    c = c / spp
    err = rel_l1_per_channel(g, c)
    st = diff_gate_stats(g, c)
    print(scene_name, engine, "rel L1", err, st)
    assert (err < FILM_TOL).all()
    assert st["frac_above_10x_mean"] < 0.005 and st["p99"] < 0.05 and st["energy_rel"] < 0.01


@pytest.mark.parametrize("engine", [A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT])
def test_scene_a_horizon_camera_film(ctx, oracle, engine):
    # scene_a_prism.rs:66-90 on the device: the horizon sky (scene.rs:44-50)
    # lights escaped rays looking up; SF11 prism and its diffuse control,
    # 128x128, 64 spp, seed 11 -- films against the oracle, and the
    # reference's gate (glass chroma > 0.05 and > 2x the control) on them
    cam, spread = scenes.scene_a_camera(), []
    for disp in (True, False):
        sc = scenes.scene_a_scene(disp)
        tr = S.HipTracer(ctx, sc, cam, 128, 128, A.ILLUM_D65, 11, engine=engine)
        tr.accumulate(64)
        g = tr.sum()
        c, _ = oracle.camera_accumulate(sc, cam, 128, 128, 64, 0, 11, opts(A.ILLUM_D65), threads=8)
        err = rel_l1_per_channel(g, c)
        print("scene A", "glass" if disp else "diffuse", engine, "rel L1", err,
              "bit-identical", np.array_equal(g, c))
        assert (err < FILM_TOL).all()
        up = g[:16, :16, 1].mean() / 64  # upper-left corner: sky above the horizon
        down = g[-16:, :16, 1].mean() / 64
        assert up > 10 * down > 0
        px = (g / np.float32(64)).reshape(-1, 3)
        s = px.sum(1)
        ok = s > 1e-5
        spread.append(float(np.sqrt((px[ok, 0] / s[ok] - 0.3127) ** 2 + (px[ok, 1] / s[ok] - 0.3290) ** 2).max()))
        tr.film.free()
    assert spread[0] > 0.05 and spread[0] > 2.0 * spread[1]


def test_camera_wavefront_equals_plain_bitwise(ctx):
    # The wavefront resolve sums each pixel's samples in sample order, as the
    # plain kernel does, so the two films are identical.
    sc, cam = scenes.rainbow_scene(), scenes.rainbow_camera(96, 80)
    films = []
    for eng in (A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT):
        tr = S.HipTracer(ctx, sc, cam, 96, 80, A.ILLUM_D65, 1, engine=eng, rr_depth=3)
        tr.accumulate(24)
        films.append(tr.sum())
    assert np.array_equal(films[0], films[1])


def test_camera_progressive(ctx):
    # accumulate(32) + accumulate(32) == accumulate(64) (tracer.rs:119, sample_base)
    sc, cam = scenes.diff_gate_scene(), scenes.gate_camera()
    a = S.HipTracer(ctx, sc, cam, 32, 32, A.ILLUM_D65, 5)
    a.accumulate(64)
    b = S.HipTracer(ctx, sc, cam, 32, 32, A.ILLUM_D65, 5)
    b.accumulate(32)
    b.accumulate(32)
    assert np.allclose(a.sum(), b.sum(), rtol=1e-5, atol=1e-4)
    b.reset()
    assert not b.sum().any() and b.samples == 0


@pytest.mark.parametrize("engine", [A.ENGINE_PLAIN, A.ENGINE_WAVEFRONT])
def test_lighttrace_film(ctx, oracle, engine):
    # lighttrace::trace onto the floor screen (prism_on_plane.rs:107 pattern)
    sc, bm, scr = scenes.prism_scene(), scenes.prism_beam(), scenes.floor_screen(256, 256)
    n = 1 << 18
    film = S.lighttrace_trace(ctx, sc, scr, bm, n, 0x0F1, illuminant=A.ILLUM_D65, engine=engine)
    g = film.image()
    c, _ = oracle.lighttrace(sc, scr, bm, n, 0, 0x0F1, opts(A.ILLUM_D65), threads=8)
    err = rel_l1_per_channel(g, c)
    print("lighttrace rel L1", err, "energy", g[..., 1].sum(), c[..., 1].sum())
    assert c[..., 1].sum() > 0 and (err < FILM_TOL).all()


def test_wedge_lighttrace_screen(ctx, oracle):
    # screen_sample_reads_deposited_energy, lighttrace.rs:231-246
    film = S.lighttrace_trace(ctx, scenes.wedge_scene(), scenes.wedge_screen(), scenes.wedge_beam(),
                              100_000, 1)
    c, _ = oracle.lighttrace(scenes.wedge_scene(), scenes.wedge_screen(), scenes.wedge_beam(),
                             100_000, 0, 1, opts(A.ILLUM_E), threads=8)
    g = film.image()
    assert g[..., 1].max() > 0 and (rel_l1_per_channel(g, c) < FILM_TOL).all()


def test_surface_rim_gate(ctx, oracle):
    # surface_rim_matches_cpu, vol_photons.rs:1547-1606: silhouette exact, max delta < 1e-4
    sc, _, cam = scenes.dsotm_scene()
    rim, depth = S.surface_rim(ctx, sc, cam, 300, 300)
    g = rim.image()
    c, cd = oracle.surface_rim(sc, cam, 300, 300)
    assert ((g > 0).any(-1) == (c > 0).any(-1)).all() and (c > 0).any()
    assert np.abs(g - c).max() < 1e-4
    gd = depth.download().reshape(300, 300)
    assert np.array_equal(np.isinf(gd), np.isinf(cd))
    fin = np.isfinite(cd)
    assert np.abs(gd[fin] - cd[fin]).max() <= 1e-6 * np.abs(cd[fin]).max()


def test_invalid_arguments_fail_loudly(ctx):
    sc = scenes.prism_scene().upload(ctx)
    cam = scenes.prism_camera()
    film = S.Film(ctx, 8, 8)
    lib = A.lib()
    o = opts()
    rc = lib.spx_camera_accumulate(ctx.handle, sc.handle, C.byref(cam.basis), 0, 8, 1, 0, 0,
                                   C.byref(o), 0, film.ptr, None)
    assert rc == A.SPX_EINVAL and b"zero-size" in lib.spx_last_error(ctx.handle)
    bad = A.Opts(0, 1, 0, 0)
    rc = lib.spx_camera_accumulate(ctx.handle, sc.handle, C.byref(cam.basis), 8, 8, 1, 0, 0,
                                   C.byref(bad), 0, film.ptr, None)
    assert rc == A.SPX_EINVAL
    nan_cam = A.CameraBasis()
    nan_cam.origin[0] = float("nan")
    rc = lib.spx_camera_accumulate(ctx.handle, sc.handle, C.byref(nan_cam), 8, 8, 1, 0, 0, None, 0,
                                   film.ptr, None)
    assert rc == A.SPX_EINVAL
    # the context still works after rejected calls
    tr = S.HipTracer(ctx, scenes.diff_gate_scene(), scenes.gate_camera(), 8, 8)
    tr.accumulate(2)
    assert tr.sum().sum() > 0


@pytest.mark.parametrize("splat_mode,bin_log2", [(1, 30), (2, 30), (2, 15), (3, 30), (3, 15)])
def test_deposit_modes_wavefront(ctx, oracle, splat_mode, bin_log2):
    # splat mode 1: LDS hash in front of film atomics; mode 2: deposits as
    # records, sorted by 32x32 tile and accumulated in LDS as fixed point (no
    # film atomics); mode 3: the LDS hash draining to those records;
    # bin_log2 = 15 overflows the pools into the atomic fallback.
    old = ctx.tuning()
    ctx.set_tuning(splat_mode=splat_mode, bin_log2=bin_log2)
    try:
        sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera()
        vp = A.VolParams(**scenes.HAZE)
        g, c = _volume_pair(ctx, oracle, sc, cam, bm, 320, 180, 1 << 18, vp, 0, ill=A.ILLUM_D65,
                            engine=A.ENGINE_WAVEFRONT)
        err = rel_l1_per_channel(g, c)
        print("prism volume", splat_mode, bin_log2, err)
        assert (err < FILM_TOL).all()
        w, h = 192, 108
        fsc, fcam = scenes.prism_scene(floor=True), scenes.floor_camera(w, h)
        _, depth = oracle.surface_rim(fsc, fcam, w, h)
        g, c = _volume_pair(ctx, oracle, fsc, fcam, bm, w, h, 1 << 16, vp, 3, ill=A.ILLUM_A,
                            engine=A.ENGINE_WAVEFRONT, zbuf=depth.ravel())
        assert (rel_l1_per_channel(g, c) < FILM_TOL).all()
        scr = scenes.floor_screen(256, 256)
        film = S.lighttrace_trace(ctx, sc, scr, bm, 1 << 17, 0x0F1, illuminant=A.ILLUM_D65,
                                  engine=A.ENGINE_WAVEFRONT)
        c, _ = oracle.lighttrace(sc, scr, bm, 1 << 17, 0, 0x0F1, opts(A.ILLUM_D65), threads=8)
        assert (rel_l1_per_channel(film.image(), c) < FILM_TOL).all()
    finally:
        ctx.set_tuning(splat_mode=old.splat_mode, bin_log2=old.bin_log2)


def test_binned_fixed_point_scaling_and_determinism(ctx):
    # Binned deposits sum in 64-bit fixed point scaled to each tile's largest
    # deposit.  sigma_s multiplies every deposit (volume.rs point estimator),
    # so sigma_s * 2^k must give the film * 2^k exactly, for k far outside
    # any fixed scale; and two runs must agree bit for bit.
    old = ctx.tuning()
    ctx.set_tuning(splat_mode=2, bin_log2=30)
    try:
        sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera(160, 90)
        base = None
        for k in (0, 0, 90, -90):
            f = S.render_volume(ctx, sc, cam, bm, 160, 90, 1 << 16, 0.05 * 2.0 ** k, 0.05, 0.0, 14.0,
                                None, 21, illuminant=A.ILLUM_D65, engine=A.ENGINE_WAVEFRONT)
            img = f.image().astype(np.float64)
            if base is None:
                base = img
                assert base.sum() > 0
            else:
                assert np.array_equal(img, base * 2.0 ** k), k
    finally:
        ctx.set_tuning(splat_mode=old.splat_mode, bin_log2=old.bin_log2)


def test_beam_wavefront_groups_larger_than_the_tile(ctx, oracle):
    # a camera 0.15-1.35 units above the DSOTM beam sheet, looking down on
    # it: tent boxes of 11-50 pixels a side (4 % of the steps 45 or more, at
    # the 24 px clamp; the oracle's counts), so the boxes of the groups of 64
    # steps with the largest tents exceed the LDS tile (1280 pixels): those
    # groups are split into halves, quarters, ... until each part fits, and
    # single steps whose own box exceeds the tile splat straight to the film;
    # the film still equals the per-photon form to summation order and the
    # oracle
    sc, bm, _ = scenes.dsotm_scene()
    cam = S.Camera.look_at((-1.93, -0.52, 0.75), (-1.93, -0.52, 0.0), (0, 1, 0), 60.0, 1.0)
    vp = A.VolParams(0.5, 0.06, 0.5, 14.0)
    w = h = 256
    # the oracle's own steps: some group of 64 has a box above the tile
    box, _ = oracle.volume_beam_max_box(sc, cam, bm, w, h, 200, 0, vp, None, 5, opts(A.ILLUM_E), 64)
    box1, _ = oracle.volume_beam_max_box(sc, cam, bm, w, h, 200, 0, vp, None, 5, opts(A.ILLUM_E), 1)
    print("largest 64-step box", box, "pixels, largest single-step box", box1)
    assert box > BEAM_TILE_PIX and box1 > BEAM_TILE_PIX
    g, c = _volume_pair(ctx, oracle, sc, cam, bm, w, h, 200, vp, 5, est=A.EST_BEAM, engine=A.ENGINE_WAVEFRONT)
    p, _ = _volume_pair(ctx, oracle, sc, cam, bm, w, h, 200, vp, 5, est=A.EST_BEAM)
    err, err_p = rel_l1_per_channel(g, c), rel_l1_per_channel(g, p)
    print("close-up beam: rel L1 vs oracle", err, "vs per-photon", err_p, "lit pixels", (g[..., 1] > 0).mean())
    assert c[..., 1].sum() > 0 and (g[..., 1] > 0).mean() > 0.05
    assert (err < 2e-5).all() and (err_p < 1e-4).all()
