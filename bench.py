"""Benchmark: Mphotons/s + Mpaths/s of the light-transport core on one MI355X.

One step = the mixed stream of 512 uneven jobs (half rainbow camera renders,
half prism-on-floor volume renders; spectral_amd.jobs.mixed_jobs, seed 0)
traced by the batched engine, inputs resident in HBM: the table's schedule is
captured once as a hipGraph (spx_batch_plan_create) and each step replays it.
The launch shape is the autotuner's pick on this table (spx_autotune,
recorded in TUNED_MIXED512; --autotune runs the search again first).
value = (paths + photons) per second / 1e6 over exactly K timed steps.

Also reported on the same line: the plain one-job-per-launch form on the same
stream, the eager (ungraphed) step, per-kernel times of the batched step (HIP
events around each kernel group of K eager steps whose rounds run in
sequence, on the stream each group is launched on) and the roofline of the
dominant kernel, the per-configuration throughputs (DSOTM prism 2^20 photons at
320x180; rainbow 512^2 x 64 spp; full-size prism-on-floor 2^28 photons at
1920x1080 under D65, A and dispersion off, plus its floor caustic), the
served stream's p50/p99 job latency and the CPU oracle timed on this host.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
                       [--autotune | --no-tune | --tuning JSON] [--sequential]
                       [--no-cpu] [--no-configs] [--no-serve] [--no-plain]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mphotons/s + Mpaths/s on 1 MI355X, batched vs plain; p50/p99 job latency"
from spectral_amd.jobs import TUNED_MIXED512  # noqa: E402  spx_autotune's pick on the 512-job table

UNIT = "Mphotons/s + Mpaths/s"
HBM_PEAK_GBS = 8000.0
# VALU lane operations per second: 256 CUs x 4 SIMDs x 16 lanes per clock
# (a wave64 instruction issues over 4 clocks; packed fp32 is compiled out)
# x 2.4 GHz peak engine clock.
VALU_LANE_PEAK = 256 * 4 * 16 * 2.4e9
# Algorithmic bytes (per unit of work): a camera path segment moves a 44 B
# state in and out (88 B) and a finished path deposits one XYZ (12 B); a
# photon segment moves a 40 B state in and out (80 B) and each film deposit
# is 3 fp32 atomics (12 B).
CAM_SEG_BYTES, CAM_PATH_BYTES, PH_SEG_BYTES, SPLAT_BYTES = 88, 12, 80, 12
# Segments per camera path, segments and film deposits per photon of the
# mixed stream's two job classes, as the oracle counts them on the samples
# cpu_baseline runs (used only when that leg is skipped; the CPU test
# tests/test_work_model.py re-derives them from the oracle).
WORK_MODEL_PINNED = {"seg_per_path": 1.787, "seg_per_photon": 4.773, "splat_per_photon": 13.94}
# Volume samples of the CPU leg: (prism rotation deg, beam offset).
VOLUME_SAMPLES = ((0.0, 0.0), (20.0, 0.05), (45.0, -0.08))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    def __init__(self, ctx, stream):
        self.ctx, self.stream = ctx, stream
        self.e0, self.e1 = ctx.event(), ctx.event()

    def run(self, fn, reps=1):
        self.ctx.sync(self.stream)
        t0 = time.perf_counter()
        self.ctx.record(self.e0, self.stream)
        for _ in range(reps):
            fn()
        self.ctx.record(self.e1, self.stream)
        ms = self.ctx.elapsed_ms(self.e0, self.e1)
        self.ctx.sync(self.stream)
        return ms / reps, (time.perf_counter() - t0) * 1e3 / reps


def profile_figure(summary: str, key: str):
    """A `key: value` figure of a committed rocprofv3 summary under profiles/
    (None when the summary has no figure for it)."""
    try:
        for line in (ROOT / "profiles" / summary).read_text().splitlines():
            if line.startswith(f"{key}:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


SQ_PROFILE = "batched_step_mixed512_sq_counters.md"
HBM_PROFILE = "batched_step_mixed512_hbm_counters.md"


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload (FETCH_SIZE and
    WRITE_SIZE, copy-calibrated)."""
    return profile_figure(HBM_PROFILE, f"traffic_bytes_per_launch[{kernel}]")


def _median_time(fn, reps=5):
    """One warm-up run, then the median wall time of `reps` timed runs (s)."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(paths_total, photons_total, reps=5):
    """The oracle (the C restatement of spectral-core) on this host, each
    figure one warm-up run plus the median of `reps`: the configurations
    one by one -- DSOTM prism 2^20 photons at 320x180 on one thread (the
    reference's single-threaded photon loop, volume.rs:118) and on all
    threads; rainbow 512^2 x 64 spp on all threads (its rayon loop,
    tracer.rs:116-134); the prism-on-floor volume at 1920x1080 per variant
    on a 2^18-photon sample on one thread and a 2^22-photon sample on all
    threads (photons are independent, so a rate measured on the first n of
    a render's photons is its rate) -- and the mixed stream at the rainbow's
    all-thread path rate and the single-threaded rate of its volume class
    (three prism rotations at 512x512).  Also returns the oracle's segment
    and deposit counts on those runs (the roofline's work model)."""
    from oracle import oracle as O
    from spectral_amd import _abi as A
    from spectral_amd import jobs, scenes
    from spectral_amd.core import opts
    cores = min(os.cpu_count() or 1, 16)
    vp = A.VolParams(**scenes.HAZE)
    rows, counts = {}, {}

    def volume(sc, cam, bm, w, h, n, ill, disp, threads, zbuf=None, key=None):
        def run():
            _, st = O.volume(sc, cam, bm, w, h, n, 0, vp, zbuf, 0, A.EST_POINT, opts(ill, disp), threads=threads)
            if key:
                counts[key] = {"seg": st.segments / st.paths, "splat": st.splats / st.paths}
        return n / _median_time(run, reps) / 1e6

    # DSOTM prism, the whole configuration
    sc, bm, cam = scenes.prism_scene(), scenes.prism_beam(), scenes.prism_camera(320, 180)
    rows["prism_2^20_320x180"] = {
        "mphotons_s_1thread": round(volume(sc, cam, bm, 320, 180, 1 << 20, A.ILLUM_D65, True, 1, key="prism"), 4),
        "mphotons_s_all_threads": round(volume(sc, cam, bm, 320, 180, 1 << 20, A.ILLUM_D65, True, cores), 4)}
    # rainbow, the whole configuration
    rs, rc = scenes.rainbow_scene(), scenes.rainbow_camera(512, 512)

    def rainbow():
        _, st = O.camera_accumulate(rs, rc, 512, 512, 64, 0, 1, opts(A.ILLUM_D65, True, 8, 3), threads=cores)
        counts["rainbow"] = {"seg": st.segments / st.paths}
    cam_rate = 512 * 512 * 64 / _median_time(rainbow, reps)
    rows["rainbow_512x512x64"] = {"mpaths_s_all_threads": round(cam_rate / 1e6, 4)}
    # prism on floor, per variant, on the first 2^18 / 2^22 photons
    fs, fc = scenes.prism_scene(floor=True), scenes.floor_camera(1920, 1080)
    _, depth = O.surface_rim(fs, fc, 1920, 1080)
    z = depth.ravel()
    for label, ill, disp in (("D65", A.ILLUM_D65, True), ("A", A.ILLUM_A, True), ("disp_off", A.ILLUM_D65, False)):
        rows[f"floor_volume_1920x1080_{label}"] = {
            "mphotons_s_1thread": round(volume(fs, fc, bm, 1920, 1080, 1 << 18, ill, disp, 1, z,
                                               key="floor_volume" if label == "D65" else None), 4),
            "mphotons_s_all_threads": round(volume(fs, fc, bm, 1920, 1080, 1 << 22, ill, disp, cores, z), 4),
            "sample_photons": {"1thread": 1 << 18, "all_threads": 1 << 22}}
    # the floor caustic's work model (lighttrace onto the floor screen)
    _, st = O.lighttrace(scenes.prism_scene(), scenes.floor_screen(1024, 1024), bm, 1 << 16, 0, 0x0F1,
                         opts(A.ILLUM_D65), threads=cores)
    counts["floor_caustic"] = {"seg": st.segments / st.paths, "splat": st.splats / st.paths}
    # the mixed stream's volume class, one thread
    tot = {"seg": 0, "splat": 0, "n": 0}

    def mixed_volume(n):
        def run():
            tot.update(seg=0, splat=0, n=0)
            for rot, off in VOLUME_SAMPLES:
                sp = jobs.JobSpec(A.JOB_VOLUME, 512, n_photons=n, rotation=rot, offset=off, seed=2)
                s, c, b = jobs.job_geometry(sp)
                _, st = O.volume(s, c, b, 512, 512, n, 0, vp, None, 2, A.EST_POINT, jobs.job_opts(sp), threads=1)
                tot["seg"] += st.segments
                tot["splat"] += st.splats
                tot["n"] += st.paths
        return run
    n_mix = 1 << 17
    ph_rate = 3 * n_mix / _median_time(mixed_volume(n_mix), reps)
    # the mixed stream's camera class is rainbow renders: its paths per
    # segment from a sample of the stream's own job shape
    spec = jobs.JobSpec(A.JOB_CAMERA, 256, spp=2, seed=1)
    s, c, _ = jobs.job_geometry(spec)
    _, cst = O.camera_accumulate(s, c, 256, 256, 2, 2, 1, jobs.job_opts(spec), threads=cores)
    step_s = paths_total / cam_rate + photons_total / ph_rate
    work = {"seg_per_path": cst.segments / cst.paths, "seg_per_photon": tot["seg"] / tot["n"],
            "splat_per_photon": tot["splat"] / tot["n"],
            "configs": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in counts.items()}}
    return {"value": round((paths_total + photons_total) / step_s / 1e6, 3), "unit": UNIT,
            "cores": cores, "kind": "port",
            "sample": (f"oracle (C restatement of spectral-core), warm-up + median of {reps}: the mixed "
                       f"stream's paths at the rainbow 512x512x64 rate on {cores} threads "
                       f"({cam_rate / 1e6:.3f} Mpaths/s) and its photons at the single-threaded rate of "
                       f"3 x {n_mix} prism-on-floor photons at 512x512 ({ph_rate / 1e6:.3f} Mphotons/s); "
                       f"per-configuration rates under `configs`"),
            "mpaths_s": round(cam_rate / 1e6, 4), "mphotons_s_1thread": round(ph_rate / 1e6, 4),
            "configs": rows, "host": host_cpu()}, work


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu": model}


def config_work(O):
    """Segments and deposits per unit of each single-job configuration, from
    the oracle's counters on small samples of the same scenes (used when the
    CPU leg is skipped)."""
    from spectral_amd import _abi as A
    from spectral_amd import scenes
    from spectral_amd.core import opts
    vp, n = A.VolParams(**scenes.HAZE), 1 << 14
    out = {}
    _, st = O.volume(scenes.prism_scene(), scenes.prism_camera(320, 180), scenes.prism_beam(), 320, 180, n, 0, vp,
                     None, 0, A.EST_POINT, opts(A.ILLUM_D65), threads=4)
    out["prism"] = {"seg": st.segments / st.paths, "splat": st.splats / st.paths}
    _, st = O.camera_accumulate(scenes.rainbow_scene(), scenes.rainbow_camera(128, 128), 128, 128, 4, 0, 1,
                                opts(A.ILLUM_D65, True, 8, 3), threads=4)
    out["rainbow"] = {"seg": st.segments / st.paths}
    fs, fc = scenes.prism_scene(floor=True), scenes.floor_camera(480, 270)
    _, depth = O.surface_rim(fs, fc, 480, 270)
    _, st = O.volume(fs, fc, scenes.prism_beam(), 480, 270, n, 0, vp, depth.ravel(), 0, A.EST_POINT,
                     opts(A.ILLUM_D65), threads=4)
    out["floor_volume"] = {"seg": st.segments / st.paths, "splat": st.splats / st.paths}
    _, st = O.lighttrace(scenes.prism_scene(), scenes.floor_screen(1024, 1024), scenes.prism_beam(), n, 0, 0x0F1,
                         opts(A.ILLUM_D65), threads=4)
    out["floor_caustic"] = {"seg": st.segments / st.paths, "splat": st.splats / st.paths}
    return {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in out.items()}


def beam_roofline(out, a_peak, chunk=1 << 19):
    """The product image's beam-estimator deposit kernel: its VALU lane
    operations and global atomics per dispatch of `chunk` photons (rocprofv3
    counters of that dispatch size) over the volume pass's time per dispatch
    measured here (the pass is 98.5 % deposit kernel)."""
    ms = out.get("prism_on_plane_800x600_volume_ms")
    n_v = out.get("prism_on_plane_800x600_volume_photons", 0)
    prof = "prism_on_plane_beam_deposit_counters.md"
    insts = profile_figure(prof, "valu_insts[k_beam_deposit]")
    lanes = profile_figure(prof, "lanes_active[k_beam_deposit]")
    atomics = profile_figure(prof, "global_atomics_per_launch[k_beam_deposit]")
    if not ms or insts is None or lanes is None:
        return None
    sec = ms / 1e3 / -(-n_v // chunk)
    row = {"kernel": "k_beam_deposit", "bound": "valu", "unit": "T lane-ops/s",
           "achieved": round(insts * lanes / sec / 1e12, 3), "peak": round(VALU_LANE_PEAK / 1e12, 2)}
    row["frac"] = round(row["achieved"] / row["peak"], 4)
    row["valu_busy"] = profile_figure(prof, "valu_busy[k_beam_deposit]")
    row["lanes_active"] = lanes
    if atomics is not None and a_peak:
        row["atomics_gops"] = round(atomics / sec / 1e9, 2)
        row["atomic_frac"] = round(row["atomics_gops"] / a_peak, 4)
    return row


def config_roofline(out, work, a_peak):
    """Per-configuration model fractions: the wavefront state model's HBM
    bytes over 8 TB/s for every form, and for the plain photon kernels --
    whose every deposit is three global fp32 atomics, so the oracle's
    deposit count is their atomic count -- the atomics over the measured
    scattered-atomic peak.  The LDS-aggregated forms (wavefront, batched)
    issue far fewer global atomics than deposits, and camera kernels none
    per path, so no atomic fraction is given for them."""
    rows = {}
    for key, rate in out.items():
        if not key.endswith(("_mphotons_s", "_mpaths_s")):
            continue
        cls = ("rainbow" if key.startswith("rainbow") else "prism" if key.startswith("prism_2") else
               "floor_volume" if key.startswith("floor_volume") else
               "floor_caustic" if key.startswith("floor_caustic") else None)
        if cls is None or cls not in work:
            continue
        wm = work[cls]
        units = rate * 1e6  # per second
        row = {}
        if cls == "rainbow":
            bytes_s = units * (wm["seg"] * CAM_SEG_BYTES + CAM_PATH_BYTES)
        else:
            bytes_s = units * (wm["seg"] * PH_SEG_BYTES + wm["splat"] * SPLAT_BYTES)
            if "_plain_" in key and a_peak:
                row["atomic_frac"] = round(units * wm["splat"] * 3 / 1e9 / a_peak, 4)
        row["hbm_model_frac"] = round(bytes_s / 1e9 / HBM_PEAK_GBS, 4)
        rows[key[:key.rfind("_m")]] = row
    return rows


def per_config(ctx, timer, quick):
    """Throughput of each single-job configuration under every engine: plain
    (one thread per path / photon), wavefront (per-bounce queues) and the
    batched engine running the one job (spx_batch_run)."""
    import ctypes as C

    import spectral_amd as S
    from spectral_amd import _abi as A
    from spectral_amd import scenes
    from spectral_amd.core import _check, opts
    out = {}
    st = timer.stream

    def batched(job):
        tbl = (A.Job * 1)(job)
        return lambda: _check(ctx, A.lib().spx_batch_run(ctx.handle, tbl, 1, None, st))

    def volume_job(ds, cam, bm, w, h, n, ill, disp, film, zbuf=None, hero=0):
        j = A.Job()
        j.kind, j.width, j.height, j.n_photons, j.seed = A.JOB_VOLUME, w, h, n, 0
        j.estimator = A.EST_POINT
        j.opts = opts(ill, disp, hero_lanes=hero)
        j.camera, j.beam, j.vol = cam.basis, bm, A.VolParams(**scenes.HAZE)
        j.scene, j.film = ds.handle.value, film.ptr
        j.zbuf = zbuf.ptr if zbuf is not None else None
        return j

    # C0: DSOTM prism, 2^20 photons, 320x180
    sc = scenes.prism_scene().upload(ctx)
    bm, cam = scenes.prism_beam(), scenes.prism_camera()
    film = S.Film(ctx, 320, 180)
    forms = [(nm, (lambda eng=eng: S.render_volume(ctx, sc, cam, bm, 320, 180, 1 << 20, 0.05, 0.05, 0.0, 14.0,
                                                   None, 0, film=film, illuminant=A.ILLUM_D65, engine=eng,
                                                   stream=st)))
             for eng, nm in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront"))]
    forms.append(("batched", batched(volume_job(sc, cam, bm, 320, 180, 1 << 20, A.ILLUM_D65, True, film))))
    # the same with 4 stratified hero wavelengths per photon (hero_lanes = 4)
    forms += [(f"hero4_{nm}", (lambda eng=eng: S.render_volume(ctx, sc, cam, bm, 320, 180, 1 << 20, 0.05, 0.05, 0.0,
                                                              14.0, None, 0, film=film, illuminant=A.ILLUM_D65,
                                                              engine=eng, stream=st, hero_lanes=4)))
              for eng, nm in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront"))]
    forms.append(("hero4_batched", batched(volume_job(sc, cam, bm, 320, 180, 1 << 20, A.ILLUM_D65, True, film,
                                                      hero=4))))
    for nm, f in forms:
        f()
        ms, _ = timer.run(f, 3)
        out[f"prism_2^20_320x180_{nm}_mphotons_s"] = round((1 << 20) / ms / 1e3, 1)
    film.free()
    # C1: rainbow 512^2 x 64 spp
    rs, rc = scenes.rainbow_scene().upload(ctx), scenes.rainbow_camera(512, 512)
    for eng, nm in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront")):
        tr = S.HipTracer(ctx, rs, rc, 512, 512, A.ILLUM_D65, 1, engine=eng, rr_depth=3)
        f = lambda: tr.accumulate(64, st)
        f()
        ms, _ = timer.run(f, 2)
        out[f"rainbow_512x512x64_{nm}_mpaths_s"] = round(512 * 512 * 64 / ms / 1e3, 1)
        tr.film.free()
    film = S.Film(ctx, 512, 512)
    j = A.Job()
    j.kind, j.width, j.height, j.spp, j.seed = A.JOB_CAMERA, 512, 512, 64, 1
    j.opts, j.camera, j.scene, j.film = opts(A.ILLUM_D65, True, 8, 3), rc.basis, rs.handle.value, film.ptr
    f = batched(j)
    f()
    ms, _ = timer.run(f, 2)
    out["rainbow_512x512x64_batched_mpaths_s"] = round(512 * 512 * 64 / ms / 1e3, 1)
    film.free()
    # C2: full-size prism on floor, 1920x1080 with the surface z-buffer, and
    # its floor caustic on a 1024x1024 screen
    n = 1 << (24 if quick else 28)
    fs = scenes.prism_scene(floor=True).upload(ctx)
    ps = scenes.prism_scene().upload(ctx)
    fc = scenes.floor_camera(1920, 1080)
    rim, zbuf = S.surface_rim(ctx, fs, fc, 1920, 1080)
    rim.free()
    film = S.Film(ctx, 1920, 1080)
    scr = scenes.floor_screen(1024, 1024)
    sfilm = S.Film(ctx, 1024, 1024)
    tag = f"2^{n.bit_length() - 1}"
    for label, ill, disp in (("D65", A.ILLUM_D65, True), ("A", A.ILLUM_A, True), ("disp_off", A.ILLUM_D65, False)):
        f = lambda: S.render_volume(ctx, fs, fc, bm, 1920, 1080, n, 0.05, 0.05, 0.0, 14.0, zbuf, 0,
                                    film=film, illuminant=ill, dispersion=disp, engine=A.ENGINE_PLAIN,
                                    stream=st)
        ms, _ = timer.run(f, 1)
        out[f"floor_volume_{tag}_1920x1080_{label}_plain_mphotons_s"] = round(n / ms / 1e3, 1)
        f = batched(volume_job(fs, fc, bm, 1920, 1080, n, ill, disp, film, zbuf))
        ms, _ = timer.run(f, 1)
        out[f"floor_volume_{tag}_1920x1080_{label}_batched_mphotons_s"] = round(n / ms / 1e3, 1)
        for eng, nm in ((A.ENGINE_PLAIN, "plain"), (A.ENGINE_WAVEFRONT, "wavefront")):
            g = lambda: S.lighttrace_trace(ctx, ps, scr, bm, n, 0x0F1, film=sfilm, illuminant=ill,
                                           dispersion=disp, engine=eng, stream=st)
            ms, _ = timer.run(g, 1)
            out[f"floor_caustic_{tag}_{label}_{nm}_mphotons_s"] = round(n / ms / 1e3, 1)
    film.free()
    sfilm.free()
    zbuf.free()
    # the floor-caustic product image of prism_on_plane.rs at its own sizes:
    # 4e7 caustic photons on a 512^2 floor, 3.5e7 beam-estimator photons at
    # 800x600, every pass on the device (wall time of the whole pipeline)
    from spectral_amd import product
    n_c, n_v = (1 << 22, 1 << 22) if quick else (40_000_000, 35_000_000)
    ctx.sync(st)
    t0 = time.perf_counter()
    img = product.prism_on_plane(ctx, 800, 600, n_c, n_v, stream=st)
    wall = time.perf_counter() - t0
    out["prism_on_plane_800x600_ms"] = round(wall * 1e3, 1)
    out["prism_on_plane_800x600_mphotons_s"] = round((n_c + n_v) / wall / 1e6, 1)
    out["prism_on_plane_800x600_exposure"] = round(img.exposure, 4)
    # its volume pass alone (the beam estimator, z-buffered by the surface
    # pass), timed with HIP events: the deposit kernel's share of the product
    sc = scenes.wedge_scene().upload(ctx)
    zb = S.DeviceBuffer(ctx, 800 * 600, None)
    zb.upload(img.depth)
    vfilm = S.Film(ctx, 800, 600)
    hz = scenes.PLANE_HAZE
    f = lambda: S.render_volumetric_scene(ctx, sc, scenes.plane_camera(800, 600), scenes.plane_beam(), 800, 600, n_v,
                                          hz["sigma_s"], hz["sigma_t"], hz["g"], hz["max_dist"], zb, 0xBEE5,
                                          film=vfilm, engine=A.ENGINE_WAVEFRONT, stream=st)
    ms, _ = timer.run(f, 1)
    out["prism_on_plane_800x600_volume_ms"] = round(ms, 1)
    out["prism_on_plane_800x600_volume_photons"] = n_v
    out["prism_on_plane_800x600_volume_mphotons_s"] = round(n_v / ms / 1e3, 1)
    vfilm.free()
    zb.free()
    sc.close()
    return out


def photon_kernel_name(t) -> str:
    """The photon trace kernel a tuning launches, as rocprofv3 names it."""
    mode = t["splat_mode"]
    wide = t.get("photon_threads", 256) == 1024 and mode in (1, 3)
    return f"k_batch_photons<{mode}, {1024 if wide else 256}>"


def kernel_roofline(name, alg_bytes_step, ms_step, launches_step, bound="valu"):
    """The roofline of one kernel group of the step.

    bound "valu" (the persistent trace kernels, which keep path state in
    registers and issue VALU work on most SIMD cycles): achieved = the VALU
    lane operations of one launch (SQ_INSTS_VALU x lanes active per
    instruction, counted by rocprofv3 on this workload) over the launch's
    duration measured here, against the VALU lane peak; the algorithmic HBM
    bytes over the duration are kept as `hbm_model`.  bound "hbm" (the bin
    passes, which stream records): algorithmic deposit bytes over the
    duration against 8 TB/s."""
    per_launch = alg_bytes_step / max(launches_step, 1)
    ms_launch = ms_step / max(launches_step, 1)
    sec = ms_launch / 1e3
    hbm = per_launch / sec / 1e9 if sec > 0 else 0.0
    hbm_model = {"achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(hbm / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": round(per_launch)}
    out = {"kernel": name, "kernel_ms_per_launch": round(ms_launch, 3), "launches_per_step": launches_step,
           "traffic": profiled_traffic(name)}
    insts = profile_figure(SQ_PROFILE, f"valu_insts[{name}]")
    lanes = profile_figure(SQ_PROFILE, f"lanes_active[{name}]")
    busy = profile_figure(SQ_PROFILE, f"valu_busy[{name}]")
    atomics = profile_figure(HBM_PROFILE, f"global_atomics_per_launch[{name}]")
    if bound == "valu" and insts is not None and lanes is not None and sec > 0:
        achieved = insts * lanes / sec / 1e12
        out.update({"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_LANE_PEAK / 1e12, 2),
                    "unit": "T lane-ops/s", "frac": round(achieved / (VALU_LANE_PEAK / 1e12), 4),
                    "valu_insts_per_launch": insts, "hbm_model": hbm_model})
    else:
        out.update({"bound": "hbm", "achieved": hbm_model["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_model["frac"], "alg_bytes_per_launch": hbm_model["alg_bytes_per_launch"]})
    if busy is not None:
        out["valu_busy"] = busy
    if lanes is not None:
        out["lanes_active"] = lanes
    if busy is not None and lanes is not None:
        # the share of the SIMDs' lane slots doing useful VALU work; the
        # counters' issue ratio is taken as at most 1 (SQ_ACTIVE_INST_VALU also
        # counts the extra cycles of multi-cycle instructions, so a saturated
        # kernel can read above 1; valu_busy keeps the raw ratio)
        out["valu_lane_frac"] = round(min(busy, 1.0) * lanes / 64.0, 3)
    if atomics is not None and sec > 0:
        out["global_atomics_per_launch"] = atomics
        out["atomics_gops"] = round(atomics / sec / 1e9, 3)
    return out


def atomic_peak(ctx, timer, film_floats=3 * 1920 * 1080):
    """A_peak: scattered fp32 atomicAdd throughput (G ops/s) on a buffer of
    the largest film's size, addresses uniform over it (spx_atomic_probe)."""
    import spectral_amd as S
    from spectral_amd import _abi as A
    buf = S.DeviceBuffer(ctx, film_floats)
    threads, ops = 1 << 20, 64
    f = lambda: A.lib().spx_atomic_probe(ctx.handle, buf.ptr, film_floats, ops, threads, timer.stream)
    f()
    ms, _ = timer.run(f, 3)
    buf.free()
    return threads * ops / ms / 1e6


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="smaller per-config sizes")
    ap.add_argument("--autotune", action="store_true",
                    help="run spx_autotune on the table first instead of installing TUNED_MIXED512")
    ap.add_argument("--no-tune", action="store_true", help="keep the context's default tuning")
    ap.add_argument("--sequential", action="store_true",
                    help="rounds in sequence throughout (no overlapped bin passes, no concurrent camera "
                         "round): every launch runs alone, as profilers should see it")
    ap.add_argument("--tuning", type=str, default=None,
                    help="JSON object of spx_tuning fields to install instead of autotuning")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 256 MiB device copy 4x (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the light-transport jobs are independent (they would "
              "shard by job, or by photon / sample ranges summed with one film reduce) and no "
              "multi-GPU path is run here", flush=True)
        return 2

    import spectral_amd as S
    from spectral_amd import _abi as A
    from spectral_amd import jobs

    ctx = S.Context(0)
    stream = ctx.stream()
    timer = Timer(ctx, stream)
    if args.copy_calib:
        n = 64 << 20
        a, b = S.DeviceBuffer(ctx, n), S.DeviceBuffer(ctx, n)
        for _ in range(4):
            A.lib().spx_copy_f32(ctx.handle, b.ptr, a.ptr, n, stream)
        ctx.sync(stream)
        print(json.dumps({"copy_calib_bytes_read": 4 * n, "copy_calib_bytes_written": 4 * n}), flush=True)
        return 0
    specs = jobs.mixed_jobs()
    paths = sum(s.units for s in specs if s.kind == A.JOB_CAMERA)
    photons = sum(s.units for s in specs if s.kind == A.JOB_VOLUME)
    total = paths + photons
    log(f"mixed stream: {len(specs)} jobs, {paths / 1e9:.3f} G paths, {photons / 1e9:.3f} G photons")
    br = jobs.BatchRunner(ctx, specs)

    # launch shape: explicit, the context default, autotuned now, or the
    # autotuner's recorded pick
    if args.tuning:
        ctx.set_tuning(**json.loads(args.tuning))
        tune_info = {"source": "given"}
    elif args.no_tune:
        tune_info = {"source": "context default"}
    elif args.autotune:
        t0 = time.perf_counter()
        _, tune_ms = br.autotune(stream, sweeps=1)
        tune_info = {"source": "spx_autotune on the 512-job table", "seconds": round(time.perf_counter() - t0, 1),
                     "tuned_ms": round(tune_ms, 3)}
    else:
        ctx.set_tuning(**TUNED_MIXED512)
        tune_info = {"source": "spx_autotune's pick on the 512-job table (bench.py --autotune)"}
    if args.sequential:
        ctx.set_tuning(overlap_bins=0, camera_blocks=0)
        tune_info["sequential"] = True
    tuning = ctx.tuning().as_dict()
    log(f"tuning ({tune_info['source']}): {tuning}")

    # the timed region: K replays of the captured schedule
    t0 = time.perf_counter()
    plan = br.plan(stream)
    tune_info["plan"] = dict(plan.info(), capture_ms=round((time.perf_counter() - t0) * 1e3, 1))
    for _ in range(args.warmup):
        plan.launch(stream)
    ctx.sync(stream)
    ms_step, wall_ms = timer.run(lambda: plan.launch(stream), args.steps)
    value = total / ms_step / 1e3
    log(f"batched (graph replay): {ms_step:.2f} ms/step ({wall_ms:.2f} ms wall) -> {value:.1f} {UNIT}")
    plan.close()
    # the two job classes of the stream apart: each half of the table as its
    # own captured schedule (camera renders alone, volume renders alone)
    split = {"mpaths_s_in_mix": round(paths / ms_step / 1e3, 1),
             "mphotons_s_in_mix": round(photons / ms_step / 1e3, 1)}
    for label, kind, units in (("camera_only", A.JOB_CAMERA, paths), ("volume_only", A.JOB_VOLUME, photons)):
        sub = [j for j, s in enumerate(specs) if s.kind == kind]
        p = br.plan(stream, subset=sub)
        p.launch(stream)
        ms_sub, _ = timer.run(lambda: p.launch(stream), args.steps)
        p.close()
        split[f"{label}_ms_per_step"] = round(ms_sub, 3)
        split[f"{label}_{'mpaths_s' if kind == A.JOB_CAMERA else 'mphotons_s'}"] = round(units / ms_sub / 1e3, 1)
    log(f"classes apart: {split}")
    # the same schedule launched eagerly (no graph)
    br.run_batched(stream)
    ctx.sync(stream)
    eager_ms, _ = timer.run(lambda: br.run_batched(stream), args.steps)
    log(f"batched (eager): {eager_ms:.2f} ms/step")
    # per-kernel times: the rounds run in sequence (bin passes not overlapped
    # with the next round, camera round not concurrent), so each kernel
    # group's HIP event pair on its stream brackets that group alone
    ctx.set_tuning(**dict(tuning, overlap_bins=0, camera_blocks=0))
    br.run_batched(stream)
    ctx.sync(stream)
    ctx.kernel_timing(True)
    seq_ms, _ = timer.run(lambda: br.run_batched(stream), args.steps)
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    ctx.set_tuning(**tuning)
    log(f"batched (eager, rounds in sequence): {seq_ms:.2f} ms/step; kernel groups {kt}")
    if args.no_plain:
        ms_plain = float("nan")
    else:
        ms_plain, _ = timer.run(lambda: br.run_plain(stream), 1)
    plain_value = total / ms_plain / 1e3
    log(f"plain one-job-per-launch: {ms_plain:.2f} ms -> {plain_value:.1f}")
    br.close()

    cpu, wm = (None, dict(WORK_MODEL_PINNED)) if args.no_cpu else cpu_baseline(paths, photons)
    cam_bytes = paths * (wm["seg_per_path"] * CAM_SEG_BYTES + CAM_PATH_BYTES)
    ph_bytes = photons * (wm["seg_per_photon"] * PH_SEG_BYTES + wm["splat_per_photon"] * SPLAT_BYTES)
    dep_bytes = photons * wm["splat_per_photon"] * SPLAT_BYTES
    steps = args.steps
    groups = {
        "camera": kernel_roofline("k_batch_camera", cam_bytes, kt["camera"][0] / steps, kt["camera"][1] // steps),
        "photon": kernel_roofline(photon_kernel_name(tuning), ph_bytes, kt["photon"][0] / steps,
                                  kt["photon"][1] // steps),
    }
    if kt["bins"][1]:
        groups["bins"] = kernel_roofline("k_bins passes", dep_bytes, kt["bins"][0] / steps, kt["bins"][1] // steps,
                                         bound="hbm")
    # scattered fp32 atomic peak: the ceiling of the global atomics each
    # group issues (counted by rocprofv3, so a fraction cannot exceed 1)
    a_peak = atomic_peak(ctx, timer)
    for g in groups.values():
        if "atomics_gops" in g:
            g["atomic_frac"] = round(g["atomics_gops"] / a_peak, 4)
    dominant = max(groups, key=lambda g: groups[g]["kernel_ms_per_launch"] * groups[g]["launches_per_step"])
    roofline = dict(groups[dominant])
    roofline["kernels"] = {g: {k: v[k] for k in ("kernel", "bound", "kernel_ms_per_launch", "launches_per_step",
                                                 "achieved", "unit", "frac", "traffic", "valu_busy", "lanes_active",
                                                 "valu_lane_frac", "hbm_model", "atomics_gops", "atomic_frac")
                               if k in v}
                           for g, v in groups.items()}
    roofline["atomic_peak_gops"] = round(a_peak, 2)
    roofline["note"] = ("the persistent trace kernels keep path state in registers and issue VALU work on most "
                        "SIMD cycles: achieved = VALU lane operations per launch (SQ_INSTS_VALU x lanes active, "
                        "rocprofv3 counters of this workload) / the launch time measured here, against the VALU "
                        "lane peak; `hbm_model` is the wavefront state model's bytes over the same time and "
                        "`traffic` the counted HBM bytes per launch")
    roofline["work_model"] = {k: round(v, 4) for k, v in wm.items() if k != "configs"}
    if "configs" in wm:
        roofline["work_model_configs"] = wm["configs"]
    roofline["work_model_source"] = "oracle counts on the cpu_baseline samples" if cpu else "pinned constants"
    result = {
        "metric": METRIC, "value": round(value, 2), "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "512 mixed jobs: rainbow camera renders + prism-on-floor volume renders",
                   "jobs": len(specs), "paths": paths, "photons": photons,
                   "engine": "batched, captured hipGraph replay", "generator_seed": 0},
        "plain_value": None if args.no_plain else round(plain_value, 2),
        "batched_over_plain": None if args.no_plain else round(value / plain_value, 3),
        "eager_batched_ms_per_step": round(eager_ms, 3),
        "sequential_batched_ms_per_step": round(seq_ms, 3),
        "tuning": dict(tuning, **tune_info),
        "mixed_classes": split,
        "roofline": roofline,
    }
    if not args.no_configs:
        result["configs"] = per_config(ctx, timer, args.quick)
        if "configs" in wm:
            result["configs_roofline"] = config_roofline(result["configs"], wm["configs"], a_peak)
        beam = beam_roofline(result["configs"], a_peak)
        if beam is not None:
            result.setdefault("configs_roofline", {})["prism_on_plane_800x600_beam_deposit"] = beam
        log(json.dumps(result["configs"]))
    if not args.no_serve:
        from spectral_amd import serve
        result["serve"] = serve.bench_serve(ctx, quick=args.quick)
    if cpu is not None:
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
