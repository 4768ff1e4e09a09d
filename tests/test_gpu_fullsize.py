"""The batched and served configurations at their full distributions:

  C3  the whole 512-job mixed table (half rainbow camera renders with film
      sides log-uniform in 128..1024 and 1..64 spp, half prism-on-floor
      volume renders of 2^14..2^24 photons), run as one persistent launch
      with the benchmark's launch shape: every job's film equals its
      one-launch-per-job film (float summation order only), and a fixed
      subset -- the eight largest jobs of each kind and eight seeded picks
      of the rest -- equals the oracle at the 0.5 % bar.
  C4  one Poisson load point of 64 requests drawn from the same
      distribution, K in 1..32 progressive frames orbiting 2 deg per frame,
      served by continuous batching: every request completes, and a fixed
      sample of the finished films (the six costliest and ten seeded picks)
      equals the oracle at equal samples.

Each test prints the worst errors it saw."""
import numpy as np
import pytest

from spectral_amd import _abi as A
from spectral_amd import jobs, scenes, serve

from _metrics import rel_l1_per_channel

pytestmark = pytest.mark.gpu

FILM_TOL = 5e-3
# batched vs one launch per job: float summation order only; the table's
# largest photon jobs (2^24 photons) put ~10^5 fp32 adds on a hot pixel, in
# a different order (LDS hash and binned records against direct atomics),
# and differ by up to 7e-5
ALONE_TOL = 2e-4
THREADS = 16


def _oracle_job(oracle, spec, first=0, count=1, cam=None):
    """The oracle's film of `count` frames of a job spec starting at frame `first`."""
    sc, c0, bm = jobs.job_geometry(spec)
    cam = cam or c0
    o = jobs.job_opts(spec)
    if spec.kind == A.JOB_CAMERA:
        f, _ = oracle.camera_accumulate(sc, cam, spec.size, spec.size, spec.spp * count, spec.spp * first,
                                        spec.seed, o, threads=THREADS)
    else:
        f, _ = oracle.volume(sc, cam, bm, spec.size, spec.size, spec.n_photons * count, spec.n_photons * first,
                             A.VolParams(**scenes.HAZE), None, spec.seed, A.EST_POINT, o, threads=THREADS)
    return f


def _oracle_subset(specs, n_large=8, n_pick=8, seed=0):
    cams = [j for j, s in enumerate(specs) if s.kind == A.JOB_CAMERA]
    vols = [j for j, s in enumerate(specs) if s.kind != A.JOB_CAMERA]
    large = sorted(cams, key=lambda j: -specs[j].units)[:n_large] + \
        sorted(vols, key=lambda j: -specs[j].units)[:n_large]
    rest = [j for j in range(len(specs)) if j not in large]
    picks = np.random.default_rng(seed).choice(rest, n_pick, replace=False).tolist()
    return large + picks


def test_config3_full_table(ctx, oracle):
    specs = jobs.mixed_jobs()
    assert len(specs) == 512
    old = ctx.tuning()
    ctx.set_tuning(**jobs.TUNED_MIXED512)
    br = jobs.BatchRunner(ctx, specs)
    try:
        br.clear()
        br.run_batched()
        ctx.sync()
        mix = [f.image() for f in br.films]
        br.clear()
        br.run_plain()
        ctx.sync()
        worst = np.zeros(3)
        for j, f in enumerate(br.films):
            alone = f.image()
            assert mix[j][..., 1].sum() > 0, f"job {j} ({specs[j]}) empty"
            err = rel_l1_per_channel(mix[j], alone)
            worst = np.maximum(worst, err)
            assert (err < ALONE_TOL).all(), f"job {j} ({specs[j]}): batched vs alone {err}"
        print(f"C3 512 jobs: batched vs one launch per job, worst rel L1 {worst}")
        worst = np.zeros(3)
        for j in _oracle_subset(specs):
            err = rel_l1_per_channel(mix[j], _oracle_job(oracle, specs[j]))
            worst = np.maximum(worst, err)
            print(f"  job {j} kind {specs[j].kind} side {specs[j].size} units {specs[j].units}: rel L1 {err}")
            assert (err < FILM_TOL).all(), f"job {j} ({specs[j]}): {err}"
        print(f"C3 oracle subset: worst rel L1 {worst}")
    finally:
        br.close()
        ctx.set_tuning(**old.as_dict())


def test_config4_poisson_load_point(ctx, oracle):
    reqs = serve.request_mix(64, 5)
    srv = serve.Server(ctx)
    try:
        res = serve.run_stream(srv, reqs, rate=60.0, seed=3, keep=True)
        assert np.isfinite(res.latencies).all()
        print(f"C4 64 requests at 60 req/s: p50 {np.percentile(res.latencies, 50) * 1e3:.1f} ms, "
              f"p99 {np.percentile(res.latencies, 99) * 1e3:.1f} ms, {res.steps} steps")
        # one frame's oracle cost per request (the orbit resets the film each frame)
        order = sorted(range(len(reqs)), key=lambda k: -reqs[k][0].units)
        rest = order[6:]
        sample = order[:6] + np.random.default_rng(1).choice(rest, 10, replace=False).tolist()
        worst = np.zeros(3)
        for k in sample:
            spec, frames = reqs[k]
            rid = res.ids[k]
            assert srv.info(rid).frames_done == frames
            g = srv.film(rid, spec.size, spec.size)
            cam = scenes.Camera(srv.frame_camera(rid, frames - 1))
            c = _oracle_job(oracle, spec, first=frames - 1, count=1, cam=cam)
            err = rel_l1_per_channel(g, c)
            worst = np.maximum(worst, err)
            print(f"  request {k} kind {spec.kind} side {spec.size} K {frames}: rel L1 {err}")
            assert g[..., 1].sum() > 0 and (err < FILM_TOL).all(), f"request {k} ({spec}, K={frames}): {err}"
        print(f"C4 served films: worst rel L1 {worst}")
    finally:
        srv.close()
