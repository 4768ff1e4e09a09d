"""Serving: continuous batching admits requests under the footprint cap, keeps
films and RNG counters resident across frames, and every completed film
matches the CPU oracle at equal samples -- alone or in the mix."""
import ctypes as C

import numpy as np
import pytest

from spectral_amd import _abi as A
from spectral_amd import jobs, scenes, serve

from _metrics import rel_l1_per_channel

pytestmark = pytest.mark.gpu

FILM_TOL = 5e-3
ALONE_TOL = 1e-4


def _specs():
    return [
        (jobs.JobSpec(A.JOB_CAMERA, 48, spp=3, seed=1), 3, 2.0),
        (jobs.JobSpec(A.JOB_CAMERA, 64, spp=2, seed=2, dispersion=False), 2, 0.0),
        (jobs.JobSpec(A.JOB_VOLUME, 64, n_photons=6000, seed=3, rotation=15.0, offset=0.03), 3, 0.0),
        (jobs.JobSpec(A.JOB_VOLUME, 80, n_photons=5000, seed=4, rotation=40.0, illuminant=A.ILLUM_A), 2, 2.0),
        (jobs.JobSpec(A.JOB_CAMERA, 32, spp=5, seed=5), 1, 2.0),
    ]


def _oracle_film(oracle, srv, rid, spec, frames, orbit):
    sc, cam, bm = jobs.job_geometry(spec)
    o = jobs.job_opts(spec)
    last = frames - 1
    if orbit != 0.0:
        cam.basis = srv.frame_camera(rid, last)
        first, count = last, 1
    else:
        first, count = 0, frames
    if spec.kind == A.JOB_CAMERA:
        c, _ = oracle.camera_accumulate(sc, cam, spec.size, spec.size, spec.spp * count,
                                        spec.spp * first, spec.seed, o, threads=8)
    else:
        vp = A.VolParams(**scenes.HAZE)
        c, _ = oracle.volume(sc, cam, bm, spec.size, spec.size, spec.n_photons * count,
                             spec.n_photons * first, vp, None, spec.seed, A.EST_POINT, o, threads=8)
    return c


def test_served_films_match_oracle(ctx, oracle):
    srv = serve.Server(ctx)
    try:
        ids = []
        for spec, k, orbit in _specs():
            ids.append(srv.submit(srv.request(spec, k, orbit)))
        steps = 0
        while any(srv.info(i).state != A.REQ_DONE for i in ids):
            assert srv.step() > 0
            steps += 1
        assert steps == max(k for _, k, _ in _specs())
        for rid, (spec, k, orbit) in zip(ids, _specs()):
            g = srv.film(rid, spec.size, spec.size)
            c = _oracle_film(oracle, srv, rid, spec, k, orbit)
            err = rel_l1_per_channel(g, c)
            print(spec.kind, spec.size, k, orbit, err)
            assert (err < FILM_TOL).all()
            srv.release(rid)
        assert srv.bytes_in_use() == 0
    finally:
        srv.close()


def test_alone_equals_in_mix(ctx):
    target = (jobs.JobSpec(A.JOB_VOLUME, 72, n_photons=8000, seed=9, rotation=25.0), 3, 2.0)
    films = []
    for mix in (False, True):
        srv = serve.Server(ctx)
        try:
            others = []
            if mix:
                for spec, k, orbit in _specs():
                    others.append(srv.submit(srv.request(spec, k, orbit)))
            rid = srv.submit(srv.request(*target))
            while srv.info(rid).state != A.REQ_DONE:
                srv.step()
            films.append(srv.film(rid, 72, 72))
        finally:
            srv.close()
    assert films[0].sum() > 0
    assert (rel_l1_per_channel(films[1], films[0]) < ALONE_TOL).all()


def test_admission_under_cap_and_failures(ctx):
    spec = jobs.JobSpec(A.JOB_CAMERA, 64, spp=1, seed=1)
    film_bytes = 12 * 64 * 64
    srv = serve.Server(ctx, cap_bytes=2 * film_bytes)  # room for two films at a time
    try:
        ids = [srv.submit(srv.request(spec, 2)) for _ in range(5)]
        srv.step()
        states = [srv.info(i).state for i in ids]
        assert states.count(A.REQ_ACTIVE) == 2 and states.count(A.REQ_PENDING) == 3
        assert srv.bytes_in_use() == 2 * film_bytes
        # too large for the cap: refused, others unaffected
        with pytest.raises(A.SpxError) as e:
            srv.submit(srv.request(jobs.JobSpec(A.JOB_CAMERA, 128, spp=1), 1))
        assert e.value.code == A.SPX_EBUSY
        # invalid request fails alone
        bad = srv.request(spec, 1)
        bad.job.width = 0
        with pytest.raises(A.SpxError):
            srv.submit(bad)
        bad2 = srv.request(spec, 0)
        with pytest.raises(A.SpxError):
            srv.submit(bad2)
        done = set()
        for _ in range(40):
            srv.step()
            for i in ids:
                if i not in done and srv.info(i).state == A.REQ_DONE:
                    done.add(i)
                    assert srv.film(i, 64, 64).sum() > 0
                    srv.release(i)
            if len(done) == len(ids):
                break
        assert len(done) == len(ids) and srv.bytes_in_use() == 0
    finally:
        srv.close()


def test_poisson_stream_latencies(ctx):
    srv = serve.Server(ctx)
    try:
        reqs = serve.request_mix(12, 3, max_side=128, max_photons_log2=15, max_spp=8)
        res = serve.run_stream(srv, reqs, rate=200.0)
        assert np.isfinite(res.latencies).all() and (res.latencies > 0).all()
        assert res.steps >= max(k for _, k in reqs)
    finally:
        srv.close()


def test_graph_replays_equal_eager_steps(ctx):
    # Requests of equal length keep the active set unchanged: the first step
    # captures a plan, the others replay it with advanced RNG bases.
    specs = [(jobs.JobSpec(A.JOB_CAMERA, 40, spp=2, seed=21), 4, 0.0),
             (jobs.JobSpec(A.JOB_VOLUME, 56, n_photons=7000, seed=22, rotation=30.0), 4, 2.0),
             (jobs.JobSpec(A.JOB_CAMERA, 48, spp=3, seed=23), 4, 2.0)]
    films, stats = [], []
    for graphs in (True, False):
        srv = serve.Server(ctx)
        try:
            srv.set_graphs(graphs)
            ids = [srv.submit(srv.request(*s)) for s in specs]
            while any(srv.info(i).state != A.REQ_DONE for i in ids):
                srv.step()
            films.append([srv.film(i, s.size, s.size) for i, (s, _, _) in zip(ids, specs)])
            stats.append(srv.stats())
        finally:
            srv.close()
    assert stats[0] == {"captures": 1, "replays": 3}, stats[0]
    assert stats[1] == {"captures": 0, "replays": 0}, stats[1]
    for g, e in zip(*films):
        assert g.sum() > 0
        assert (rel_l1_per_channel(g, e) < ALONE_TOL).all()
