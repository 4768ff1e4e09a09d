"""Serving: Poisson request streams over the continuous-batching server.

A request is K progressive frames of one job drawn from the mixed-job
distribution (spectral_amd.jobs), its camera orbiting 2 deg per frame.  The
server (spx_server_*, spx_serve.hip) admits requests by film + replica
footprint under a memory cap, keeps films and RNG counters resident, and
runs every active request's next frame in one batched launch per step (a
captured graph, replayed while the set of active requests is unchanged).
``run_stream`` replays arrivals in real time and returns each request's
latency (arrival -> last frame done).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _abi as A
from . import jobs, scenes
from .core import Context, DeviceScene, _check

ORBIT_DEG = 2.0
CAP_BYTES = 200 * (1 << 30)


def orbit_center(spec: jobs.JobSpec):
    return (0.0, 0.0, 0.0) if spec.kind == A.JOB_CAMERA else (0.8, -0.8, 0.0)


class Server:
    def __init__(self, ctx: Context, cap_bytes: int = CAP_BYTES):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx, A.lib().spx_server_create(ctx.handle, cap_bytes, C.byref(h)))
        self.handle = h
        self._scenes = {}
        self._rainbow: Optional[DeviceScene] = None

    def close(self) -> None:
        if self.handle:
            A.lib().spx_server_destroy(self.handle)
            self.handle = None
        for s in self._scenes.values():
            s.close()
        self._scenes.clear()
        if self._rainbow:
            self._rainbow.close()
            self._rainbow = None

    def request(self, spec: jobs.JobSpec, frames: int, orbit: float = ORBIT_DEG) -> A.Request:
        sc, cam, bm = jobs.job_geometry(spec)
        if spec.kind == A.JOB_CAMERA:
            self._rainbow = self._rainbow or sc.upload(self.ctx)
            ds = self._rainbow
        else:
            ds = sc.upload(self.ctx)
        r = A.Request()
        j = r.job
        j.kind, j.width, j.height = spec.kind, spec.size, spec.size
        j.spp, j.n_photons, j.seed = spec.spp, spec.n_photons, spec.seed
        j.estimator = A.EST_POINT
        j.opts = jobs.job_opts(spec)
        j.camera = cam.basis
        if bm is not None:
            j.beam = bm
        j.vol = A.VolParams(**scenes.HAZE)
        j.scene = ds.handle.value
        r.frames = frames
        r.orbit_deg = orbit
        r.orbit_center = A.f3(orbit_center(spec))
        r._scene = ds
        return r

    def submit(self, req: A.Request) -> int:
        rid = C.c_uint64()
        _check(self.ctx, A.lib().spx_server_submit(self.handle, C.byref(req), C.byref(rid)))
        if req.job.kind != A.JOB_CAMERA:
            self._scenes[rid.value] = req._scene
        return rid.value

    def step(self, stream=None) -> int:
        n = C.c_uint32()
        _check(self.ctx, A.lib().spx_server_step(self.handle, stream, C.byref(n)))
        return n.value

    def info(self, rid: int) -> A.RequestInfo:
        i = A.RequestInfo()
        _check(self.ctx, A.lib().spx_server_info(self.handle, rid, C.byref(i)))
        return i

    def frame_camera(self, rid: int, frame: int) -> A.CameraBasis:
        c = A.CameraBasis()
        _check(self.ctx, A.lib().spx_server_frame_camera(self.handle, rid, frame, C.byref(c)))
        return c

    def film(self, rid: int, w: int, h: int) -> np.ndarray:
        out = np.empty(3 * w * h, np.float32)
        _check(self.ctx, A.lib().spx_server_film(self.handle, rid,
                                                 out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out.reshape(h, w, 3)

    def release(self, rid: int) -> None:
        _check(self.ctx, A.lib().spx_server_release(self.handle, rid))
        s = self._scenes.pop(rid, None)
        if s is not None:
            s.close()

    def set_graphs(self, enable: bool) -> None:
        """Replay captured plans while the active set is unchanged (default on)."""
        _check(self.ctx, A.lib().spx_server_set_graphs(self.handle, int(bool(enable))))

    def stats(self) -> dict:
        c, r = C.c_uint64(), C.c_uint64()
        _check(self.ctx, A.lib().spx_server_stats(self.handle, C.byref(c), C.byref(r)))
        return {"captures": c.value, "replays": r.value}

    def bytes_in_use(self) -> int:
        return A.lib().spx_server_bytes_in_use(self.handle)

    def clock(self) -> float:
        return A.lib().spx_server_clock(self.handle)


def request_mix(n: int, seed: int, **kw) -> List[tuple]:
    """(spec, frames) pairs: specs from the mixed-job distribution, K uniform in 1..32."""
    rng = np.random.default_rng(seed + 1000)
    specs = jobs.mixed_jobs(n=n, seed=seed, **kw)
    return [(s, int(rng.integers(1, 33))) for s in specs]


@dataclass
class StreamResult:
    latencies: np.ndarray
    wall_s: float
    units: int
    steps: int
    ids: list
    paths: int = 0     # camera paths traced (all frames)
    photons: int = 0   # photons traced (all frames)


def run_stream(srv: Server, reqs: List[tuple], rate: float, seed: int = 0, stream=None,
               keep: bool = False) -> StreamResult:
    """Replays Poisson arrivals at `rate` requests/s in real time."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    arrival = np.cumsum(gaps) - gaps[0]
    built = [srv.request(s, k) for s, k in reqs]
    t0 = srv.clock()
    nxt, ids, steps = 0, [], 0
    live = {}
    lat = np.full(len(reqs), np.nan)
    while nxt < len(reqs) or live:
        now = srv.clock() - t0
        while nxt < len(reqs) and arrival[nxt] <= now:
            rid = srv.submit(built[nxt])
            live[rid] = nxt
            ids.append(rid)
            nxt += 1
        if not live:
            time.sleep(max(0.0, arrival[nxt] - (srv.clock() - t0)))
            continue
        srv.step(stream)
        steps += 1
        for rid in list(live):
            inf = srv.info(rid)
            if inf.state == A.REQ_DONE:
                k = live.pop(rid)
                lat[k] = (inf.t_done - t0) - arrival[k]
                if not keep:
                    srv.release(rid)
    wall = srv.clock() - t0
    paths = sum(s.units * k for s, k in reqs if s.kind == A.JOB_CAMERA)
    photons = sum(s.units * k for s, k in reqs if s.kind != A.JOB_CAMERA)
    return StreamResult(lat, wall, paths + photons, steps, ids, paths, photons)


def bench_serve(ctx: Context, quick: bool = False, n_req: int = 400) -> dict:
    """Saturation throughput, then p50/p99 latency at 0.25x .. 2x of it, over
    n_req requests per load point (p99 is then the 4th-largest latency of 400,
    not the maximum); paths and photons per second are reported apart."""
    kw = dict(max_side=512, max_photons_log2=22) if quick else {}
    srv = Server(ctx)
    try:
        sat_reqs = request_mix(n_req, 7, **kw)
        res = run_stream(srv, sat_reqs, rate=1e6)  # all at once: saturation
        sat_rate = len(sat_reqs) / res.wall_s
        out = {"saturation_req_s": round(sat_rate, 3),
               "saturation_value": round(res.units / res.wall_s / 1e6, 1),
               "saturation_mpaths_s": round(res.paths / res.wall_s / 1e6, 1),
               "saturation_mphotons_s": round(res.photons / res.wall_s / 1e6, 1),
               "requests_per_point": n_req, "points": []}
        for mult in (0.25, 0.5, 1.0, 2.0):
            reqs = request_mix(n_req, 11, **kw)
            r = run_stream(srv, reqs, rate=mult * sat_rate, seed=int(mult * 100))
            out["points"].append({
                "load": mult, "rate_req_s": round(mult * sat_rate, 3),
                "p50_ms": round(float(np.nanpercentile(r.latencies, 50)) * 1e3, 2),
                "p99_ms": round(float(np.nanpercentile(r.latencies, 99)) * 1e3, 2),
                "throughput": round(r.units / r.wall_s / 1e6, 1),
                "mpaths_s": round(r.paths / r.wall_s / 1e6, 1), "mphotons_s": round(r.photons / r.wall_s / 1e6, 1),
                "steps": r.steps})
        out["graph_steps"] = srv.stats()
        return out
    finally:
        srv.close()
